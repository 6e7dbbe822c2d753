"""The allocation-free time-step entry point against the allocating one.

pdg_critical_time_step_noalloc runs the same kernels as
pdg_critical_time_step with the per-workgroup partial minima in the
library's fixed device array instead of memory allocated for the call. Two
calls in a row, the second on a smaller block, must each equal the
allocating entry point's result bitwise (same kernels, same reduction order)
and the FP64 oracle's within 1e-14. The output is device memory, allocated
and read here through the HIP runtime the library itself loads.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc_dt(ctx, hip, dbuf, rho):
    ctx.check(ctx.lib.pdg_critical_time_step_noalloc(ctx.h, C.c_double(rho), dbuf))
    out = C.c_double(0.0)
    assert hip.hipMemcpy(C.byref(out), dbuf, 8, 2) == 0  # device to host
    return out.value


def test_noalloc_time_step_equals_allocating_entry():
    from peridigm_amd import Context
    hip = _hip()
    dbuf = C.c_void_p()
    x1, v1, n1 = O.quickgrid(30, 8, 6, [0, 0, 0], [30e-3, 8e-3, 6e-3], 3.015e-3)
    rng = np.random.default_rng(9)
    x2 = rng.uniform(0.0, 5e-3, size=(700, 3))
    n2 = O.radius_search(x2, 1.2e-3)
    v2 = rng.uniform(0.5e-9, 1.5e-9, size=700)
    cases = [(x1, v1, n1, 3.015e-3, 7800.0), (x2, v2, n2, 1.2e-3, 2200.0)]
    assert len(v2) < len(v1)  # the second call is the smaller input
    from peridigm_amd import Context as _C  # noqa: F401
    ctxs = []
    try:
        for x, v, nl, delta, rho in cases:
            ctx = Context(0)
            ctxs.append(ctx)
            ctx.set_topology(nl, len(v))
            ctx.set_constants(x, v, 130e9, 78e9, delta, 0, 0.0)
        assert hip.hipMalloc(C.byref(dbuf), 8) == 0
        for ctx, (x, v, nl, delta, rho) in zip(ctxs, cases):
            ref = ctx.critical_time_step(rho)
            got = _noalloc_dt(ctx, hip, dbuf, rho)
            orc = O.critical_time_step(x, v, nl, rho, 130e9, delta)
            print(f"N={len(v)}: noalloc {got!r} allocating {ref!r} oracle {orc!r}")
            assert got == ref
            assert abs(got - orc) <= 1e-14 * orc
    finally:
        if dbuf.value:
            hip.hipFree(dbuf)
        for ctx in ctxs:
            ctx.close()
