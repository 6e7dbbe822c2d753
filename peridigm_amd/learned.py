"""Learned preconditioner for the quasi-static linear solves.

A 3-layer message-passing network (width 32) over the bond graph, run and
trained on the device by hand-written BF16-MFMA kernels (libpdgpu.so). It
outputs a correction added to the block-Jacobi step and is used inside
flexible CG; block-Jacobi CG stays the fallback and the judge of the
converged residual. Training pairs (r, e = K^-1 r) come from the plain
solver's own solves under random end loads, and the loss is the energy-norm
error of the preconditioned step, (z - e)^T K (z - e) / e^T K e.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import quasistatic as Q
from .core import Context, _f64, _p

DIMS = (8, 32, 32, 3)


def num_params(ctx: Context) -> int:
    return int(ctx.lib.pdg_gnn_num_params())


def init(ctx: Context, seed=1):
    ctx.check(ctx.lib.pdg_gnn_init(ctx.h, C.c_uint(seed)))


def get_params(ctx: Context):
    p = np.zeros(num_params(ctx), np.float32)
    ctx.check(ctx.lib.pdg_gnn_get_params(ctx.h, _p(p)))
    return p


def set_params(ctx: Context, p):
    p = np.ascontiguousarray(p, np.float32)
    ctx.check(ctx.lib.pdg_gnn_set_params(ctx.h, _p(p)))


def forward_backward(ctx: Context, features, gout=None):
    f = np.ascontiguousarray(features, np.float32).reshape(ctx.N, 8)
    out = np.zeros((ctx.N, 3), np.float32)
    dp = None
    g = None
    if gout is not None:
        g = np.ascontiguousarray(gout, np.float32).reshape(ctx.N, 3)
        dp = np.zeros(num_params(ctx), np.float32)
    ctx.check(ctx.lib.pdg_gnn_forward_backward(ctx.h, _p(f), None if g is None else _p(g),
                                               _p(out), None if dp is None else _p(dp)))
    return out, dp


def _iterate(ctx: Context, u, b, k, lin_tol):
    """The k-th block-Jacobi CG iterate x_k for K(u) x = b from x_0 = 0."""
    x = np.zeros((ctx.N, 3))
    it = C.c_int(); rel = C.c_double()
    rc = ctx.lib.pdg_qs_solve(ctx.h, _p(_f64(u, (ctx.N, 3))), _p(_f64(b, (ctx.N, 3))), _p(x),
                              Q.PRECOND["block_jacobi"], float(lin_tol), int(k), C.byref(it),
                              C.byref(rel))
    if rc not in (0, -5):  # -5: stopped at k iterations before the tolerance
        ctx.check(rc)
    return x


def generate_pairs(ctx: Context, p, nsamples=100, seed=0, lin_tol=1e-8, krylov=2):
    """Residual/correction pairs from the plain solver's own solves.

    Each sample displaces the two end layers by a random rigid translation
    (of order strain * Lx / 2); r is the residual there and block-Jacobi CG
    solves K e = r. Besides (r, e), `krylov` intermediate iterates x_k of the
    same solve (k drawn uniformly below its iteration count) give the pairs
    (r - K x_k, e - x_k): the residuals flexible CG meets after its first
    step."""
    rng = np.random.default_rng(seed)
    N = ctx.N
    fixed = Q.end_dofs(p["left"], p["right"])
    Q.set_fixed(ctx, fixed)
    amp = 0.5 * p["strain"] * p["length"][0]
    u0 = np.zeros((N, 3))
    R, E, iters = [], [], []
    for s in range(nsamples):
        u = np.zeros((N, 3))
        u[p["left"]] = rng.normal(0.0, amp, size=3)
        u[p["right"]] = rng.normal(0.0, amp, size=3)
        r, _ = Q.residual(ctx, u)
        x, it, rel = Q.solve(ctx, u0, r, precond="block_jacobi", rel_tol=lin_tol)
        R.append(r); E.append(x); iters.append(it)
        for k in rng.integers(1, max(2, it), size=krylov):
            xk = _iterate(ctx, u0, r, int(k), lin_tol)
            R.append(r - Q.tangent_apply(ctx, u0, xk))
            E.append(x - xk)
    return np.array(R), np.array(E), iters


def train(ctx: Context, R, E, u=None, epochs=20, lr=2e-3, seed=0):
    n = R.shape[0]
    u = np.zeros((ctx.N, 3)) if u is None else _f64(u, (ctx.N, 3))
    R = np.ascontiguousarray(R, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    loss = C.c_float()
    ctx.check(ctx.lib.pdg_gnn_train(ctx.h, _p(u), n, _p(R), _p(E), int(epochs), C.c_float(lr),
                                    C.c_uint(seed), C.byref(loss)))
    return float(loss.value)
