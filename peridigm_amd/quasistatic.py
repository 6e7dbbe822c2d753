"""Quasi-static Newton solve on the device (Peridigm::executeQuasiStatic).

Load stepping as in src/core/Peridigm.cpp:2578-2877: at each load step the
prescribed displacements are applied, the Newton loop solves K(u) du = r(u)
with the matrix-free FP64 tangent, and each linear solve runs to the given
relative residual with block-Jacobi CG or the learned preconditioner inside
flexible CG.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import pdg_qs_params, pdg_qs_step
from .core import Context, _f64, _i32, _p

PRECOND = {"none": 0, "block_jacobi": 1, "learned": 2, "two_level": 3}


def set_fixed(ctx: Context, dofs):
    d = _i32(dofs)
    ctx.check(ctx.lib.pdg_qs_set_fixed(ctx.h, _p(d), d.size))


def residual(ctx: Context, u):
    u = _f64(u, (ctx.N, 3)); r = np.zeros((ctx.N, 3)); n = C.c_double()
    ctx.check(ctx.lib.pdg_qs_residual(ctx.h, _p(u), _p(r), C.byref(n)))
    return r, n.value


def tangent_apply(ctx: Context, u, w):
    u = _f64(u, (ctx.N, 3)); w = _f64(w, (ctx.N, 3)); out = np.zeros((ctx.N, 3))
    ctx.check(ctx.lib.pdg_qs_tangent_apply(ctx.h, _p(u), _p(w), _p(out)))
    return out


def block_diagonal_inverse(ctx: Context, u):
    u = _f64(u, (ctx.N, 3)); out = np.zeros((ctx.N, 3, 3))
    ctx.check(ctx.lib.pdg_qs_block_diagonal(ctx.h, _p(u), _p(out)))
    return out


def solve(ctx: Context, u, rhs, x0=None, precond="block_jacobi", rel_tol=1e-8, max_iter=20000):
    u = _f64(u, (ctx.N, 3)); b = _f64(rhs, (ctx.N, 3))
    x = np.zeros((ctx.N, 3)) if x0 is None else _f64(x0, (ctx.N, 3)).copy()
    it = C.c_int(); rel = C.c_double()
    ctx.check(ctx.lib.pdg_qs_solve(ctx.h, _p(u), _p(b), _p(x), PRECOND[precond], float(rel_tol),
                                   int(max_iter), C.byref(it), C.byref(rel)))
    return x, it.value, rel.value


def newton(ctx: Context, u, tol=1e-6, max_newton=20, precond="block_jacobi", lin_tol=1e-8,
           max_lin=20000):
    u = _f64(u, (ctx.N, 3)).copy()
    ni = C.c_int(); li = C.c_int(); rn = C.c_double()
    ctx.check(ctx.lib.pdg_qs_newton(ctx.h, _p(u), float(tol), int(max_newton), PRECOND[precond],
                                    float(lin_tol), int(max_lin), C.byref(ni), C.byref(li),
                                    C.byref(rn)))
    return u, dict(newton=ni.value, linear=li.value, residual=rn.value)


def force_and_dilatation(ctx: Context, u):
    """FP64 force density (N,3, every DOF) and dilatation (N) at u."""
    u = _f64(u, (ctx.N, 3)); f = np.zeros((ctx.N, 3)); th = np.zeros(ctx.N)
    ctx.check(ctx.lib.pdg_qs_force_dilatation(ctx.h, _p(u), _p(f), _p(th)))
    return f, th


def set_absolute_tolerance(ctx: Context, on=True):
    """Newton tolerance absolute (multiplier 1) instead of scaled by the
    reaction norm (the reference's "Absolute Tolerance")."""
    ctx.check(ctx.lib.pdg_qs_set_absolute_tolerance(ctx.h, int(bool(on))))


def tolerance_multiplier(ctx: Context):
    """Reaction norm (or 1) that scaled the last Newton call's tolerance."""
    m = C.c_double()
    ctx.check(ctx.lib.pdg_qs_tolerance_multiplier(ctx.h, C.byref(m)))
    return m.value


def set_graphs(ctx: Context, enable=True):
    """Replay the CG iteration sequences as HIP graphs, or launch every
    kernel directly (the default)."""
    ctx.check(ctx.lib.pdg_qs_set_graphs(ctx.h, int(bool(enable))))


def fallback_count(ctx: Context):
    """Learned-preconditioner solves finished by block-Jacobi CG so far."""
    n = C.c_int()
    ctx.check(ctx.lib.pdg_qs_fallback_count(ctx.h, C.byref(n)))
    return n.value


def end_dofs(left, right):
    """All three components of the end-layer nodes are kinematic."""
    nodes = np.concatenate([left, right])
    return (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()


def run_tension(ctx: Context, p, precond="block_jacobi", tol=1e-6, lin_tol=1e-8,
                keep_entry=False):
    """Tensile load stepping of the qs_bar problem: end layers displaced by
    -/+ strain*Lx/2 * s (y, z held) for s = k/load_steps. Each step's stats
    hold the Newton and linear iteration counts, the final residual norm and
    the tolerance multiplier (the reaction norm at the step's entry state);
    with keep_entry, also that entry state (u_entry)."""
    N = ctx.N
    set_fixed(ctx, end_dofs(p["left"], p["right"]))
    Lx = p["length"][0]
    u = np.zeros((N, 3))
    stats = []
    t0 = time.perf_counter()
    for k in range(1, p["load_steps"] + 1):
        s = k / p["load_steps"]
        u[p["left"], :] = 0.0
        u[p["right"], :] = 0.0
        u[p["left"], 0] = -0.5 * p["strain"] * Lx * s
        u[p["right"], 0] = 0.5 * p["strain"] * Lx * s
        entry = u.copy() if keep_entry else None
        u, st = newton(ctx, u, tol=tol, precond=precond, lin_tol=lin_tol)
        st["multiplier"] = tolerance_multiplier(ctx)
        if keep_entry:
            st["u_entry"] = entry
        stats.append(st)
    wall = time.perf_counter() - t0
    return u, stats, wall


def default_params(**overrides):
    """pdg_qs_params with the reference's defaults, then the given fields
    (maxSolverIterations, tolerance, absoluteTolerance, disableHeuristics,
    adaptiveLoadStepping, ..., precond as a name or code, linearTolerance)."""
    prm = pdg_qs_params()
    from ._lib import load
    load().pdg_qs_params_default(C.byref(prm))
    for k, v in overrides.items():
        if k == "precond" and isinstance(v, str):
            v = PRECOND[v]
        setattr(prm, k, v)
    return prm


def run(ctx: Context, bcs, times, u0=None, record=False, **params):
    """Peridigm::executeQuasiStatic on the device (pdg_qs_run): load steps at
    `times`, prescribed displacements bcs = [(node_ids, coord, c0, c1)] with
    u(t) = c1 t + c0, parameters as default_params. Returns (u, steps) with
    one dict per accepted load step (with `record`, each also holds that
    step's displacement "u"); raises PdgError if the run aborts."""
    N = ctx.N
    u = np.zeros((N, 3)) if u0 is None else _f64(u0, (N, 3)).copy()
    t = np.ascontiguousarray(times, dtype=np.float64)
    prm = default_params(**params)
    cap = 4 * len(t) + 64
    steps = (pdg_qs_step * cap)()
    ns = C.c_int()
    arr = ctx.bc_array(bcs)
    us = np.zeros((cap, N, 3)) if record else None
    rc = ctx.lib.pdg_qs_run_record(ctx.h, arr, len(bcs), _p(t), t.size, C.byref(prm), _p(u),
                                   steps, cap, C.byref(ns), _p(us))
    out = [dict(t0=st.timePrevious, t1=st.timeCurrent, newton=st.newtonIterations,
                linear=st.linearIterations, residual=st.residualNorm,
                multiplier=st.toleranceMultiplier, status=st.status,
                damped=bool(st.dampedNewton), alpha=st.lastAlpha)
           for st in steps[:min(ns.value, cap)]]
    if record:
        for k, d in enumerate(out):
            d["u"] = us[k].copy()
    ctx.check(rc)
    return u, out
