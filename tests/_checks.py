"""Shared parity checks for the GPU tests (imported by test modules).

Bond-break masks follow the rule the tests enforce everywhere: at the same
state (the same displacements and the same prior mask) the GPU's decisions
equal the FP64 oracle's for every bond whose FP64 stretch lies outside
1e-6 * s_c of the critical stretch s_c
(src/damage/Peridigm_CriticalStretchDamageModel.cpp:147-161). Bonds inside
that band may go either way and are counted.
"""
from __future__ import annotations

import numpy as np

import oracle as O

BAND = 1e-6


def bond_sc(sc, nlist, N):
    """Critical stretch of every bond in list order (a bond i->j uses node i's)."""
    if np.ndim(sc) == 0:
        return float(sc)
    return np.repeat(np.asarray(sc, float), O.degrees(nlist, N))


def same_state_masks(x, u, nlist, bd_prev, bd_gpu, sc):
    """Checks one damage evaluation: the oracle decides every bond at y = x + u
    from the prior mask bd_prev; every bond where the GPU's mask bd_gpu differs
    must lie in the 1e-6 band. Returns (mismatches, oracle near-critical count)."""
    N = x.shape[0]
    y = x + u
    bd_o, _, near_o = O.compute_damage(x, y, bd_prev, nlist, sc, near_rel=BAND)
    mism = bd_o != bd_gpu
    n = int(np.count_nonzero(mism))
    if n:
        s = O.bond_stretch(x, y, nlist)
        scb = bond_sc(sc, nlist, N)
        band = np.abs(s - scb) <= BAND * scb
        outside = int(np.count_nonzero(mism & ~band))
        assert outside == 0, f"{outside} mask mismatches outside the 1e-6 band ({n} in total)"
    return n, near_o


def rel_max(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def rel_rms(a, b):
    den = np.sqrt(np.mean(np.square(b)))
    return float(np.sqrt(np.mean(np.square(a - b))) / (den if den > 0 else 1.0))


def stepwise_run(ctx, x, nlist, sc, nsteps, dt, check_every=1):
    """Runs the context's explicit state nsteps Verlet steps one step per
    call, and checks the damage decisions of every check_every-th evaluation
    (the initial one included) against the oracle at the same state. Returns
    (total mismatches, evaluations checked, GPU near-critical total)."""
    N = x.shape[0]
    B = ctx.B
    bd_prev = np.zeros(B)
    total = 0
    checked = 0
    near_gpu = 0
    for k in range(nsteps + 1):
        if k == 0:
            ctx.verlet_run(0, dt, t0=0.0, initial_eval=True)
        else:
            ctx.verlet_run(1, dt, t0=(k - 1) * dt, initial_eval=False)
        near_gpu += ctx.near_critical_count()
        st = ctx.download_state()
        if k % check_every == 0 or k == nsteps:
            n, _ = same_state_masks(x, st["u"], nlist, bd_prev, st["bond_damage"], sc)
            total += n
            checked += 1
        bd_prev = st["bond_damage"]
    return total, checked, near_gpu, st


def last_step_check(ctx, x, nlist, sc, nsteps, dt):
    """Full-size variant: runs nsteps - 1 steps on the device, takes the mask,
    runs the last step and checks its damage decisions against the oracle at
    the same state. Returns (mismatches, GPU near-critical count of the last
    step, oracle near count, final state)."""
    ctx.verlet_run(nsteps - 1, dt)
    bd_prev = ctx.download_state()["bond_damage"]
    ctx.verlet_run(1, dt, t0=(nsteps - 1) * dt, initial_eval=False)
    near_gpu = ctx.near_critical_count()
    st = ctx.download_state()
    n, near_o = same_state_masks(x, st["u"], nlist, bd_prev, st["bond_damage"], sc)
    return n, near_gpu, near_o, st
