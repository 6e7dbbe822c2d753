"""FP64 CPU restatement of Peridigm's quasi-static load stepping (test
infrastructure only; imported by tests).

Follows Peridigm::executeQuasiStatic (src/core/Peridigm.cpp:2470-3043):
prescribed-displacement increments per load step (DirichletIncrementBC,
src/core/Peridigm_BoundaryCondition.cpp:202-241 and :136-143), the residual
norm ||r||_2 + 20 ||r||_inf with kinematic rows zeroed (:4176-4317), the
reaction-scaled tolerance (:2716-2732), the predictor from the previous load
step's velocity (:2764-2769), damped Newton (:2772-2822, :3302-3319), the
candidate-alpha line search (:3364-3618) and adaptive halving (:2880-2934).
The tangent is the oracle's forward-mode derivative of the force, assembled
densely (small problems only) and solved exactly.
"""
from __future__ import annotations

import numpy as np

from . import force_directional_thermal, weighted_volume


class Problem:
    def __init__(self, x, vol, nlist, K, mu, horizon, alpha_t=0.0):
        self.x, self.vol, self.nl = x, vol, nlist
        self.K, self.mu, self.hz = K, mu, horizon
        self.N = len(vol)
        self.m = weighted_volume(x, vol, nlist, horizon)
        self.bd = np.zeros(nlist.size - self.N)
        self.alpha_t = alpha_t
        self.dT = None  # temperature change per point (thermal strains), or None

    def force_v(self, u):
        """f(u) V, reference order, flattened (3N)."""
        u3 = u.reshape(-1, 3)
        f, _ = force_directional_thermal(self.x, u3, np.zeros_like(u3), self.m, self.vol, self.bd,
                                         self.nl, self.K, self.mu, self.hz, self.alpha_t, self.dT)
        return (f * self.vol[:, None]).ravel()

    def tangent(self, u, fixed):
        """K = -V df/du with identity rows on the kinematic DOFs (dense)."""
        n3 = 3 * self.N
        Kd = np.zeros((n3, n3))
        u3 = u.reshape(-1, 3)
        vv = np.repeat(self.vol, 3)
        for q in range(n3):
            w = np.zeros(n3)
            w[q] = 1.0
            _, jw = force_directional_thermal(self.x, u3, w.reshape(-1, 3), self.m, self.vol,
                                              self.bd, self.nl, self.K, self.mu, self.hz,
                                              self.alpha_t, self.dT)
            Kd[:, q] = -vv * jw.ravel()
        Kd[fixed, :] = 0.0
        Kd[fixed, fixed] = 1.0
        return Kd


def residual_norm(P, u, fixed):
    r = P.force_v(u)
    r[fixed] = 0.0
    return r, np.linalg.norm(r) + 20.0 * np.abs(r).max()


def line_search(P, ub, du, lhs, fixed):
    """quasiStaticsLineSearch (src/core/Peridigm.cpp:3364-3618)."""
    r, r0 = residual_norm(P, ub + du, fixed)
    if r0 == 0.0:
        return 0.0
    best, bestres = 1.0, 1.0e50
    eps = 1.0e-4
    rp, _ = residual_norm(P, ub + (du + eps * lhs), fixed)
    SR, SPR = float(lhs @ r), float(lhs @ rp)
    cand = []
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = float(np.float64(-1.0 * eps * SR) / np.float64(SPR - SR))  # inf/nan as in C++
    if -0.1 < ta < 10.0:
        cand.append(ta)
    cand += [0.1, 0.2, 0.3, 0.5, 0.75, 1.0]
    for a in cand:
        _, rn = residual_norm(P, ub + (du + a * lhs), fixed)
        if np.isfinite(rn) and rn < bestres:
            best, bestres = a, rn
    if (r0 - bestres) / r0 > 0.02:
        return best
    cand = [0.002 * (i + 1) for i in range(25)] + [0.01 * (i + 1) + 0.05 for i in range(15)] + \
           [0.05 * (i + 1) + 0.2 for i in range(6)] + [0.75, 1.0, 1.5, 2.0, 5.0, 10.0]
    up = []
    for a in cand:
        _, rn = residual_norm(P, ub + (du + a * lhs), fixed)
        if np.isfinite(rn):
            if rn < bestres:
                best, bestres = a, rn
            elif rn > r0:
                up.append((a, (rn - r0) / r0))
    if (r0 - bestres) / r0 > 0.0001:
        return best
    dev = 1.0e50
    for a, inc in up:
        if abs(inc - 0.1) < dev:
            best, dev = a, abs(inc - 0.1)
    return best


def qs_run(x, vol, nlist, K, mu, horizon, bcs, times, u0=None, maxSolverIterations=10,
           tolerance=1e-6, absoluteTolerance=False, disableHeuristics=False,
           dampedNewtonDiagonalScaleFactor=1.0001, dampedNewtonDiagonalShiftFactor=1e-5,
           adaptiveLoadStepping=False, maxLoadStepReductionsInOneStep=0,
           maxTotalLoadStepReductions=0, reduceAllRemainingLoadSteps=False, thermal=None):
    """Returns (u, steps, aborted): u (N,3) of the last accepted load step and
    one dict per accepted step (t0, t1, newton, multiplier, status, alpha).
    thermal = (alpha_t, dT(t)): thermal strains with the temperature change
    per point at each load step's time (a prescribed temperature,
    DirichletIncrementBC relative to the initial value,
    src/core/Peridigm_BoundaryAndInitialConditionManager.cpp:148-159)."""
    P = Problem(x, vol, nlist, K, mu, horizon, thermal[0] if thermal else 0.0)
    n3 = 3 * P.N
    kin = {}
    for ids, coord, c0, c1 in bcs:
        for i in ids:
            kin[3 * int(i) + int(coord)] = (float(c0), float(c1))
    kdof = np.array(sorted(kin), dtype=int)
    fixed = np.zeros(n3, bool)
    fixed[kdof] = True
    ub = np.zeros(n3) if u0 is None else np.array(u0, float).ravel().copy()
    pred = np.zeros(n3)
    ts = list(map(float, times))
    tfinal = ts[-1]
    tc = tp = ts[0]
    failed = failed_in_step = False
    fail_step = fail_total = 0
    steps = []
    step = 1
    while step < len(ts):
        if not adaptiveLoadStepping or not failed:
            tp = tc
        failed = False
        tc = ts[step]
        dt = tc - tp
        du = np.zeros(n3)
        if thermal is not None and tc != tp:
            P.dT = np.asarray(thermal[1](tc), float)
        if tc != tp:
            for d in kdof:
                c0, c1 = kin[d]
                cur = c1 * tc + c0
                prev = c1 * tp + c0
                if cur - prev == 0.0:
                    prev = ub[d]
                du[d] = 1.0 * (cur - prev) + 0.0 * (cur - prev) * (1.0 / dt)
        mult = 1.0
        if not absoluteTolerance:
            mult = float(np.linalg.norm(P.force_v(ub + du)[fixed]))
        _, rnorm = residual_norm(P, ub + du, fixed)
        tol = tolerance * mult
        it, damped, alpha = 1, False, 0.0
        while rnorm > tol and it <= maxSolverIterations:
            if it == 1 and step > 1 and not disableHeuristics:
                lhs = pred * dt
                lhs[fixed] = 0.0
            else:
                if it > 50 and not damped and not disableHeuristics:
                    damped = True
                Kd = P.tangent(ub + du, fixed)
                if damped:
                    d = np.diag(Kd) * dampedNewtonDiagonalScaleFactor
                    d = d + dampedNewtonDiagonalShiftFactor * np.abs(d).max()
                    np.fill_diagonal(Kd, d)
                r, _ = residual_norm(P, ub + du, fixed)
                lhs = np.linalg.solve(Kd, r)
            lhs[fixed] = 0.0
            alpha = 1.0 if disableHeuristics else line_search(P, ub, du, lhs, fixed)
            du = du + alpha * lhs
            _, rnorm = residual_norm(P, ub + du, fixed)
            it += 1
        status = 0
        if rnorm > tol:
            if rnorm < 100.0 * tol:
                status = 1
            elif adaptiveLoadStepping:
                fail_step += 1
                fail_total += 1
                if fail_total > maxTotalLoadStepReductions or \
                        fail_step > maxLoadStepReductionsInOneStep:
                    return ub.reshape(-1, 3), steps, True
                failed = failed_in_step = True
            else:
                return ub.reshape(-1, 3), steps, True
        if adaptiveLoadStepping and failed:
            ts.insert(step, tp + dt / 2.0)
            step -= 1
        if adaptiveLoadStepping and reduceAllRemainingLoadSteps and not failed and failed_in_step:
            del ts[step + 1:]
            i = 0
            while ts[step + i] < tfinal:
                ts.append(tc + (i + 1) * dt)
                i += 1
        if not adaptiveLoadStepping or not failed:
            failed_in_step = False
            fail_step = 0
            pred = du / dt
            ub = ub + du
            steps.append(dict(t0=tp, t1=tc, newton=it - 1, multiplier=mult, status=status,
                              alpha=alpha, damped=damped))
        step += 1
    return ub.reshape(-1, 3), steps, False
