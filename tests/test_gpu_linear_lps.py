"""GPU parity of the linearized LPS ("Linear LPS Partial Volume" without
partial volumes, src/materials/Peridigm_LinearLPSPVMaterial.cpp:229-394)
against the FP64 oracle and the LinearLPSBar gold, through the C ABI.

The model is linear in u, which fixes its tangent: K w = -V f(w) on free
DOFs. Tolerances (FP64 on both sides, sums in another order): force and
dilatation within 1e-11 of the oracle's largest value, tangent and block
diagonal within 1e-10; the gold at its .comp tolerances."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def _bar(nx=12, ny=5, nz=4, h=1e-3, infl=0):
    x, vol, nl = O.quickgrid(nx, ny, nz, [0, 0, 0], [nx * h, ny * h, nz * h], 3.015 * h)
    return x, vol, nl, 3.015 * h


def _ctx(nl, x, vol, hz, K=130e9, mu=78e9, infl=0, sc=0.0):
    from peridigm_amd import Context
    c = Context(0)
    c.set_topology(nl, len(vol))
    c.set_material_model("Linear LPS")
    c.set_constants(x, vol, K, mu, hz, infl, sc)
    return c


def _sym_damage(nl, n, frac, seed):
    rng = np.random.default_rng(seed)
    rp, col = O.list_to_csr(nl, n)
    bd = np.zeros(col.size)
    pick = rng.random(col.size) < frac
    for i in range(n):
        for b in range(rp[i], rp[i + 1]):
            j = col[b]
            if i < j and pick[b]:
                bd[b] = 1.0
                bd[rp[j] + np.searchsorted(col[rp[j]:rp[j + 1]], i)] = 1.0
    return bd


@pytest.mark.parametrize("infl", [0, 1, 2])
def test_linear_lps_force_matches_oracle(infl):
    x, vol, nl, hz = _bar()
    N = len(vol)
    rng = np.random.default_rng(3)
    u = 1e-5 * rng.standard_normal(x.shape) + 2e-3 * x
    bd = _sym_damage(nl, N, 0.05, 9)
    with _ctx(nl, x, vol, hz, infl=infl) as ctx:
        m = ctx.initialize()
        th, f = ctx.compute_force(x + u, bd)
    mo = O.weighted_volume(x, vol, nl, hz, infl)
    tho, fo = O.compute_force_linear(x, x + u, mo, vol, bd, nl, 130e9, 78e9, hz, infl)
    assert _rel(m, mo) <= 1e-14
    assert _rel(th, tho) <= 1e-11
    assert _rel(f, fo) <= 1e-11
    # the linearization is visible at this strain: the elastic LPS differs
    _, fe = O.compute_force(x, x + u, mo, vol, bd, nl, 130e9, 78e9, hz, infl)
    assert _rel(fe, fo) > 1e-6


def test_linear_lps_tangent_and_block_diagonal():
    from peridigm_amd import quasistatic as Q
    x, vol, nl, hz = _bar(6, 4, 4)
    N = len(vol)
    rng = np.random.default_rng(5)
    u = 1e-5 * rng.standard_normal(x.shape)
    w = rng.standard_normal(x.shape)
    fixed = np.array([0, 1, 2, 3 * (N - 1) + 1])
    bd = np.zeros(O.num_bonds(nl, N))
    with _ctx(nl, x, vol, hz) as ctx:
        m = ctx.initialize()
        Q.set_fixed(ctx, fixed)
        kw = Q.tangent_apply(ctx, u, w)
        dinv = Q.block_diagonal_inverse(ctx, u)
    # linear model: K w = -V f(w) on free DOFs, w on kinematic ones
    _, fw = O.compute_force_linear(x, x + w, m, vol, bd, nl, 130e9, 78e9, hz)
    ref = (-fw * vol[:, None]).ravel()
    ref[fixed] = w.ravel()[fixed]
    assert _rel(kw.ravel(), ref) <= 1e-10
    # diagonal blocks from unit probes of the oracle force
    for i in range(N):
        Kb = np.zeros((3, 3))
        for c in range(3):
            e = np.zeros((N, 3)); e[i, c] = 1e-3
            _, fe = O.compute_force_linear(x, x + e, m, vol, bd, nl, 130e9, 78e9, hz)
            Kb[:, c] = -fe[i] * vol[i] / 1e-3
        for c in range(3):
            if 3 * i + c in fixed:
                Kb[c, :] = 0.0; Kb[:, c] = 0.0; Kb[c, c] = 1.0
        assert np.abs(dinv[i] @ Kb - np.eye(3)).max() <= 1e-9, i


def test_linear_lps_quasi_static_solve():
    """A tension bar solved by the load-stepping driver: the oracle's force
    at the solution vanishes on the free DOFs (relative to the reactions),
    in one Newton step per load step (the tangent is exact)."""
    from peridigm_amd import quasistatic as Q
    x, vol, nl, hz = _bar(16, 5, 5)
    N = len(vol)
    h = 1e-3
    i = np.rint(x[:, 0] / h - 0.5).astype(int)
    left, right = np.nonzero(i <= 2)[0], np.nonzero(i >= 13)[0]
    bcs = [(left, 0, 0.0, -1e-5), (right, 0, 0.0, 1e-5)]
    for c in (1, 2):
        bcs += [(left, c, 0.0, 0.0), (right, c, 0.0, 0.0)]
    with _ctx(nl, x, vol, hz) as ctx:
        m = ctx.initialize()
        u, steps = Q.run(ctx, bcs, [0.0, 0.5, 1.0], tolerance=1e-10)
    assert len(steps) == 2 and all(s["newton"] <= 2 for s in steps)
    assert np.allclose(u[left, 0], -1e-5) and np.allclose(u[right, 0], 1e-5)
    _, f = O.compute_force_linear(x, x + u, m, vol, np.zeros(O.num_bonds(nl, N)), nl,
                                  130e9, 78e9, hz)
    F = f * vol[:, None]
    free = np.ones(N, bool); free[left] = False; free[right] = False
    assert np.abs(F[free]).max() <= 1e-8 * np.abs(F[~free]).max()
    assert np.abs(u[free, 0]).max() > 1e-7


def test_explicit_run_refuses_linear_lps():
    from peridigm_amd import PdgError
    x, vol, nl, hz = _bar(6, 4, 4)
    with _ctx(nl, x, vol, hz) as ctx:
        ctx.initialize()
        ctx.set_density(np.full(len(vol), 7800.0))
        ctx.set_bcs([])
        with pytest.raises(PdgError):
            ctx.verlet_run(1, 1e-8)


def test_linear_lps_bar_deck_against_gold():
    """LinearLPSBar through the deck runner: the two bars in contexts of
    their own (linearized and elastic LPS), five quasi-static load steps, all
    six gold frames at the .comp tolerances."""
    from peridigm_amd.deck import run_deck
    g = np.load(os.path.join(HERE, "golden", "linear_lps_bar.npz"))
    res = run_deck(os.path.join(HERE, "golden", "decks", "LinearLPSBar.yaml"))
    fr = res["frames"]
    assert np.allclose(res["times"], g["time"], rtol=0, atol=1e-14)
    for ax in "XYZ":
        assert np.abs(fr["Displacement" + ax] - g["Displacement" + ax]).max() <= float(g["tol_displacement"]), ax
        assert np.abs(fr["Force" + ax] - g["Force" + ax]).max() <= float(g["tol_force"]), ax
    assert np.abs(fr["Dilatation"] - g["Dilatation"]).max() <= float(g["tol_dilatation"])
    assert np.abs(fr["Weighted_Volume"] - g["Weighted_Volume"]).max() <= float(g["tol_weighted_volume"])
    assert np.abs(fr["Volume"] - g["Volume"]).max() <= float(g["tol_volume"])
