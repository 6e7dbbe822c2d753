"""GPU parity of the thermal strains and the partial stress of the elastic
model (src/materials/elastic.cxx:55-146, material_utilities.cxx:235-288)
against the FP64 oracle and the ThermalExpansionCube gold, through the C ABI.

Tolerances (FP64 on both sides, sums in another order): force, dilatation
and partial stress within 1e-11 of the oracle's largest value, the tangent
within 1e-10; the gold at its .comp tolerances."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
K, MU = 130e9, 78e9


def _rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def _bar(nx=12, ny=5, nz=4, h=1e-3):
    x, vol, nl = O.quickgrid(nx, ny, nz, [0, 0, 0], [nx * h, ny * h, nz * h], 3.015 * h)
    return x, vol, nl, 3.015 * h


def _ctx(nl, x, vol, hz, alpha_t=0.0, dT=None):
    from peridigm_amd import Context
    c = Context(0)
    c.set_topology(nl, len(vol))
    c.set_constants(x, vol, K, MU, hz, 0, 0.0)
    if alpha_t:
        c.set_thermal_expansion(alpha_t)
    if dT is not None:
        c.set_temperature_change(dT)
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


def test_thermal_force_and_partial_stress_match_oracle():
    x, vol, nl, hz = _bar()
    N = len(vol)
    rng = np.random.default_rng(2)
    u = 1e-6 * rng.standard_normal(x.shape) + 5e-4 * x
    dT = rng.uniform(-50.0, 150.0, N)
    at = 1.1e-5
    bd = _sym_damage(nl, N, 0.05, 4)
    with _ctx(nl, x, vol, hz, at, dT) as ctx:
        m = ctx.initialize()
        th, f = ctx.compute_force(x + u, bd)
        ps = ctx.compute_partial_stress(x + u, bd)
        # without the field the model is the plain elastic one
        ctx.set_temperature_change(None)
        th0, f0 = ctx.compute_force(x + u, bd)
    tho, fo, pso = O.compute_force_thermal(x, x + u, m, vol, bd, nl, K, MU, hz, at, dT,
                                           partial_stress=True)
    assert _rel(th, tho) <= 1e-11
    assert _rel(f, fo) <= 1e-11
    assert _rel(ps, pso) <= 1e-11
    th_e, f_e = O.compute_force(x, x + u, m, vol, bd, nl, K, MU, hz)
    assert _rel(f0, f_e) <= 1e-5          # FP32-compensated bond passes
    assert _rel(fo, f_e) > 1e-3           # the thermal term is not negligible here


def test_thermal_tangent_matches_oracle():
    from peridigm_amd import quasistatic as Q
    x, vol, nl, hz = _bar(8, 4, 4)
    N = len(vol)
    rng = np.random.default_rng(6)
    u = 1e-5 * rng.standard_normal(x.shape)
    w = rng.standard_normal(x.shape)
    dT0 = rng.uniform(0.0, 100.0, N)
    dT1 = rng.uniform(0.0, 1.0, N)
    at, t = 1e-5, 40.0
    with _ctx(nl, x, vol, hz, at) as ctx:
        m = ctx.initialize()
        ctx.set_temperature_change(dT0, dT1)
        ctx.set_temperature_time(t)
        Q.set_fixed(ctx, np.array([], np.int64))
        kw = Q.tangent_apply(ctx, u, w)
        r, _ = Q.residual(ctx, u)
    bd = np.zeros(O.num_bonds(nl, N))
    f, jw = O.force_directional_thermal(x, u, w, m, vol, bd, nl, K, MU, hz, at, dT1 * t + dT0)
    assert _rel(kw, -jw * vol[:, None]) <= 1e-10
    assert _rel(r, f * vol[:, None]) <= 1e-11


def test_explicit_run_refuses_thermal_strains():
    from peridigm_amd import PdgError
    x, vol, nl, hz = _bar(6, 4, 4)
    with _ctx(nl, x, vol, hz, 1e-5, np.full(len(vol), 10.0)) as ctx:
        ctx.initialize()
        ctx.set_density(np.full(len(vol), 7800.0))
        ctx.set_bcs([])
        with pytest.raises(PdgError):
            ctx.verlet_run(1, 1e-8)


def test_thermal_expansion_cube_deck_against_gold():
    """ThermalExpansionCube through the deck runner: ten quasi-static load
    steps of a prescribed temperature, all eleven gold frames (displacement,
    velocity, force, dilatation, weighted volume, temperature and its change)
    at the .comp tolerances."""
    from peridigm_amd.deck import run_deck
    g = np.load(os.path.join(HERE, "golden", "thermal_expansion_cube.npz"))
    res = run_deck(os.path.join(HERE, "golden", "decks", "ThermalExpansionCube.xml"))
    fr = res["frames"]
    assert np.allclose(res["times"], g["time"], rtol=0, atol=1e-12)
    for ax in "XYZ":
        assert np.abs(fr["Displacement" + ax] - g["Displacement" + ax]).max() <= float(g["tol_displacement"]), ax
        assert np.abs(fr["Velocity" + ax] - g["Velocity" + ax]).max() <= float(g["tol_velocity"]), ax
        assert np.abs(fr["Force" + ax] - g["Force" + ax]).max() <= float(g["tol_force"]), ax
    assert np.abs(fr["Dilatation"] - g["Dilatation"]).max() <= float(g["tol_dilatation"])
    assert np.abs(fr["Weighted_Volume"] - g["Weighted_Volume"]).max() <= float(g["tol_weighted_volume"])
    assert np.abs(fr["Temperature"] - g["Temperature"]).max() <= float(g["tol_temperature"])
    assert np.abs(fr["Temperature_Change"] - g["Temperature_Change"]).max() <= float(g["tol_temperature"])
