"""GPU parity of the HIP force path against the FP64 oracle, through the C ABI.

Tolerances (written here): integer results (neighbour
lists, bond-break masks, weighted volume for the "One" influence function)
are bit-exact; FP32-compensated forces and dilatations are checked as
max |f_gpu - f_oracle| / max |f_oracle| <= 1e-5 (and the same for theta);
mask mismatches are allowed only for bonds whose FP64 stretch lies within
1e-6 * s_c of s_c.
"""
import json
import os

import numpy as np
import pytest

import _checks as C
import oracle as O

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _ctx(nl, x, vol, K=130e9, mu=78e9, horizon=1.0, sc=0.0, lanes=16, infl=0, variant=0,
         tile_nodes=128):
    """Context with the given kernel shape. The tile size defaults to 128 so
    that the requested lanes per node run on these small meshes too."""
    from peridigm_amd import Context
    c = Context(0, lanes=lanes, variant=variant, tile_nodes=tile_nodes)
    c.set_topology(nl, len(vol))
    c.set_constants(x, vol, K, mu, horizon, infl, sc)
    return c


def _rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def load(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as fh:
        return json.load(fh)


# ------------------------------------------------------- reference unit tests
@pytest.mark.parametrize("case", ["two", "eight", "three"])
def test_elastic_unit_cases(golden_dir, case):
    g = load(golden_dir, "elastic_unit_tests.json")
    p = g["params"]
    c = g[case]
    x = np.array(c["x"], float); y = np.array(c["y"], float)
    vol = np.array(c["vol"], float); nl = np.array(c["nlist"], np.int32)
    with _ctx(nl, x, vol, p["K"], p["mu"], p["horizon"]) as ctx:
        m = ctx.initialize()
        th, f = ctx.compute_force(y)
    mo = O.weighted_volume(x, vol, nl, p["horizon"])
    tho, fo = O.compute_force(x, y, mo, vol, np.zeros(O.num_bonds(nl, len(vol))), nl,
                              p["K"], p["mu"], p["horizon"])
    assert np.array_equal(m, mo)
    assert _rel(th, tho) <= FORCE_TOL
    assert _rel(f, fo) <= FORCE_TOL
    if case == "eight":
        assert np.allclose(m, c["m"], rtol=c["m_tol"])
        assert np.allclose(th, c["theta"], rtol=1e-6)


# ---------------------------------------------------------- neighbour lists
@pytest.mark.parametrize("n,h,delta,sph", [((20, 12, 8), 1e-3, 3.015, True),
                                          ((2, 2, 1), 1.0, 0.51, False),
                                          ((11, 1, 1), 1.0, 3.1, True),
                                          ((9, 7, 5), 0.37, 2.5, True)])
def test_quickgrid_bit_exact(n, h, delta, sph):
    from peridigm_amd.discretization import quickgrid
    L = [k * h for k in n]
    x, vol, nl = quickgrid(n, [0.1, -0.2, 0.0], L, delta * h, spherical=sph)
    xo, vo, nlo = O.quickgrid(*n, [0.1, -0.2, 0.0], L, delta * h, spherical=sph)
    assert np.array_equal(nl, nlo)
    assert np.array_equal(x, xo)
    assert np.array_equal(vol, vo)


def test_radius_search_and_notch_bit_exact():
    from peridigm_amd.discretization import radius_search
    from peridigm_amd.problems import plate, plate_notch_plane
    p = plate(nx=60, ny=60, nz=4, notch_cells=20, nsteps=10)
    nl = radius_search(p["x"], p["horizon"], p["planes"])
    nlo = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    assert np.array_equal(nl, nlo)
    rng = np.random.default_rng(9)
    xr = rng.uniform(0, 1, size=(4000, 3))
    assert np.array_equal(radius_search(xr, 0.07), O.radius_search(xr, 0.07))


# ----------------------------------------------------------- force / damage
def _bar(nx=30, ny=8, nz=6, h=1e-3):
    x, vol, nl = O.quickgrid(nx, ny, nz, [0, 0, 0], [nx * h, ny * h, nz * h], 3.015 * h)
    return x, vol, nl, 3.015 * h


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("mu", [78e9, 45e9])
@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_compute_force_matches_oracle(mu, lanes, variant):
    x, vol, nl, hz = _bar()
    rng = np.random.default_rng(1)
    u = 1e-3 * 1e-3 * rng.standard_normal(x.shape) + 1e-4 * x  # strain ~1e-4 + noise
    u += 0.3e-3  # a large rigid offset: FP64 differences keep eta exact
    y = x + u
    with _ctx(nl, x, vol, 130e9, mu, hz, lanes=lanes, variant=variant) as ctx:
        assert ctx.tiling()[0] > 0
        m = ctx.initialize()
        th, f = ctx.compute_force(y)
    tho, fo = O.compute_force(x, y, m, vol, np.zeros(O.num_bonds(nl, len(vol))), nl, 130e9, mu, hz)
    assert _rel(th, tho) <= FORCE_TOL
    assert _rel(f, fo) <= FORCE_TOL


@pytest.mark.parametrize("n", [0, 1, 2])
def test_empty_and_tiny_blocks(n):
    """Blocks with no owned points, one isolated point, or two points out of
    each other's reach (an MPI rank can own such a block): every drop-in call
    and a short Verlet run succeed and give zero dilatation and force."""
    x = np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])[:n]
    vol = np.full(n, 1e-9)
    nl = np.zeros(n, np.int32)  # every neighbourhood empty
    with _ctx(nl, x, vol, 130e9, 60e9, 3e-3, sc=5e-4, tile_nodes=0) as ctx:
        m = ctx.initialize()
        assert m.shape == (n,) and not np.any(m)
        th, f = ctx.compute_force(x + 1e-6)
        assert th.shape == (n,) and f.shape == (n, 3)
        assert not np.any(th) and not np.any(f)
        ctx.set_density(np.full(n, 7800.0))
        ctx.upload_state(v=np.full((n, 3), 1.0))
        ctx.verlet_run(3, 1e-8)
        st = ctx.download_state()
        assert np.allclose(st["u"], 3e-8, rtol=1e-12, atol=0.0)
        assert not np.any(st["bond_damage"])


def test_critical_time_step_matches_oracle():
    """Stable time step (src/core/Peridigm_CriticalTimeStep.cpp:54-126) on the
    device against the oracle: a lattice bar with isolated nodes appended
    (no bonds: 1e50, excluded from the minimum), a perturbed random cloud, and
    a block without points (1e50). FP64; the device sums each row in its own
    order, so the tolerance is 1e-14 relative."""
    rng = np.random.default_rng(4)
    x, vol, nl, hz = _bar(nx=30, ny=7, nz=6)
    iso = np.array([[1.0, 1.0, 1.0], [2.0, -1.0, 0.5]])
    x = np.vstack([x, iso]); vol = np.concatenate([vol, np.full(2, vol[0])])
    nl = np.concatenate([nl, np.zeros(2, np.int32)])
    xc = rng.uniform(0.0, 1e-2, size=(3000, 3))
    nlc = O.radius_search(xc, 1.1e-3)
    volc = rng.uniform(0.5e-9, 1.5e-9, size=3000)
    for xx, vv, ll, delta in ((x, vol, nl, hz), (xc, volc, nlc, 1.1e-3)):
        with _ctx(ll, xx, vv, 130e9, 78e9, delta) as ctx:
            for rho in (7800.0, 2200.0):
                dt = ctx.critical_time_step(rho)
                ref = O.critical_time_step(xx, vv, ll, rho, 130e9, delta)
                assert abs(dt - ref) <= 1e-14 * ref, (dt, ref)
    with _ctx(np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0), tile_nodes=0) as ctx:
        assert ctx.critical_time_step(7800.0) == 1.0e50


@pytest.mark.parametrize("tile_nodes, lanes, want", [(0, 4, (32, 16)), (32, 4, (32, 16)),
                                                    (64, 4, (64, 8)), (128, 4, (128, 4)),
                                                    (64, 16, (64, 16))])
def test_tile_sizes_match_oracle(tile_nodes, lanes, want):
    """Tiles of 32, 64 and 128 nodes (and the automatic choice, which picks 32
    for a mesh this small) against the oracle; the tiled passes run at least
    512 / tile size lanes per node."""
    x, vol, nl, hz = _bar()
    rng = np.random.default_rng(5)
    u = 1e-6 * rng.standard_normal(x.shape) + 1e-4 * x
    y = x + u
    with _ctx(nl, x, vol, 130e9, 60e9, hz, lanes=lanes, tile_nodes=tile_nodes) as ctx:
        assert ctx.tile_config() == want
        m = ctx.initialize()
        th, f = ctx.compute_force(y)
    tho, fo = O.compute_force(x, y, m, vol, np.zeros(O.num_bonds(nl, len(vol))), nl, 130e9, 60e9, hz)
    assert _rel(th, tho) <= FORCE_TOL
    assert _rel(f, fo) <= FORCE_TOL


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_isolated_nodes_and_degree_over_128(lanes):
    """Edge cases of the neighbour structure: (a) isolated nodes (empty
    neighbourhoods, zero weighted volume) appended to a lattice bar run the
    tiled kernels with theta = 0 and f = 0 there; (b) a dense random cloud
    with degrees above 128 disables tiling and runs the untiled kernels
    (4 and 8 lanes per node map to 16 there). Both against the oracle."""
    rng = np.random.default_rng(11)
    x, vol, nl, hz = _bar(nx=20, ny=6, nz=5)
    iso = np.array([[1.0, 1.0, 1.0], [2.0, -1.0, 0.5], [-3.0, 0.0, 0.0]])
    x2 = np.vstack([x, iso]); vol2 = np.concatenate([vol, np.full(3, vol[0])])
    nl2 = np.concatenate([nl, np.zeros(3, np.int32)])
    y2 = x2 + 1e-7 * rng.standard_normal(x2.shape) + 2e-4 * x2
    bd = np.zeros(O.num_bonds(nl2, len(vol2)))
    with _ctx(nl2, x2, vol2, 130e9, 60e9, hz, lanes=lanes, sc=5e-4) as ctx:
        assert ctx.tiling()[0] > 0
        m = ctx.initialize()
        th, f = ctx.compute_force(y2, bd)
        bd1, dmg, _ = ctx.compute_damage(y2, bd)
    mo = O.weighted_volume(x2, vol2, nl2, hz)
    tho, fo = O.compute_force(x2, y2, mo, vol2, bd, nl2, 130e9, 60e9, hz)
    bdo, dmgo, _ = O.compute_damage(x2, y2, bd, nl2, 5e-4)
    assert np.array_equal(m, mo) and np.all(m[-3:] == 0.0)
    assert _rel(th, tho) <= FORCE_TOL and _rel(f, fo) <= FORCE_TOL
    assert np.all(th[-3:] == 0.0) and np.all(f[-3:] == 0.0)
    assert np.array_equal(bd1, bdo) and np.array_equal(dmg, dmgo)

    n = 1500
    xc = rng.uniform(0.0, 1.0, size=(n, 3))
    R = 0.3
    nlc = O.radius_search(xc, R)
    deg = O.list_to_csr(nlc, n)[0]
    assert np.diff(deg).max() > 128
    volc = np.full(n, 1.0 / n)
    yc = xc + 1e-4 * xc + 1e-6 * rng.standard_normal(xc.shape)
    with _ctx(nlc, xc, volc, 130e9, 50e9, R, lanes=lanes) as ctx:
        assert ctx.tiling()[0] == 0
        m = ctx.initialize()
        th, f = ctx.compute_force(yc)
    tho, fo = O.compute_force(xc, yc, m, volc, np.zeros(O.num_bonds(nlc, n)), nlc, 130e9, 50e9, R)
    assert _rel(th, tho) <= FORCE_TOL and _rel(f, fo) <= FORCE_TOL


def test_compute_force_with_broken_bonds_and_influence():
    x, vol, nl, hz = _bar(16, 8, 6)
    rng = np.random.default_rng(4)
    y = x + 2e-7 * rng.standard_normal(x.shape)
    rp, col = O.list_to_csr(nl, len(vol))
    B = col.size
    bd = np.zeros(B)
    # break a symmetric random subset
    pick = rng.random(B) < 0.1
    for i in range(len(vol)):
        for b in range(rp[i], rp[i + 1]):
            j = col[b]
            if i < j and pick[b]:
                bd[b] = 1.0
                rb = rp[j] + np.searchsorted(col[rp[j]:rp[j + 1]], i)
                bd[rb] = 1.0
    for infl, variant in ((0, 0), (1, 0), (2, 0), (0, 1), (2, 1)):
        with _ctx(nl, x, vol, 130e9, 50e9, hz, infl=infl, variant=variant) as ctx:
            m = ctx.initialize()
            th, f = ctx.compute_force(y, bd)
        mo = O.weighted_volume(x, vol, nl, hz, infl)
        assert np.allclose(m, mo, rtol=1e-14, atol=0)
        tho, fo = O.compute_force(x, y, mo, vol, bd, nl, 130e9, 50e9, hz, infl)
        assert _rel(th, tho) <= FORCE_TOL
        assert _rel(f, fo) <= FORCE_TOL


def test_asymmetric_damage_rejected():
    from peridigm_amd import PdgError
    x, vol, nl, hz = _bar(6, 4, 4)
    bd = np.zeros(O.num_bonds(nl, len(vol)))
    bd[0] = 1.0
    with _ctx(nl, x, vol, horizon=hz) as ctx:
        with pytest.raises(PdgError):
            ctx.compute_force(x, bd)


def test_damage_field_checks_on_a_large_block():
    """The bond-damage field of a block above the host threading threshold
    (72,000 points): a value other than 0/1 and an asymmetric pair are
    rejected, and an all-intact field gives bitwise the force of no field."""
    from peridigm_amd import PdgError
    x, vol, nl, hz = _bar(60, 60, 20)
    nb = O.num_bonds(nl, len(vol))
    y = x * (1.0 + 1e-4)
    with _ctx(nl, x, vol, horizon=hz, lanes=4) as ctx:
        th0, f0 = ctx.compute_force(y)
        th1, f1 = ctx.compute_force(y, np.zeros(nb))
        assert np.array_equal(f0, f1) and np.array_equal(th0, th1)
        bad = np.zeros(nb); bad[nb // 2] = 0.5
        with pytest.raises(PdgError):
            ctx.compute_force(y, bad)
        asym = np.zeros(nb); asym[nb - 1] = 1.0
        with pytest.raises(PdgError):
            ctx.compute_force(y, asym)


@pytest.mark.parametrize("variant", [0, 1])
def test_compute_damage_bit_exact_outside_band(variant):
    x, vol, nl, hz = _bar(24, 10, 6)
    sc = 5e-4
    rng = np.random.default_rng(2)
    y = x.copy()
    y += 6e-4 * x * np.array([1.0, 0.3, 0.0])  # stretches straddle sc
    y += 1e-8 * rng.standard_normal(x.shape)
    bdn = np.zeros(O.num_bonds(nl, len(vol)))
    with _ctx(nl, x, vol, horizon=hz, sc=sc, variant=variant) as ctx:
        bd, dmg, near = ctx.compute_damage(y, bdn)
    bdo, dmgo, nearo = O.compute_damage(x, y, bdn, nl, sc)
    s = O.bond_stretch(x, y, nl)
    mism = np.nonzero(bd != bdo)[0]
    band = np.abs(s - sc) <= 1e-6 * sc
    assert np.all(band[mism]), "mask mismatch outside the near-critical band"
    assert bdo.sum() > 0 and (bdo == 0).sum() > 0
    if mism.size == 0:
        assert np.array_equal(dmg, dmgo)


# ------------------------------------------------------------- explicit run
def test_bond_breaking_initial_velocity_gold(golden_dir):
    from peridigm_amd import Context
    g = load(golden_dir, "bond_breaking_initial_velocity.json")
    dk = g["deck"]
    from peridigm_amd.discretization import quickgrid
    x, vol, nl = quickgrid(dk["n"], dk["origin"], dk["length"], dk["horizon"], spherical=False)
    v0 = np.zeros((4, 3)); v0[:, 1] = dk["initial_velocity_y"]
    with Context(0) as ctx:
        ctx.set_topology(nl, 4)
        ctx.set_constants(x, vol, dk["K"], dk["mu"], dk["horizon"], 0, dk["critical_stretch"])
        ctx.set_density(np.full(4, dk["density"]))
        ctx.upload_state(v=v0)
        for k in range(1, len(g["time"])):
            ctx.verlet_run(dk["output_every"], dk["dt"], t0=g["time"][k - 1], initial_eval=(k == 1))
            st = ctx.download_state()
            for c, nm in enumerate("XYZ"):
                gu = np.array(g["Displacement" + nm][k]); gv = np.array(g["Velocity" + nm][k])
                assert np.abs(st["u"][:, c] - gu).max() <= 1e-9
                assert np.abs(st["v"][:, c] - gv).max() <= 1e-6 * max(1.0, np.abs(gv).max())
            assert np.abs(st["theta"] - np.array(g["Dilatation"][k])).max() <= 1e-6
            assert np.array_equal(st["damage"], np.array(g["Damage"][k]))


def test_explicit_bar_matches_oracle():
    from peridigm_amd import Context
    from peridigm_amd.problems import bar
    p = bar(nx=40, ny=8, nz=8)
    xo, vo, nl = O.quickgrid(*p["n"], p["origin"], p["length"], p["horizon"])
    nsteps = 200
    r = O.explicit_run(xo, vo, p["density"], nl, p["K"], p["mu"], p["horizon"], 0.0, p["dt"],
                       nsteps, v0=p["v0"])
    with Context(0) as ctx:
        ctx.set_topology(nl, len(vo))
        ctx.set_constants(xo, vo, p["K"], p["mu"], p["horizon"], 0, 0.0)
        ctx.set_density(np.full(len(vo), p["density"]))
        ctx.upload_state(v=p["v0"])
        ctx.verlet_run(nsteps, p["dt"])
        st = ctx.download_state()
    assert _rel(st["u"], r["u"]) <= 1e-5
    assert _rel(st["v"], r["v"]) <= 1e-5
    assert _rel(st["force"], r["force"]) <= 1e-4


def test_explicit_notched_plate_damage_matches_oracle():
    """Damaged notched plate, 150 steps. Every damage evaluation is checked at
    the GPU's own state against the oracle (mismatches only inside the 1e-6
    band); the whole trajectory is then compared with the oracle's own run,
    where a bond breaking one step earlier or later moves its neighbourhood
    (a bounded mismatch count is the second check)."""
    from peridigm_amd import Context
    from peridigm_amd.problems import plate
    p = plate(nx=48, ny=48, nz=4, notch_cells=16, pull=17.3, dt=2.5e-8)
    nl = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    nsteps = 150
    r = O.explicit_run(p["x"], p["volume"], p["density"], nl, p["K"], p["mu"], p["horizon"],
                       p["critical_stretch"], p["dt"], nsteps, bcs=p["bcs"])
    with Context(0) as ctx:
        ctx.set_topology(nl, len(p["volume"]))
        ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, p["critical_stretch"])
        ctx.set_density(np.full(len(p["volume"]), p["density"]))
        ctx.set_bcs(p["bcs"])
        ctx.upload_state()
        mism_steps, checked, near_gpu, st = C.stepwise_run(ctx, p["x"], nl, p["critical_stretch"],
                                                           nsteps, p["dt"])
    assert checked == nsteps + 1
    nb = r["bond_damage"].sum()
    assert nb > 0, "the workload must break bonds"
    mism = np.count_nonzero(st["bond_damage"] != r["bond_damage"])
    du = st["u"] - r["u"]
    umax = np.abs(r["u"]).max()
    err_max = np.abs(du).max() / umax
    err_rms = np.sqrt((du ** 2).mean()) / np.sqrt((r["u"] ** 2).mean())
    print(f"broken={int(nb)} same-state mismatches={mism_steps} near(gpu)={near_gpu} "
          f"trajectory mask mismatch={mism} u_err max={err_max:.2e} rms={err_rms:.2e}")
    assert mism <= max(2, 1e-3 * nb)
    assert err_rms <= 1e-5
    assert err_max <= 1e-4


def test_near_critical_count_in_runs():
    """Explicit runs count intact bonds whose FP32 stretch lies in the 1e-6
    band, over every damage evaluation of the run: the count restarts with
    each run and equals the sum of the counts of the same steps run one at a
    time. A critical stretch set on purpose at a bond's stretch gives a
    nonzero count that the oracle confirms (its own FP64 near count)."""
    from peridigm_amd import Context
    x, vol, nl = O.quickgrid(12, 6, 6, [0, 0, 0], [12e-3, 6e-3, 6e-3], 3.015e-3)
    N = len(vol)
    rng = np.random.default_rng(3)
    u = 1e-4 * x + 1e-8 * rng.standard_normal(x.shape)
    s = O.bond_stretch(x, x + u, nl)
    # the critical stretch set to one bond's own FP64 stretch: that bond (both
    # directions) sits in the band
    sc = float(np.sort(s)[len(s) // 2])
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(x, vol, 130e9, 78e9, 3.015e-3, 0, sc)
        ctx.set_density(np.full(N, 7800.0))
        ctx.upload_state(u=u)
        ctx.verlet_run(0, 1e-12)
        n0 = ctx.near_critical_count()
        _, _, near_o = O.compute_damage(x, x + u, np.zeros(ctx.B), nl, sc, near_rel=1e-6)
        assert n0 > 0 and near_o >= 2
        assert abs(n0 - near_o) <= max(2, 0.05 * near_o)
        ctx.upload_state(u=u)
        ctx.verlet_run(3, 1e-12)
        n3 = ctx.near_critical_count()
        ctx.upload_state(u=u)
        parts = []
        ctx.verlet_run(0, 1e-12)
        parts.append(ctx.near_critical_count())
        for k in range(3):
            ctx.verlet_run(1, 1e-12, t0=k * 1e-12, initial_eval=False)
            parts.append(ctx.near_critical_count())
        assert n3 == sum(parts)
        print(f"near-critical: gpu {n0}, oracle {near_o}; 3-step run {n3} = {parts}")


def test_deterministic_and_lane_independent():
    from peridigm_amd import Context
    from peridigm_amd.problems import plate
    p = plate(nx=40, ny=40, nz=4, notch_cells=10, pull=17.3)
    nl = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    outs = []
    for lanes, variant in ((16, 0), (16, 0), (16, 2), (4, 0), (8, 0), (32, 0), (64, 0), (16, 1), (64, 1)):
        with Context(0, lanes=lanes, variant=variant, tile_nodes=128) as ctx:
            ctx.set_topology(nl, len(p["volume"]))
            ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, p["critical_stretch"])
            ctx.set_density(np.full(len(p["volume"]), p["density"]))
            ctx.set_bcs(p["bcs"])
            ctx.upload_state()
            ctx.verlet_run(60, p["dt"])
            outs.append(ctx.download_state())
    # repeat run, and pipelined vs plain bond loops (same arithmetic order)
    for k in ("u", "v", "bond_damage", "theta"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
        assert np.array_equal(outs[0][k], outs[2][k]), k
    for o in outs[3:]:
        assert _rel(o["u"], outs[0]["u"]) <= 1e-5
        assert np.array_equal(o["bond_damage"], outs[0]["bond_damage"])


# ------------------------------------------------- full-size invariants (C2)
def test_full_plate_invariants():
    from peridigm_amd import Context
    from peridigm_amd.problems import plate, neighbours
    from peridigm_amd.discretization import radius_search
    p = plate()
    nl = neighbours(p)
    x = p["x"]
    full = radius_search(x, p["horizon"])
    assert full.size - len(x) == 55350840  # 122-offset stencil, no notch
    with Context(0) as ctx:
        ctx.set_topology(nl, len(p["volume"]))
        assert 0 < 55350840 - ctx.B < 100000  # bonds removed by the notch
        ctx.set_constants(x, p["volume"], p["K"], 50e9, p["horizon"], 0, 0.0)
        assert ctx.tiling()[0] > 0
        eps = 1e-4
        th, f = ctx.compute_force(x * (1 + eps))
        assert np.abs(th - 3 * eps).max() <= 1e-5 * 3 * eps
        rng = np.random.default_rng(0)
        u = 1e-7 * rng.standard_normal(x.shape)
        th, f = ctx.compute_force(x + u)
        fV = (f * p["volume"][:, None]).sum(0)
        assert np.abs(fV).max() <= 1e-5 * np.abs(f).max() * p["volume"][0] * np.sqrt(len(x))



# ------------------------------------------------ batched persistent launch
def test_persistent_launch_bitwise_equals_per_step_launches():
    from peridigm_amd import Context
    from peridigm_amd.problems import plate
    p = plate(nx=44, ny=40, nz=4, notch_cells=12, pull=17.3)
    nl = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    outs = []
    for persistent in (False, True):
        with Context(0) as ctx:
            ctx.set_topology(nl, len(p["volume"]))
            ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, p["critical_stretch"])
            ctx.set_density(np.full(len(p["volume"]), p["density"]))
            ctx.set_bcs(p["bcs"])
            ctx.upload_state()
            if persistent:
                g = ctx.verlet_run_persistent(40, p["dt"])
                assert g >= 1
                ctx.verlet_run_persistent(25, p["dt"], t0=40 * p["dt"], initial_eval=False)
            else:
                ctx.verlet_run(40, p["dt"])
                ctx.verlet_run(25, p["dt"], t0=40 * p["dt"], initial_eval=False)
            outs.append(ctx.download_state())
    for k in ("u", "v", "a", "bond_damage", "theta", "force"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_batched_specimens_match_oracle():
    """Five uneven plates with their own critical stretch in one persistent
    launch: the last step's damage decisions at the same state, and the
    trajectory against the oracle's run with per-node s_c."""
    from peridigm_amd import Context
    from peridigm_amd.problems import batched_plates
    p = batched_plates(nspec=5, seed=7, nodes=(1500, 5000), notch=(5, 15), pull=17.3)
    nl = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    nsteps = 80
    sc = p["critical_stretch"]
    r = O.explicit_run(p["x"], p["volume"], p["density"], nl, p["K"], p["mu"], p["horizon"],
                       sc, p["dt"], nsteps, bcs=p["bcs"])
    N = len(p["volume"])
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, float(sc.max()))
        ctx.set_critical_stretch(sc)
        ctx.set_density(np.full(N, p["density"]))
        ctx.set_bcs(p["bcs"])
        ctx.upload_state()
        ctx.verlet_run_persistent(nsteps - 1, p["dt"])
        bd_prev = ctx.download_state()["bond_damage"]
        ctx.verlet_run_persistent(1, p["dt"], t0=(nsteps - 1) * p["dt"], initial_eval=False)
        st = ctx.download_state()
    last, _ = C.same_state_masks(p["x"], st["u"], nl, bd_prev, st["bond_damage"], sc)
    nb = r["bond_damage"].sum()
    mism = np.count_nonzero(st["bond_damage"] != r["bond_damage"])
    du = st["u"] - r["u"]
    print(f"specimens={len(p['specimens'])} N={N} broken={int(nb)} last-step mismatches={last} "
          f"trajectory mismatch={mism} u_err max={np.abs(du).max() / np.abs(r['u']).max():.2e}")
    assert nb > 0
    assert mism <= max(2, 1e-3 * nb)
    assert np.abs(du).max() <= 1e-4 * np.abs(r["u"]).max()


def test_expanding_cylinder_matches_oracle():
    """Perturbed hollow cylinder (radius search, radial initial velocity,
    critical stretch 1.5e-3) at test scale against the oracle: every damage
    evaluation at the same state, then the trajectory."""
    from peridigm_amd import Context
    from peridigm_amd.discretization import radius_search
    from peridigm_amd.problems import cylinder
    p = cylinder(r_in=8, r_out=12, length=16, seed=99)
    nl = radius_search(p["x"], p["horizon"])
    assert np.array_equal(nl, O.radius_search(p["x"], p["horizon"]))
    nsteps = 40
    r = O.explicit_run(p["x"], p["volume"], p["density"], nl, p["K"], p["mu"], p["horizon"],
                       p["critical_stretch"], p["dt"], nsteps, v0=p["v0"])
    N = len(p["volume"])
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, p["critical_stretch"])
        ctx.set_density(np.full(N, p["density"]))
        ctx.upload_state(v=p["v0"])
        mism_steps, checked, near_gpu, st = C.stepwise_run(ctx, p["x"], nl, p["critical_stretch"],
                                                           nsteps, p["dt"])
    nb = r["bond_damage"].sum()
    mism = np.count_nonzero(st["bond_damage"] != r["bond_damage"])
    du = st["u"] - r["u"]
    print(f"cylinder N={N} broken={int(nb)} same-state mismatches={mism_steps} "
          f"trajectory mismatch={mism} u_err max={np.abs(du).max() / np.abs(r['u']).max():.2e}")
    assert nb > 0
    assert mism <= max(2, 1e-3 * nb)
    assert np.abs(du).max() <= 1e-4 * np.abs(r["u"]).max()
