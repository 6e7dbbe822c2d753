"""GPU parity of short-range contact against the FP64 oracle and the
reference's contact decks, through the C ABI: the contact search
(src/core/Peridigm_ContactManager.cpp:700-750), the short-range force model
(src/contact/Peridigm_ShortRangeForceContactModel.cpp:98-258) and their place
in the explicit run (src/core/Peridigm.cpp:1376-1565).

Tolerances: contact lists identical; the drop-in contact force within
1e-12 of the oracle's (relative to its largest component: FP64 throughout,
sums in another order); explicit runs with bonds as max |u_gpu - u_oracle| /
max |u_oracle| <= 1e-5 (the internal force is FP32-compensated), velocity
and contact force the same, the total force density within 1e-4 (it is the
output most sensitive to the trajectory: stiff contact springs and the
friction direction amplify the internal force's rounding over the run); the
two-point decks (no bonds) at their gold tolerances."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def _rel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def _two_blocks(nx=8, ny=5, nz=4, h=1e-3, gap=1.1, hz_cells=3.015):
    """Two lattice blocks side by side along x, the facing layers `gap`
    spacings apart, bonds only inside each block."""
    L = [nx * h, ny * h, nz * h]
    x1, v1, nl1 = O.quickgrid(nx, ny, nz, [0, 0, 0], L, hz_cells * h)
    n1 = len(v1)
    x2 = x1.copy()
    x2[:, 0] += (nx - 1 + gap) * h
    nl2 = nl1.copy()
    p = 0
    while p < nl2.size:  # shift the second block's neighbour ids
        n = int(nl2[p])
        nl2[p + 1:p + 1 + n] += n1
        p += n + 1
    x = np.vstack([x1, x2])
    vol = np.concatenate([v1, v1])
    nl = np.concatenate([nl1, nl2]).astype(np.int32)
    return x, vol, nl, hz_cells * h, n1


def _ctx(nl, x, vol, hz, K=140e9, mu=48e9, variant=0):
    from peridigm_amd import Context
    c = Context(0, lanes=4, variant=variant)
    c.set_topology(nl, len(vol))
    c.set_constants(x, vol, K, mu, hz, 0, 0.0)
    return c


@pytest.mark.parametrize("hz_cells,R_cells", [(3.015, 1.6), (1.1, 2.3)])
def test_contact_search_matches_oracle(hz_cells, R_cells):
    h = 1e-3
    x, vol, nl, hz, n1 = _two_blocks(hz_cells=hz_cells, h=h)
    rng = np.random.default_rng(11)
    y = x + 0.05 * h * rng.standard_normal(x.shape)
    y[:n1, 0] += 0.2 * h  # the blocks close in
    with _ctx(nl, x, vol, hz) as ctx:
        ctx.set_contact(R_cells * h, 10, 0.9 * h, 1e12, 0.0)
        got = ctx.contact_search(y)
    ref = O.contact_search(y, R_cells * h, nl)
    assert got.size == ref.size and np.array_equal(got, ref)
    # contacts exist, and with the short horizon some lie inside a block
    rp, col = O.list_to_csr(ref, len(vol))
    rows = np.repeat(np.arange(len(vol)), np.diff(rp))
    assert np.any((rows < n1) != (col < n1))
    if hz_cells < 2:
        assert np.any((rows < n1) == (col < n1))


def test_contact_search_hash_collisions_and_outliers():
    """Scattered points in many more grid cells than hash buckets, negative
    and far-away coordinates: the list is still the brute-force one."""
    rng = np.random.default_rng(5)
    n = 60
    x = rng.uniform(-25.0, 25.0, (n, 3))
    x[:8] = x[8:16] + 0.7 * rng.standard_normal((8, 3))  # close pairs
    # the current configuration carries two far-away points (clamped cells)
    y = x.copy()
    y[16] = [1e12, -3e11, 5.0]
    y[17] = [1e12 + 0.5, -3e11, 5.0]
    vol = np.full(n, 1e-3)
    nl = np.zeros(n, np.int32)  # no bonds
    # bond a few pairs so the exclusion is exercised
    pairs = [(0, 8), (1, 9), (2, 10)]
    rows = [[] for _ in range(n)]
    for a, b in pairs:
        rows[a].append(b); rows[b].append(a)
    nl = np.array([v for r in rows for v in [len(r)] + r], np.int32)
    with _ctx(nl, x, vol, 0.05) as ctx:
        ctx.set_contact(3.0, 1, 1.0, 1.0, 0.0)
        got = ctx.contact_search(y)
    ref = O.contact_search(y, 3.0, nl)
    assert np.array_equal(got, ref)
    rp, col = O.list_to_csr(ref, n)
    assert 17 in col[rp[16]:rp[17]]                 # the far pair
    assert 8 not in col[rp[0]:rp[1]]                # bonded: excluded
    assert rp[-1] > 10


@pytest.mark.parametrize("friction", [0.0, 0.3])
@pytest.mark.parametrize("symmetric", [True, False])
def test_contact_compute_force_matches_oracle(friction, symmetric):
    h = 1e-3
    x, vol, nl, hz, n1 = _two_blocks(hz_cells=1.1, h=h)
    rng = np.random.default_rng(3)
    y = x + 0.05 * h * rng.standard_normal(x.shape)
    y[:n1, 0] += 0.2 * h
    v = rng.standard_normal(x.shape)
    clist = O.contact_search(y, 2.3 * h, nl)
    if not symmetric:  # drop a third of the entries: the reference scatters any list
        rp, col = O.list_to_csr(clist, len(vol))
        out = []
        for i in range(len(vol)):
            keep = [j for j in col[rp[i]:rp[i + 1]] if rng.random() > 0.33]
            out += [len(keep)] + keep
        clist = np.array(out, np.int32)
    rc, ks = 1.3 * h, 1e12
    with _ctx(nl, x, vol, hz) as ctx:
        ctx.set_contact(2.3 * h, 10, rc, ks, friction)
        f = ctx.contact_compute_force(y, v, clist)
    fo = O.contact_force(y, v, vol, clist, rc, ks, friction, hz)
    assert np.abs(fo).max() > 0.0
    assert _rel(f, fo) <= 1e-12


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("friction", [0.0, 0.3])
def test_verlet_contact_matches_oracle(variant, friction):
    """Two bonded blocks driven into each other: the explicit run with
    contact (search every 4 steps) against the oracle's."""
    h = 1e-3
    x, vol, nl, hz, n1 = _two_blocks(hz_cells=3.015, h=h, gap=1.15)
    N = len(vol)
    rho = 8960.0
    v0 = np.zeros((N, 3))
    v0[:n1, 0] = 30.0
    v0[n1:, 0] = -30.0
    v0[:n1, 1] = 10.0  # sliding, for the friction term
    contact = dict(search_radius=2.0 * h, search_frequency=4, contact_radius=1.2 * h,
                   spring_constant=1e12, friction=friction)
    nsteps = 30
    with _ctx(nl, x, vol, hz, variant=variant) as ctx:
        dt = 0.5 * ctx.critical_time_step(rho)
        ctx.set_contact(**contact)
        ctx.initialize()
        ctx.set_density(np.full(N, rho))
        ctx.set_bcs([])
        ctx.upload_state(v=v0)
        ctx.verlet_run(0, dt, initial_eval=True)
        ctx.verlet_run(nsteps // 2, dt, initial_eval=False)
        ctx.verlet_run(nsteps - nsteps // 2, dt, t0=(nsteps // 2) * dt, initial_eval=False)
        st = ctx.download_state()
        fc = ctx.contact_force_density()
    r = O.explicit_run_ext(x, vol, rho, nl, 140e9, 48e9, hz, 0.0, dt, nsteps, v0=v0,
                           contact=contact)
    assert np.abs(r["contact_force"]).max() > 0.0
    assert _rel(st["u"], r["u"]) <= 1e-5
    assert _rel(st["v"] - v0, r["v"] - v0) <= 1e-5
    assert _rel(fc, r["contact_force"]) <= 1e-5
    assert _rel(st["force"], r["force"]) <= 1e-4


def test_contact_off_is_unchanged():
    """Removing the model restores the run without contact, and the
    persistent run refuses contact."""
    from peridigm_amd import PdgError
    h = 1e-3
    x, vol, nl, hz, n1 = _two_blocks(h=h, gap=1.1)
    N = len(vol)
    v0 = np.zeros((N, 3)); v0[:n1, 0] = 20.0
    outs = []
    for on in (False, True):
        with _ctx(nl, x, vol, hz) as ctx:
            dt = 0.5 * ctx.critical_time_step(8960.0)
            if on:
                ctx.set_contact(2e-3, 4, 1.2e-3, 1e12, 0.0)
                with pytest.raises(PdgError):
                    ctx.verlet_run_persistent(2, dt)
                ctx.set_contact(0, 1, 0.0, 0.0, 0.0)
            ctx.initialize()
            ctx.set_density(np.full(N, 8960.0))
            ctx.set_bcs([])
            ctx.upload_state(v=v0)
            ctx.verlet_run(10, dt)
            outs.append(ctx.download_state()["u"])
    assert np.array_equal(outs[0], outs[1])


def _gold(name):
    return np.load(os.path.join(HERE, "golden", name))


@pytest.mark.parametrize("deck,fixture", [("Contact_2x1x1.xml", "contact_2x1x1.npz"),
                                          ("Contact_Friction.xml", "contact_friction.npz")])
def test_contact_decks_against_gold(deck, fixture):
    from peridigm_amd.deck import run_deck
    g = _gold(fixture)
    res = run_deck(os.path.join(DECKS, deck))
    assert np.allclose(res["times"], g["time"], rtol=0, atol=1e-12)
    fr = res["frames"]
    for ax in "XYZ":
        assert np.abs(fr["Displacement" + ax] - g["Displacement" + ax]).max() <= float(g["tol_displacement"]), ax
        assert np.abs(fr["Velocity" + ax] - g["Velocity" + ax]).max() <= float(g["tol_velocity"]), ax
        if "tol_contact_force" in g:
            assert np.abs(fr["Contact_Force_Density" + ax] - g["Contact_Force_Density" + ax]).max() \
                <= float(g["tol_contact_force"]), ax
