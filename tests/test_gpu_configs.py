"""The five benchmark configurations at their stated sizes, on the GPU,
against the FP64 oracle.

- bar 100x10x10, 1000 Verlet steps: the whole trajectory against the
  oracle's own run (u and v within 1e-6 / 1e-5 of their maxima, force within
  1e-5);
- notched plate 400x400x4 (2000 steps), expanding cylinder (567k nodes, 500
  steps) and the 64-plate sweep (1000 steps, per-specimen critical stretch):
  the last step's damage decisions at the GPU's own state (mismatches only
  inside 1e-6 s_c of s_c), then force and dilatation of the drop-in call at
  the final damaged state against the oracle's (max error over max |f| and
  RMS error over RMS |f| below 1e-5);
- quasi-static bar 60x20x20, 10 load steps to 1e-3 strain: block-Jacobi CG
  and learned flexible CG, each Newton solution re-checked with the oracle's
  residual on the free DOFs.

The explicit runs print their broken-bond and near-critical counts. Seeds are
the workload builders' defaults (cylinder 1234, sweep 2024, training 11/5/3).
"""
import numpy as np
import pytest

import _checks as C
import oracle as O

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _context(p, nl):
    from peridigm_amd import Context
    N = len(p["volume"])
    sc = p["critical_stretch"]
    per_node = np.ndim(sc) > 0
    ctx = Context(0)
    ctx.set_topology(nl, N)
    ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0,
                      float(np.max(sc)) if per_node else sc)
    if per_node:
        ctx.set_critical_stretch(sc)
    ctx.set_density(np.full(N, p["density"]))
    ctx.set_bcs(p["bcs"])
    ctx.upload_state(v=p["v0"])
    return ctx


def test_bar_100x10x10_full_trajectory():
    from peridigm_amd.problems import bar, neighbours
    p = bar()
    nl = neighbours(p)
    assert len(p["volume"]) == 10000 and nl.size - 10000 == 943608
    with _context(p, nl) as ctx:
        ctx.verlet_run(p["nsteps"], p["dt"])
        st = ctx.download_state()
    r = O.explicit_run(p["x"], p["volume"], p["density"], nl, p["K"], p["mu"], p["horizon"], 0.0,
                       p["dt"], p["nsteps"], v0=p["v0"])
    eu, ev, ef = C.rel_max(st["u"], r["u"]), C.rel_max(st["v"], r["v"]), C.rel_max(st["force"], r["force"])
    print(f"bar 1000 steps: u max {eu:.2e} rms {C.rel_rms(st['u'], r['u']):.2e}, v max {ev:.2e}, "
          f"force max {ef:.2e} rms {C.rel_rms(st['force'], r['force']):.2e}")
    assert eu <= 1e-6
    assert ev <= 1e-5
    assert ef <= FORCE_TOL


def _end_state(name, p, nl):
    """Full run with the last step checked at the same state, then the
    drop-in force at the final damaged state against the oracle."""
    x = p["x"]
    sc = p["critical_stretch"]
    with _context(p, nl) as ctx:
        mism, near_gpu, near_o, st = C.last_step_check(ctx, x, nl, sc, p["nsteps"], p["dt"])
        y = x + st["u"]
        bd = st["bond_damage"]
        th_g, f_g = ctx.compute_force(y, bd)
    broken = int(bd.sum())
    m = O.weighted_volume(x, p["volume"], nl, p["horizon"])
    th_o, f_o = O.compute_force(x, y, m, p["volume"], bd, nl, p["K"], p["mu"], p["horizon"])
    ef, efr = C.rel_max(f_g, f_o), C.rel_rms(f_g, f_o)
    et, etr = C.rel_max(th_g, th_o), C.rel_rms(th_g, th_o)
    print(f"{name}: N={len(x)} steps={p['nsteps']} broken={broken} last-step mismatches "
          f"(inside band)={mism} near-critical last step gpu={near_gpu} oracle={near_o}; "
          f"force max {ef:.2e} rms {efr:.2e}; dilatation max {et:.2e} rms {etr:.2e}")
    assert broken > 0
    assert ef <= FORCE_TOL and efr <= FORCE_TOL
    assert et <= FORCE_TOL and etr <= FORCE_TOL


def test_notched_plate_400x400x4_2000_steps():
    from peridigm_amd.problems import plate, neighbours
    p = plate()
    nl = neighbours(p)
    assert len(p["volume"]) == 640000
    _end_state("plate 400x400x4", p, nl)


def test_expanding_cylinder_full_500_steps():
    from peridigm_amd.problems import cylinder, neighbours
    p = cylinder()
    nl = neighbours(p)
    assert len(p["volume"]) > 560000
    _end_state("cylinder 40h-50h x 200h", p, nl)


def test_batched_64_plates_1000_steps():
    from peridigm_amd.problems import batched_plates, neighbours
    p = batched_plates()
    nl = neighbours(p)
    assert len(p["specimens"]) == 64
    sizes = [s["nodes"] for s in p["specimens"]]
    assert min(sizes) >= 4000 and max(sizes) <= 61000
    _end_state("64-plate sweep", p, nl)


@pytest.mark.parametrize("precond", ["block_jacobi", "learned"])
def test_qs_bar_60x20x20_ten_load_steps(precond):
    from peridigm_amd import Context
    from peridigm_amd import learned as Lr
    from peridigm_amd import quasistatic as Q
    from peridigm_amd.problems import qs_bar, neighbours
    p = qs_bar()
    nl = neighbours(p)
    x, vol = p["x"], p["volume"]
    N = len(vol)
    assert N == 24000 and nl.size - N == 2550648
    tol = 1e-6
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(x, vol, p["K"], p["mu"], p["horizon"], 0, 0.0)
        if precond == "learned":
            R, E, _ = Lr.generate_pairs(ctx, p, nsamples=100, seed=11)
            Lr.init(ctx, 5)
            Lr.train(ctx, R, E, epochs=20, lr=2e-3, seed=3)
        u, stats, wall = Q.run_tension(ctx, p, precond=precond, tol=tol, lin_tol=1e-8,
                                       keep_entry=True)
        m = ctx.initialize()
    print(f"QS bar {precond}: {wall:.3f} s, Newton {sum(s['newton'] for s in stats)}, "
          f"linear {sum(s['linear'] for s in stats)}")
    assert len(stats) == 10
    bd0 = np.zeros(nl.size - N)
    free = np.ones(N, bool)
    free[p["left"]] = False
    free[p["right"]] = False
    # the last load step's tolerance is tol times the reaction norm (f V on
    # the kinematic DOFs) at that step's entry state: the previous solution
    # with the new end displacements (src/core/Peridigm.cpp:2716-2732)
    last = stats[-1]
    fe = O.force_at(x, last["u_entry"], m, vol, bd0, nl, p["K"], p["mu"], p["horizon"])
    reac = np.linalg.norm((fe * vol[:, None])[~free])
    mult = last["multiplier"]
    assert abs(mult - reac) <= 1e-8 * reac
    f = O.force_at(x, u, m, vol, bd0, nl, p["K"], p["mu"], p["horizon"])
    rf = (f * vol[:, None])[free]
    assert np.linalg.norm(rf) + 20 * np.abs(rf).max() <= tol * mult * (1 + 1e-6)
    Lx = p["length"][0]
    assert np.allclose(u[p["right"], 0], 0.5 * p["strain"] * Lx, rtol=0, atol=1e-15)
