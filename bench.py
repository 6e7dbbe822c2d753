#!/usr/bin/env python3
"""Explicit-step throughput of the LPS + critical-stretch force path.

Workload (default): the notched plate, 400x400x4 lattice (640k nodes),
h = 0.5 mm, horizon 3.015h, critical stretch 5e-4, edge notch 40h (bonds
crossing it removed), top/bottom 3-node layers pulled at +-2 m/s; synthetic
inputs built on the device. A step is one velocity-Verlet step: the
damage+dilatation pass and the force+Verlet pass over every directed bond.

Prints ONE JSON line: value = directed-bond evaluations per second
(B bonds per step; each bond is evaluated from both ends, as the CSR holds
both directions) with all inputs resident in HBM. `roofline` reports the
dominant kernel (the force+Verlet pass) from HIP events recorded on the
stream it runs on. `quasi_static` is the second half of the metric: the
60x20x20 tensile bar solved to 1e-8 linear residuals with block-Jacobi CG
and with the learned preconditioner. The cpu_baseline leg times the FP64 CPU
path (the reference loops, multi-threaded, every host thread OpenMP is
given) on a bounded sample and checks the GPU against FP64: `accuracy` holds
the force/dilatation error at the timed run's end state and the displacement
and velocity error of the sample's trajectory.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--workload plate|plate_contact|bar|cylinder|batched]
                    [--lanes 4|8|16|32|64] [--persistent] [--no-qs] [--no-cpu-baseline]
                    [--no-accuracy]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Measured HBM bytes per launch of the default configuration's kernels (plate
# 400x400x4, 4 lanes per node, LDS-tiled pipelined loops): rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch, mean of 73 dispatches), each
# scaled by what the counter reads for a plain elementwise kernel that reads
# and writes 256 MiB (FETCH_SIZE reads half: x2.0; WRITE_SIZE exact: x1.0).
# Source: profiles/bond_passes_tiled_plate_400x400x4.md. Reported only for
# that configuration.
MEASURED_TRAFFIC = {
    "pass1": (94792.0 * 2.0 + 25002.7 * 1.0) * 1024,
    "pass2": (101676.7 * 2.0 + 36232.9 * 1.0) * 1024,
}


def bytes_pass1(B, N):
    """Algorithmic HBM bytes of the damage+dilatation pass: the int32 CSR
    index (4B), the bond-break bits read and written (B/8 each) and per node
    the FP32 reference position and displacement (12 + 12) and the
    dilatation written (4)."""
    return 4 * B + B / 8 + B / 8 + 28 * N


def bytes_pass2(B, N):
    """Algorithmic HBM bytes of the force+Verlet pass: the CSR index (4B),
    the bond-break bits (B/8), and per node 192 bytes: reference position and
    displacement (12 + 12), dilatation, weighted volume and volume (4 + 4 + 4),
    FP64 displacement, velocity and acceleration read and written (3 x 48)
    and the FP32 displacement written (12)."""
    return 4 * B + B / 8 + 192 * N


def build(workload):
    from peridigm_amd import problems
    if workload == "plate":
        p = problems.plate()
    elif workload == "plate_contact":
        p = problems.plate_contact()
    elif workload == "bar":
        p = problems.bar()
    elif workload == "cylinder":
        p = problems.cylinder()
    elif workload == "batched":
        p = problems.batched_plates()
    else:
        raise SystemExit(f"unknown workload {workload}")
    nl = problems.neighbours(p)
    return p, nl


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(target_s=12.0):
    """Multi-threaded FP64 CPU baseline (oracle/pd_baseline.cpp: the reference
    loops in reference order, OpenMP over owned points, per-thread force
    buffers) on a 200x200x4 sample of the plate workload (13.7M bonds), on
    every thread OpenMP is given (OMP_NUM_THREADS). The step count is chosen
    for about target_s seconds of stepping; per-step time is the run minus a
    zero-step run (setup and initial evaluation). Returns the baseline record
    and the sample's final state (for the trajectory error)."""
    import oracle as O
    from peridigm_amd import problems
    p = problems.plate(nx=200, ny=200, nz=4, notch_cells=20)
    nl = O.filter_planes(p["x"], O.radius_search(p["x"], p["horizon"]), p["planes"])
    B = nl.size - len(p["x"])
    m = O.weighted_volume(p["x"], p["volume"], nl, p["horizon"])
    threads = O.baseline_threads()

    def run(n):
        t = time.perf_counter()
        r = O.baseline_explicit_run(p["x"], p["volume"], p["density"], nl, m, p["K"], p["mu"],
                                    p["critical_stretch"], p["dt"], n, bcs=p["bcs"])
        return time.perf_counter() - t, r

    t_zero, _ = run(0)
    t_probe, _ = run(5)
    per = max((t_probe - t_zero) / 5, 1e-6)
    steps = int(min(2000, max(20, target_s / per)))
    t_run, r = run(steps)
    per_step = (t_run - t_zero) / steps
    rec = dict(value=B / per_step, unit="bond_evals/s", cores=threads, kind="port",
               cpu=_cpu_model(),
               sample=f"plate 200x200x4 (N={len(p['x'])}, B={B}), {steps} FP64 Verlet steps "
                      f"(damage, dilatation, scatter force with per-thread buffers) on {threads} "
                      f"threads, {per_step*1e3:.2f} ms/step")
    return rec, dict(p=p, nl=nl, steps=steps, state=r)


def trajectory_error(sample):
    """The GPU run of the CPU baseline's sample (same steps) against the
    baseline's FP64 end state: max and RMS displacement and velocity error
    (relative to max / RMS of the FP64 values) and bond-mask mismatches."""
    from peridigm_amd import Context
    p, nl, steps, r = sample["p"], sample["nl"], sample["steps"], sample["state"]
    N = len(p["volume"])
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0,
                          p["critical_stretch"])
        ctx.set_density(np.full(N, p["density"]))
        ctx.set_bcs(p["bcs"])
        ctx.upload_state()
        ctx.verlet_run(steps, p["dt"])
        st = ctx.download_state()

    def err(a, b):
        return (float(np.abs(a - b).max() / np.abs(b).max()),
                float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))))

    eu, ev = err(st["u"], r["u"]), err(st["v"], r["v"])
    return {"sample": f"plate 200x200x4, {steps} steps, GPU vs the FP64 CPU run",
            "u_max": eu[0], "u_rms": eu[1], "v_max": ev[0], "v_rms": ev[1],
            "broken_fp64": int(r["bond_damage"].sum()),
            "mask_mismatches": int(np.count_nonzero(st["bond_damage"] != r["bond_damage"]))}


def force_error(p, nl, st, ctx_force):
    """Force and dilatation of the GPU drop-in call at the bench run's final
    state (its displacements and broken bonds) against the FP64 oracle at the
    same state: max error over max |f| and RMS error over RMS |f|."""
    import oracle as O
    x = p["x"]
    y = x + st["u"]
    m = O.weighted_volume(x, p["volume"], nl, p["horizon"])
    th_o, f_o = O.compute_force(x, y, m, p["volume"], st["bond_damage"], nl, p["K"], p["mu"],
                                p["horizon"])
    th_g, f_g = ctx_force

    def err(a, b):
        return (float(np.abs(a - b).max() / np.abs(b).max()),
                float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))))

    ef, et = err(f_g, f_o), err(th_g, th_o)
    return {"state": "end of the timed run (same displacements and broken bonds)",
            "force_max": ef[0], "force_rms": ef[1], "dilatation_max": et[0],
            "dilatation_rms": et[1]}


def copy_bandwidth(mib=1024, reps=10):
    """Measured device copy bandwidth (GB/s, bytes read + written) of a plain
    device-to-device copy of `mib` MiB, timed with events on the current
    stream: the practical ceiling beside the 8 TB/s spec."""
    import torch
    src = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    del src, dst
    return 2.0 * mib * 1024 * 1024 / sec / 1e9


def qs_figure(nsamples=100, epochs=20, large=False):
    """Quasi-static tensile bar 60x20x20 (24k nodes; 125x40x40, 200k nodes,
    with `large`), 10 load steps to 1e-3 strain, every linear solve to 1e-8
    relative residual: wall time with block-Jacobi CG, with block-Jacobi CG
    plus the coarse-grid correction (two-level), and with the learned
    preconditioner in flexible CG (trained on the device on
    residual/correction pairs from `nsamples` block-Jacobi solves under random
    end loads, two intermediate iterates each)."""
    from peridigm_amd import Context
    from peridigm_amd import learned as Lr
    from peridigm_amd import quasistatic as Q
    from peridigm_amd.problems import qs_bar, qs_bar_200k, neighbours
    p = qs_bar_200k() if large else qs_bar()
    nl = neighbours(p)
    N = len(p["volume"])
    out = {}
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0, 0.0)
        def note(msg):  # progress on stderr (the JSON line stays alone on stdout)
            print("bench.py: QS N=%d: %s" % (N, msg), file=sys.stderr, flush=True)
        u_j, st_j, wall_j = Q.run_tension(ctx, p, precond="block_jacobi", tol=1e-6, lin_tol=1e-8)
        note("block-Jacobi run %.2f s" % wall_j)
        u_t, st_t, wall_t = Q.run_tension(ctx, p, precond="two_level", tol=1e-6, lin_tol=1e-8)
        note("two-level run %.2f s" % wall_t)
        t0 = time.perf_counter()
        R, E, it_gen = Lr.generate_pairs(ctx, p, nsamples=nsamples, seed=11)
        t_gen = time.perf_counter() - t0
        note("training pairs %.2f s" % t_gen)
        Lr.init(ctx, 5)
        t0 = time.perf_counter()
        loss = Lr.train(ctx, R, E, epochs=epochs, lr=2e-3, seed=3)
        t_train = time.perf_counter() - t0
        note("training %.2f s" % t_train)
        u_l, st_l, wall_l = Q.run_tension(ctx, p, precond="learned", tol=1e-6, lin_tol=1e-8)
        note("learned run %.2f s" % wall_l)
        fallbacks = Q.fallback_count(ctx)
    out = {
        "workload": "QS tensile bar %dx%dx%d (N=%d), 10 load steps to 1e-3 strain, "
                    "linear solves to 1e-8 relative residual" % (*p["n"], N),
        "jacobi_pcg_s": wall_j, "learned_fcg_s": wall_l, "speedup": wall_j / wall_l,
        "jacobi_linear_iters": int(sum(s["linear"] for s in st_j)),
        "learned_linear_iters": int(sum(s["linear"] for s in st_l)),
        "learned_fallbacks_to_jacobi": fallbacks,
        "newton_iters": [int(sum(s["newton"] for s in st_j)), int(sum(s["newton"] for s in st_l))],
        "two_level_pcg_s": wall_t, "two_level_speedup_vs_jacobi": wall_j / wall_t,
        "two_level_linear_iters": int(sum(s["linear"] for s in st_t)),
        "two_level_max_u_diff_rel": float(np.abs(u_t - u_j).max() / np.abs(u_j).max()),
        "max_u_diff_rel": float(np.abs(u_l - u_j).max() / np.abs(u_j).max()),
        "training": {"load_samples": nsamples, "pairs": int(len(R)), "epochs": epochs, "final_rel_loss": loss,
                     "data_s": t_gen, "train_s": t_train},
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="plate")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--variant", type=int, default=0, help="0: LDS-staged tiles, 1: untiled, 2: tiles with plain loops")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-qs", action="store_true",
                    help="skip the quasi-static solve figure (block-Jacobi CG vs learned FCG)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the force and trajectory errors against the FP64 CPU path")
    ap.add_argument("--qs-large", action="store_true",
                    help="also the 200k-node quasi-static bar (125x40x40) with both solvers")
    ap.add_argument("--persistent", action="store_true",
                    help="time the steps as one persistent cooperative launch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (3x) to calibrate HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from peridigm_amd import Context

    if args.copy_calibration:
        # an elementwise kernel that reads 256 MiB and writes 256 MiB
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    p, nl = build(args.workload)
    N = len(p["volume"])
    ctx = Context(0, lanes=args.lanes, variant=args.variant)
    ctx.set_topology(nl, N)
    B = ctx.B
    sc = p["critical_stretch"]
    per_node_sc = np.ndim(sc) > 0
    ctx.set_constants(p["x"], p["volume"], p["K"], p["mu"], p["horizon"], 0,
                      float(np.max(sc)) if per_node_sc else sc)
    if per_node_sc:
        ctx.set_critical_stretch(sc)
    if "contact" in p:
        ctx.set_contact(**p["contact"])
    ctx.set_density(np.full(N, p["density"]))
    ctx.set_bcs(p["bcs"])
    ctx.upload_state(v=p["v0"])
    ntiles, halo_total = ctx.tiling()
    tile_nodes, lanes_run = ctx.tile_config()
    lds_groups, lds_before, lds_after = ctx.lds_schedule()
    dt = p["dt"]

    # warmup: initial force evaluation + W steps
    ctx.verlet_run(args.warmup, dt, t0=0.0, initial_eval=True)
    t_start = args.warmup * dt

    # timed region: exactly K steps, device-resident state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grid = None
    if args.persistent:
        grid = ctx.verlet_run_persistent(args.steps, dt, t0=t_start, initial_eval=False)
    else:
        ctx.verlet_run(args.steps, dt, t0=t_start, initial_eval=False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    near_timed = ctx.near_critical_count()
    t_start += args.steps * dt

    # per-kernel durations from HIP events on the library's stream
    ctx.set_per_pass_timing(True)
    ksteps = max(50, args.steps // 4)
    ctx.verlet_run(ksteps, dt, t0=t_start, initial_eval=False)
    _, p1_ms, p2_ms = ctx.last_timing()
    p1_avg = p1_ms / ksteps * 1e-3
    p2_avg = p2_ms / ksteps * 1e-3
    st = ctx.download_state()
    fg = None
    if not args.no_accuracy:
        fg = ctx.compute_force(p["x"] + st["u"], st["bond_damage"])
    ctx.close()

    copy_gbs = copy_bandwidth()
    profiled = args.workload == "plate" and lanes_run == 4 and tile_nodes == 128 and args.variant == 0
    traffic2 = MEASURED_TRAFFIC["pass2"] if profiled else None
    ms_per_step = elapsed / args.steps * 1e3
    value = B * args.steps / elapsed
    a2 = bytes_pass2(B, N) / p2_avg / 1e9
    a1 = bytes_pass1(B, N) / p1_avg / 1e9
    out = {
        "metric": "bonds/s per explicit step",
        "value": value,
        "unit": "bond_evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": {
                "plate": "notched plate 400x400x4 LPS + critical stretch 5e-4, h=0.5mm, "
                         "horizon 3.015h, +-2 m/s pull, fixed dt 2.5e-8 s",
                "plate_contact": "notched plate 400x400x4 as 'plate', with short-range contact "
                                 "on every node: search radius = horizon every 100 steps, "
                                 "contact radius 0.9h, spring constant 15K",
                "bar": "LPS elastic bar 100x10x10, h=1mm, horizon 3.015h, dt 1e-8 s",
                "cylinder": "expanding hollow cylinder 40h-50h x 200h, h=1mm, perturbed, sc 1.5e-3",
                "batched": "64 notched plates 5k-60k nodes (seed 2024), sc U[3e-4,1e-3], "
                           "notch U[10h,60h], one node/bond space",
            }[args.workload],
            "nodes": N, "bonds": B, "lanes_per_node": lanes_run,
            "launch": ("persistent cooperative, %d workgroups" % grid) if args.persistent
                      else "two launches per step",
            "kernels": ("untiled" if (args.variant == 1 or ntiles == 0) else
                        "lds-tiled" + (", plain loops" if args.variant == 2 else ", pipelined loops")),
            "tiles": ntiles, "tile_nodes": tile_nodes, "halo_per_node": halo_total / N,
            "lds_conflict_cost": [lds_before / max(lds_groups, 1), lds_after / max(lds_groups, 1)],
            "bond_arithmetic": "fp32 compensated", "integration": "fp64",
            "broken_bonds_end": int(st["bond_damage"].sum()),
        },
        "roofline": {
            "bound": "hbm", "achieved": a2, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a2 / HBM_PEAK_GBS, "traffic": traffic2,
            "kernel": "pass 2 (force + Verlet)", "kernel_avg_us": p2_avg * 1e6,
            "pass1_achieved": a1, "pass1_avg_us": p1_avg * 1e6,
            "pass1_traffic": MEASURED_TRAFFIC["pass1"] if profiled else None,
            "algorithmic_bytes": bytes_pass2(B, N),
            "step_frac": (bytes_pass1(B, N) + bytes_pass2(B, N)) / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
            "copy_gbs": copy_gbs, "frac_of_copy": a2 / copy_gbs,
        },
    }
    out["config"]["near_critical_bonds_timed"] = near_timed
    if not args.no_qs:
        out["quasi_static"] = qs_figure()
    if args.qs_large:
        out["quasi_static_200k"] = qs_figure(nsamples=40, epochs=10, large=True)
    if not args.no_cpu_baseline:
        # the cpu_baseline leg: the FP64 CPU path timed on the host, and the
        # GPU results checked against it
        rec, sample = cpu_baseline()
        out["cpu_baseline"] = rec
        if not args.no_accuracy:
            out["accuracy"] = {"end_state": force_error(p, nl, st, fg),
                               "trajectory": trajectory_error(sample)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
