"""Accuracy of the GPU path against the FP64 oracle at the full sizes of the
explicit workloads. Prints one JSON line per workload.

Each workload first runs its full explicit run on the GPU with device-resident
Verlet. Then, at that final damaged state, the drop-in calls are compared with
the oracle on identical inputs (y = x + u, the GPU's bond damage):
- force: pdg_compute_force vs orc_compute_force; max error over max|f| and
  RMS error over RMS|f|, and the same for the dilatation;
- damage: pdg_compute_damage vs orc_compute_damage from that state; mask
  mismatches outside the 1e-6 band around the critical stretch, and the
  bonds inside it (both sides count them). The batched sweep has one critical
  stretch per specimen; the oracle's drop-in damage call takes one value, so
  there only the force is compared (its per-specimen masks are checked by the
  batched parity test).

The bar (10k nodes) is also run by the oracle for all its 1000 steps, and
the final displacements are compared.

    python tests/accuracy_report.py [bar] [plate] [cylinder] [batched]
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402


def _errs(a, b):
    a = np.asarray(a, float).ravel(); b = np.asarray(b, float).ravel()
    d = a - b
    return dict(max_rel=float(np.abs(d).max() / max(np.abs(b).max(), 1e-300)),
                rms_rel=float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(b * b)), 1e-300)))


def _problem(name):
    from peridigm_amd import problems as P
    return {"bar": P.bar, "plate": P.plate, "cylinder": P.cylinder,
            "batched": P.batched_plates}[name]()


def report(name):
    from peridigm_amd import Context
    from peridigm_amd.problems import neighbours
    p = _problem(name)
    nl = neighbours(p)
    x, vol = p["x"], p["volume"]
    N = len(vol)
    sc = p["critical_stretch"]
    per_node = np.ndim(sc) > 0
    out = dict(workload=name, nodes=N, steps=p["nsteps"])
    t0 = time.perf_counter()
    with Context(0) as ctx:
        ctx.set_topology(nl, N)
        ctx.set_constants(x, vol, p["K"], p["mu"], p["horizon"], 0,
                          float(np.max(sc)) if per_node else sc)
        if per_node:
            ctx.set_critical_stretch(sc)
        ctx.set_density(np.full(N, p["density"]))
        ctx.set_bcs(p["bcs"])
        ctx.upload_state(v=p["v0"])
        ctx.verlet_run(p["nsteps"], p["dt"])
        st = ctx.download_state()
        out["bonds"] = ctx.B
        y = x + st["u"]
        bd = st["bond_damage"]
        th_g, f_g = ctx.compute_force(y, bd)
        dmg_gpu = None
        if not per_node and sc > 0.0:
            dmg_gpu = ctx.compute_damage(y, bd)
    out["gpu_s"] = time.perf_counter() - t0
    out["broken_bonds"] = int(bd.sum())
    m = O.weighted_volume(x, vol, nl, p["horizon"])
    th_o, f_o = O.compute_force(x, y, m, vol, bd, nl, p["K"], p["mu"], p["horizon"])
    out["force"] = _errs(f_g, f_o)
    out["dilatation"] = _errs(th_g, th_o)
    if dmg_gpu is not None:
        bd_o, _, near_o = O.compute_damage(x, y, bd, nl, sc)
        s = O.bond_stretch(x, y, nl)
        band = np.abs(s - sc) <= 1e-6 * sc
        mism = (dmg_gpu[0] != bd_o)
        out["damage"] = dict(mismatch_outside_band=int((mism & ~band).sum()),
                             mismatch_inside_band=int((mism & band).sum()),
                             near_critical_gpu=int(dmg_gpu[2]), near_critical_oracle=int(near_o),
                             broken_after=int(bd_o.sum()))
    if name == "bar":
        r = O.explicit_run(x, vol, p["density"], nl, p["K"], p["mu"], p["horizon"], sc,
                           p["dt"], p["nsteps"], v0=p["v0"], bcs=p["bcs"])
        out["displacement_after_run"] = _errs(st["u"], r["u"])
        out["velocity_after_run"] = _errs(st["v"], r["v"])
    out["total_s"] = time.perf_counter() - t0
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    names = sys.argv[1:] or ["bar", "plate", "cylinder", "batched"]
    for n in names:
        report(n)
