"""Reference regression decks run end to end on the GPU (Exodus mesh in,
device solve, Exodus frames out) against the reference's gold outputs.

WaveInBar (explicit LPS, 220 Verlet steps, 12 frames): the bond arithmetic is
FP32-compensated, so the comparison uses this build's tolerances instead of
the .comp file's FP64 ones (5e-14 m on displacements): displacement and
velocity within 1e-6 of their maxima, force density and dilatation within
1e-5 of theirs; the weighted volume is FP64 (1e-12 relative).

PrecrackedPlate (one quasi-static load step to 1% strain, a rectangular-plane
bond filter, absolute Newton tolerance 1.0): the quasi-static path is FP64,
and the frames meet the .comp file's own absolute tolerances: neighbour
counts exact, displacement 1e-12, dilatation 1e-13, force 1.0."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_wave_in_bar_deck_against_gold(golden_dir, tmp_path):
    from peridigm_amd import deck as D
    from peridigm_amd import exodus as EX
    g = np.load(os.path.join(golden_dir, "wave_in_bar.npz"))
    out = str(tmp_path / "WaveInBar.e")
    res = D.run_deck(os.path.join(DECKS, "WaveInBar.xml"), output=out)
    t, x, vals = EX.read_results(out)
    assert len(t) == 12 and np.allclose(t, g["time"], rtol=1e-12, atol=1e-20)
    assert np.abs(x - g["coords"]).max() <= 1e-12
    errs = {}
    for name, tol in [("DisplacementX", 1e-6), ("DisplacementY", 1e-6), ("DisplacementZ", 1e-6),
                      ("VelocityX", 1e-6), ("VelocityY", 1e-6), ("VelocityZ", 1e-6),
                      ("Force_DensityX", 1e-5), ("Force_DensityY", 1e-5), ("Force_DensityZ", 1e-5),
                      ("Dilatation", 1e-5), ("Weighted_Volume", 1e-12)]:
        errs[name] = _rel(vals[name], g[name])
    print("WaveInBar max errors relative to each field's max:", errs)
    for name, tol in [("DisplacementX", 1e-6), ("VelocityX", 1e-6), ("Force_DensityX", 1e-5),
                      ("Dilatation", 1e-5), ("Weighted_Volume", 1e-12)]:
        assert errs[name] <= tol, (name, errs[name])
    assert np.all(vals["Damage"] == 0.0)


def test_precracked_plate_deck_against_gold(golden_dir, tmp_path):
    from peridigm_amd import deck as D
    from peridigm_amd import exodus as EX
    g = np.load(os.path.join(golden_dir, "precracked_plate.npz"))
    out = str(tmp_path / "PrecrackedPlate.e")
    res = D.run_deck(os.path.join(DECKS, "PrecrackedPlate.xml"), output=out)
    t, x, vals = EX.read_results(out)
    assert np.array_equal(t, g["time"])
    assert np.array_equal(vals["Number_Of_Neighbors"], g["Number_Of_Neighbors"])
    assert np.abs(vals["Volume"] - g["Volume"]).max() <= float(g["tol_default"])
    eu = max(_rel(vals[k][1], g[k][1]) for k in ("DisplacementX", "DisplacementY", "DisplacementZ"))
    ed = float(np.abs(vals["Dilatation"][1] - g["Dilatation"][1]).max())
    ef = max(float(np.abs(vals[k][1] - g[k][1]).max()) for k in ("ForceX", "ForceY", "ForceZ"))
    print(f"PrecrackedPlate: steps {res['info']['steps']}; displacement max rel {eu:.2e}, "
          f"dilatation max abs {ed:.2e}, force max abs {ef:.3g}")
    for k in ("DisplacementX", "DisplacementY", "DisplacementZ"):
        assert np.abs(vals[k] - g[k]).max() <= float(g["tol_displacement"]), k
    assert ed <= float(g["tol_dilatation"])
    assert ef <= float(g["tol_force"])
