"""Host-side deck and Exodus handling (CPU): XML parameter lists, boundary
condition expressions, elastic constants from any two moduli, the Exodus
mesh reader and discretization against the reference's gold outputs, and
the Exodus results writer."""
import os

import numpy as np
import pytest

from peridigm_amd import deck as D
from peridigm_amd import exodus as EX

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def test_parse_wave_in_bar_deck():
    d = D.parse_deck(os.path.join(DECKS, "WaveInBar.xml"))
    assert d["Discretization"]["Type"] == "Exodus"
    mat = d["Materials"]["My Elastic Material"]
    assert mat["Bulk Modulus"] == 14.90e9 and mat["Density"] == 2200.0
    assert d["Solver"]["Verlet"]["Fixed dt"] == 9.0659e-08
    assert d["Output"]["Output Frequency"] == 20
    assert d["Output"]["Output Variables"]["Dilatation"] is True


def test_expressions():
    x = np.array([0.0, 1.0, 2.0]); y = np.array([3.0, -1.0, 0.5]); z = np.zeros(3)
    assert np.array_equal(D.evaluate("value = -100.0", x, y, z, 0.0), [-100.0] * 3)
    assert np.allclose(D.evaluate("y*0.01*t", x, y, z, 2.0), y * 0.02)
    assert np.allclose(D.evaluate("value = sin(x) + pow(y, 2.0) - (t/4)", x, y, z, 1.0),
                       np.sin(x) + y ** 2 - 0.25)
    c0, c1 = D.linear_in_t("y*0.01*t", x, y, z)
    assert np.array_equal(c0, np.zeros(3)) and np.allclose(c1, 0.01 * y)
    c0, c1 = D.linear_in_t("3.0 - 2.0*t", x, y, z)
    assert np.allclose(c0, 3.0) and np.allclose(c1, -2.0)
    with pytest.raises(D.DeckError):
        D.linear_in_t("t*t", x, y, z)
    with pytest.raises(D.DeckError):
        D.evaluate("__import__('os').getcwd()", x, y, z, 0.0)


def test_elastic_constants_from_any_two():
    """K and mu from each pair of constants agree with the closed forms
    (E = 9K mu / (3K + mu), nu = (3K - 2 mu) / (2 (3K + mu)))."""
    K, mu = 130e9, 78e9
    E = 9 * K * mu / (3 * K + mu)
    nu = (3 * K - 2 * mu) / (2 * (3 * K + mu))
    pairs = [{"Bulk Modulus": K, "Shear Modulus": mu}, {"Young's Modulus": E, "Shear Modulus": mu},
             {"Young's Modulus": E, "Poisson's Ratio": nu}, {"Shear Modulus": mu, "Poisson's Ratio": nu},
             {"Bulk Modulus": K, "Young's Modulus": E}, {"Bulk Modulus": K, "Poisson's Ratio": nu}]
    for p in pairs:
        k, g = D.elastic_constants(p)
        assert abs(k - K) <= 1e-9 * K and abs(g - mu) <= 1e-9 * mu, p
    with pytest.raises(D.DeckError):
        D.elastic_constants({"Bulk Modulus": K})


def test_exodus_discretization_matches_gold_wave_in_bar(golden_dir):
    g = np.load(os.path.join(golden_dir, "wave_in_bar.npz"))
    mesh = EX.read_mesh(os.path.join(DECKS, "WaveInBar.g"))
    x, vol, blk, sets = EX.discretize(mesh)
    assert np.abs(x - g["coords"]).max() <= float(g["tol_coordinates"])
    # 24 tetrahedra summed against the box extents' product: round-off apart
    assert np.abs(vol - g["volume"]).max() <= 1e-13 * vol.max()
    assert np.array_equal(np.sort(sets["nodelist_1"]), np.sort(g["initial_velocity_elements"]))


def test_exodus_discretization_matches_gold_precracked_plate(golden_dir):
    g = np.load(os.path.join(golden_dir, "precracked_plate.npz"))
    mesh = EX.read_mesh(os.path.join(DECKS, "Plate.g"))
    x, vol, blk, sets = EX.discretize(mesh)
    assert np.abs(x - g["coords"]).max() <= 1e-12
    assert np.abs(vol - g["Volume"][0]).max() <= float(g["tol_default"])
    assert {k: len(v) for k, v in sets.items()} == \
        {"nodelist_1": 120, "nodelist_2": 120, "nodelist_3": 10, "nodelist_4": 4}


def test_hex_volume_closed_forms():
    """An axis-aligned box gives its extent product and centre; a sheared
    (parallelepiped) hex gives |det| of its edge vectors."""
    box = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0],
                    [0, 0, 5], [2, 0, 5], [2, 3, 5], [0, 3, 5]], float)
    c, v = EX.hex_centroid_volume(box[None])
    assert abs(v[0] - 30.0) <= 1e-12 and np.allclose(c[0], [1.0, 1.5, 2.5])
    a, b, e = np.array([1.0, 0.2, 0.0]), np.array([0.3, 1.0, 0.1]), np.array([0.1, -0.2, 1.5])
    par = np.array([0 * a, a, a + b, b, e, a + e, a + b + e, b + e])
    c, v = EX.hex_centroid_volume(par[None])
    assert abs(v[0] - abs(np.linalg.det(np.stack([a, b, e])))) <= 1e-12
    assert np.allclose(c[0], 0.5 * (a + b + e))


def test_results_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((7, 3))
    nod = {"DisplacementX": rng.standard_normal((3, 7)), "DisplacementY": rng.standard_normal((3, 7))}
    elem = {"Dilatation": rng.standard_normal((3, 7))}
    p = str(tmp_path / "out.e")
    EX.write_results(p, x, [0.0, 0.5, 1.0], nod, elem)
    t, xr, vals = EX.read_results(p)
    assert np.array_equal(t, [0.0, 0.5, 1.0]) and np.array_equal(xr, x)
    for k, v in {**nod, **elem}.items():
        assert np.array_equal(vals[k], v)
