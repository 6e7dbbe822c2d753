"""Pins the oracle's contact search and short-range contact force beyond the
two contact decks' gold (tests/test_oracle_golden.py):

- the search against scipy's KD-tree pair query (distance <= R), bonded
  pairs removed (src/core/Peridigm_ContactManager.cpp:700-750);
- the force (src/contact/Peridigm_ShortRangeForceContactModel.cpp:98-258)
  against what the model's form fixes: the closed-form normal force of one
  pair, zero beyond the contact radius, linear momentum conserved
  (sum V_i f_i = 0, friction included), a friction force of magnitude mu
  times the normal one that opposes the relative tangential velocity."""
import numpy as np
from scipy.spatial import cKDTree

import oracle as O


def _bonds(n, pairs):
    rows = [[] for _ in range(n)]
    for a, b in pairs:
        rows[a].append(b)
        rows[b].append(a)
    return np.array([v for r in rows for v in [len(r)] + sorted(r)], np.int32)


def test_contact_search_against_kdtree():
    rng = np.random.default_rng(2)
    n = 400
    y = rng.uniform(0.0, 8.0, (n, 3))
    R = 0.9
    allp = sorted(cKDTree(y).query_pairs(R))
    bonded = {allp[k] for k in rng.choice(len(allp), len(allp) // 3, replace=False)}
    nl = _bonds(n, bonded)
    got = O.contact_search(y, R, nl)
    rp, col = O.list_to_csr(got, n)
    want = [[] for _ in range(n)]
    for a, b in allp:
        if (a, b) not in bonded:
            want[a].append(b)
            want[b].append(a)
    for i in range(n):
        assert list(col[rp[i]:rp[i + 1]]) == sorted(want[i]), i
    assert rp[-1] > 100


def test_contact_force_single_pair_closed_form():
    rc, ks, hz = 0.5, 3.0e4, 0.6
    y = np.array([[0.0, 0.0, 0.0], [0.24, 0.32, 0.0]])  # distance 0.4 < rc
    v = np.zeros((2, 3))
    vol = np.array([2.0, 5.0])
    cl = np.array([1, 1, 1, 0], np.int32)
    f = O.contact_force(y, v, vol, cl, rc, ks, 0.0, hz)
    c = 9.0 * ks / (np.pi * hz ** 4)
    mag = c * (rc - 0.4) / hz
    n = (y[1] - y[0]) / 0.4
    # symmetric list: each pair acts twice, pushing the points apart
    assert np.allclose(f[0], -2 * mag * vol[1] * n, rtol=1e-14, atol=0)
    assert np.allclose(f[1], 2 * mag * vol[0] * n, rtol=1e-14, atol=0)
    # a one-sided list acts once
    f1 = O.contact_force(y, v, vol, np.array([1, 1, 0], np.int32), rc, ks, 0.0, hz)
    assert np.allclose(f1, f / 2, rtol=1e-14, atol=0)
    # beyond the contact radius: nothing
    y2 = y.copy(); y2[1] = [0.3, 0.4, 0.0]             # distance 0.5 = rc (strict <)
    assert np.all(O.contact_force(y2, v, vol, cl, rc, ks, 0.0, hz) == 0.0)


def test_contact_force_momentum_and_friction():
    rng = np.random.default_rng(9)
    n = 200
    y = rng.uniform(0.0, 4.0, (n, 3))
    v = rng.standard_normal((n, 3))
    vol = rng.uniform(0.5, 1.5, n)
    nl = np.zeros(n, np.int32)
    cl = O.contact_search(y, 0.6, nl)
    rc, ks, hz = 0.45, 1.0e3, 0.5
    for mu in (0.0, 0.4):
        f = O.contact_force(y, v, vol, cl, rc, ks, mu, hz)
        p = (vol[:, None] * f).sum(0)
        assert np.abs(p).max() <= 1e-12 * np.abs(vol[:, None] * f).sum()
    # friction of one pair: magnitude mu |normal|, against the relative
    # tangential velocity
    y = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0]])
    v = np.array([[0.0, 2.0, 1.0], [5.0, -1.0, 0.0]])
    vol = np.array([1.0, 3.0])
    cl = np.array([1, 1, 0], np.int32)
    fn = O.contact_force(y, v, vol, cl, rc, ks, 0.0, hz)
    ff = O.contact_force(y, v, vol, cl, rc, ks, 0.25, hz) - fn
    assert np.allclose(ff[:, 0], 0.0, atol=1e-9 * np.abs(fn).max())  # no normal part
    for i in range(2):
        assert np.isclose(np.linalg.norm(ff[i]), 0.25 * np.linalg.norm(fn[i]), rtol=1e-13)
    vt = v[0, 1:] - v[1, 1:]
    assert ff[0, 1:] @ vt < 0 and ff[1, 1:] @ vt > 0
