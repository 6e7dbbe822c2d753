"""FP64 CPU oracle (test infrastructure).

ctypes wrapper around ``oracle/liboracle.so`` (built from ``pd_oracle.cpp``,
which restates the reference loops operation for operation). Only the tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_quickgrid.restype = C.c_longlong
        L.orc_radius_search.restype = C.c_longlong
        L.orc_radius_search_brute.restype = C.c_longlong
        L.orc_filter_planes.restype = C.c_longlong
        L.orc_critical_time_step.restype = C.c_double
        L.orc_explicit_run.restype = C.c_int
        L.orc_explicit_run_sc.restype = C.c_int
        L.orc_explicit_run_ext.restype = C.c_int
        L.orc_plane_cuts.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- neighbours
def quickgrid(nx, ny, nz, origin, length, horizon, spherical=True):
    """QuickGrid tensor-product lattice: returns (x[N,3], vol[N], nlist)."""
    L = lib()
    args = [C.c_int(nx), C.c_int(ny), C.c_int(nz)] + [C.c_double(v) for v in origin] + \
           [C.c_double(v) for v in length] + [C.c_double(horizon), C.c_int(int(spherical))]
    size = L.orc_quickgrid(*args, None, None, None, C.c_longlong(0))
    N = nx * ny * nz
    x = np.zeros((N, 3))
    vol = np.zeros(N)
    nl = np.zeros(size, np.int32)
    L.orc_quickgrid(*args, _p(x), _p(vol), _p(nl), C.c_longlong(size))
    return x, vol, nl


def radius_search(x, R, include_self=False, brute=False):
    L = lib()
    x = _c(x, np.float64)
    N = x.shape[0]
    fn = L.orc_radius_search_brute if brute else L.orc_radius_search
    size = fn(C.c_int(N), _p(x), C.c_double(R), C.c_int(int(include_self)), None, C.c_longlong(0))
    nl = np.zeros(size, np.int32)
    fn(C.c_int(N), _p(x), C.c_double(R), C.c_int(int(include_self)), _p(nl), C.c_longlong(size))
    return nl


def plane_cuts(plane, p0, p1):
    return bool(lib().orc_plane_cuts(_p(_c(plane, np.float64)), _p(_c(p0, np.float64)),
                                     _p(_c(p1, np.float64))))


def filter_planes(x, nlist, planes):
    L = lib()
    x = _c(x, np.float64)
    nlist = _c(nlist, np.int32)
    planes = _c(planes, np.float64).reshape(-1, 11)
    N = x.shape[0]
    size = L.orc_filter_planes(C.c_int(N), _p(x), _p(nlist), C.c_int(planes.shape[0]),
                               _p(planes), None, C.c_longlong(0))
    out = np.zeros(size, np.int32)
    L.orc_filter_planes(C.c_int(N), _p(x), _p(nlist), C.c_int(planes.shape[0]), _p(planes),
                        _p(out), C.c_longlong(size))
    return out


# ----------------------------------------------------------------- kernels
def num_bonds(nlist, N):
    nlist = np.asarray(nlist)
    li = 0
    B = 0
    for _ in range(N):
        nn = int(nlist[li])
        li += nn + 1
        B += nn
    return B


def weighted_volume(x, vol, nlist, horizon, infl=0):
    x = _c(x, np.float64); vol = _c(vol, np.float64); nlist = _c(nlist, np.int32)
    m = np.zeros(x.shape[0])
    lib().orc_weighted_volume(C.c_int(x.shape[0]), _p(x), _p(vol), _p(nlist),
                              C.c_double(horizon), C.c_int(infl), _p(m))
    return m


def compute_force(x, y, m, vol, bond_damage, nlist, K, mu, horizon, infl=0):
    """ElasticMaterial::computeForce: returns (theta[N], force_density[N,3])."""
    x = _c(x, np.float64); y = _c(y, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    th = np.zeros(N)
    f = np.zeros((N, 3))
    lib().orc_compute_force(C.c_int(N), _p(x), _p(y), _p(m), _p(vol), _p(bd), _p(nlist),
                            C.c_double(K), C.c_double(mu), C.c_double(horizon), C.c_int(infl),
                            _p(th), _p(f))
    return th, f


def compute_force_thermal(x, y, m, vol, bond_damage, nlist, K, mu, horizon, alpha_t, delta_t,
                          infl=0, partial_stress=False):
    """ElasticMaterial::computeForce with thermal strains (delta_t per point)
    and optionally the partial stress: (theta, force_density, ps (N,3,3) or None)."""
    x = _c(x, np.float64); y = _c(y, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    dt = None if delta_t is None else _c(np.broadcast_to(delta_t, (N,)), np.float64)
    th = np.zeros(N); f = np.zeros((N, 3))
    ps = np.zeros((N, 3, 3)) if partial_stress else None
    lib().orc_compute_force_thermal(C.c_int(N), _p(x), _p(y), _p(m), _p(vol), _p(bd), _p(nlist),
                                    C.c_double(K), C.c_double(mu), C.c_double(horizon),
                                    C.c_int(infl), C.c_double(alpha_t), _p(dt), _p(th), _p(f),
                                    _p(ps))
    return th, f, ps


def force_directional_thermal(x, u, w, m, vol, bond_damage, nlist, K, mu, horizon, alpha_t,
                              delta_t, infl=0):
    """(f(u), J(u) w) with thermal strains, by forward-mode dual numbers."""
    x = _c(x, np.float64); u = _c(u, np.float64); w = _c(w, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    dt = None if delta_t is None else _c(np.broadcast_to(delta_t, (N,)), np.float64)
    f = np.zeros((N, 3)); jw = np.zeros((N, 3))
    lib().orc_force_directional_thermal(C.c_int(N), _p(x), _p(u), _p(w), _p(m), _p(vol), _p(bd),
                                        _p(nlist), C.c_double(K), C.c_double(mu),
                                        C.c_double(horizon), C.c_int(infl), C.c_double(alpha_t),
                                        _p(dt), _p(f), _p(jw))
    return f, jw


def compute_force_linear(x, y, m, vol, bond_damage, nlist, K, mu, horizon, infl=0):
    """Linearized LPS ("Linear LPS Partial Volume" without partial volumes):
    returns (theta[N], force_density[N,3])."""
    x = _c(x, np.float64); y = _c(y, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    th = np.zeros(N)
    f = np.zeros((N, 3))
    lib().orc_compute_force_linear(C.c_int(N), _p(x), _p(y), _p(m), _p(vol), _p(bd), _p(nlist),
                                   C.c_double(K), C.c_double(mu), C.c_double(horizon),
                                   C.c_int(infl), _p(th), _p(f))
    return th, f


def force_given_theta(x, y, m, vol, theta, bond_damage, nlist, K, mu, horizon, infl=0):
    x = _c(x, np.float64); y = _c(y, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); th = _c(theta, np.float64); bd = _c(bond_damage, np.float64)
    nlist = _c(nlist, np.int32)
    f = np.zeros((x.shape[0], 3))
    lib().orc_force_given_theta(C.c_int(x.shape[0]), _p(x), _p(y), _p(m), _p(vol), _p(th),
                                _p(bd), _p(nlist), C.c_double(K), C.c_double(mu),
                                C.c_double(horizon), C.c_int(infl), _p(f))
    return f


def compute_damage(x, y, bd_n, nlist, sc, near_rel=1e-6):
    """CriticalStretchDamageModel::computeDamage: returns (bd_np1, damage, n_near).
    sc: critical stretch, scalar or one value per node (bond i->j uses node i's)."""
    x = _c(x, np.float64); y = _c(y, np.float64); bd_n = _c(bd_n, np.float64)
    nlist = _c(nlist, np.int32)
    N = x.shape[0]
    out = np.zeros_like(bd_n)
    dmg = np.zeros(N)
    near = C.c_longlong(0)
    if np.ndim(sc) > 0:
        scn = _c(np.broadcast_to(sc, (N,)), np.float64)
        lib().orc_compute_damage_sc(C.c_int(N), _p(x), _p(y), _p(bd_n), _p(out), _p(dmg),
                                    _p(nlist), _p(scn), C.byref(near), C.c_double(near_rel))
    else:
        lib().orc_compute_damage(C.c_int(N), _p(x), _p(y), _p(bd_n), _p(out), _p(dmg),
                                 _p(nlist), C.c_double(sc), C.byref(near), C.c_double(near_rel))
    return out, dmg, near.value


def degrees(nlist, N):
    """Neighbour count of every point of a flattened [n, j...] list."""
    nl = np.asarray(nlist)
    deg = np.zeros(N, np.int64)
    li = 0
    for i in range(N):
        deg[i] = nl[li]
        li += int(nl[li]) + 1
    return deg


def bond_stretch(x, y, nlist):
    x = _c(x, np.float64); y = _c(y, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    s = np.zeros(num_bonds(nlist, N))
    lib().orc_bond_stretch(C.c_int(N), _p(x), _p(y), _p(nlist), _p(s))
    return s


def critical_time_step(x, vol, nlist, density, K, horizon):
    x = _c(x, np.float64); vol = _c(vol, np.float64); nlist = _c(nlist, np.int32)
    return lib().orc_critical_time_step(C.c_int(x.shape[0]), _p(x), _p(vol), _p(nlist),
                                        C.c_double(density), C.c_double(K), C.c_double(horizon))


def _bc_arrays(bcs):
    """bcs: list of (node_ids, coord, c0, c1) meaning u_coord(t) = c1*t + c0."""
    ptr = [0]
    nodes = []
    coord, c0, c1 = [], [], []
    for ids, c, a0, a1 in bcs:
        nodes.extend(int(i) for i in ids)
        ptr.append(len(nodes))
        coord.append(int(c)); c0.append(float(a0)); c1.append(float(a1))
    return (np.array(ptr, np.int32), np.array(nodes if nodes else [0], np.int32),
            np.array(coord if coord else [0], np.int32), np.array(c0 if c0 else [0.0]),
            np.array(c1 if c1 else [0.0]))


def explicit_run(x, vol, rho, nlist, K, mu, horizon, sc, dt, nsteps, u0=None, v0=None,
                 bond_damage=None, bcs=(), t0=0.0, infl=0):
    """Velocity-Verlet run (Peridigm::executeExplicit). Returns dict of final state.
    sc: critical stretch (scalar, <= 0: no damage) or one value per node."""
    x = _c(x, np.float64); vol = _c(vol, np.float64)
    N = x.shape[0]
    rho = _c(np.broadcast_to(rho, (N,)), np.float64)
    nlist = _c(nlist, np.int32)
    B = num_bonds(nlist, N)
    u = np.zeros((N, 3)) if u0 is None else _c(u0, np.float64).copy()
    v = np.zeros((N, 3)) if v0 is None else _c(v0, np.float64).copy()
    bd = np.zeros(B) if bond_damage is None else _c(bond_damage, np.float64).copy()
    a = np.zeros((N, 3)); th = np.zeros(N); f = np.zeros((N, 3)); dmg = np.zeros(N); m = np.zeros(N)
    ptr, nodes, coord, c0, c1 = _bc_arrays(bcs)
    sc_node = None
    if np.ndim(sc) > 0:
        sc_node = _c(sc, np.float64)
        sc = float(sc_node.max())
    rc = lib().orc_explicit_run_sc(C.c_int(N), _p(x), _p(vol), _p(rho), _p(nlist), C.c_double(K),
                                C.c_double(mu), C.c_double(horizon), C.c_int(infl),
                                C.c_double(sc), None if sc_node is None else _p(sc_node),
                                C.c_double(dt), C.c_double(t0), C.c_int(nsteps),
                                C.c_int(len(bcs)), _p(ptr), _p(nodes), _p(coord), _p(c0), _p(c1),
                                _p(u), _p(v), _p(a), _p(bd), _p(dmg), _p(th), _p(f), _p(m))
    if rc != 0:
        raise FloatingPointError("oracle explicit run: non-finite force")
    return dict(u=u, v=v, a=a, bond_damage=bd, damage=dmg, theta=th, force=f, m=m)


_BLIB = None


def baseline_lib():
    """libbaseline.so: the multi-threaded CPU baseline (pd_baseline.cpp)."""
    global _BLIB
    if _BLIB is None:
        path = os.path.join(_HERE, "libbaseline.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.bl_explicit_run.restype = C.c_int
        L.bl_max_threads.restype = C.c_int
        _BLIB = L
    return _BLIB


def baseline_explicit_run(x, vol, rho, nlist, m, K, mu, sc, dt, nsteps, u0=None, v0=None,
                          bond_damage=None, bcs=(), t0=0.0, threads=0):
    """Multi-threaded FP64 velocity-Verlet run ("One" influence function,
    scalar critical stretch, weighted volume m given). threads <= 0: all the
    threads OpenMP allows. Returns the final state like explicit_run."""
    x = _c(x, np.float64); vol = _c(vol, np.float64); m = _c(m, np.float64)
    N = x.shape[0]
    rho = _c(np.broadcast_to(rho, (N,)), np.float64)
    nlist = _c(nlist, np.int32)
    B = num_bonds(nlist, N)
    u = np.zeros((N, 3)) if u0 is None else _c(u0, np.float64).copy()
    v = np.zeros((N, 3)) if v0 is None else _c(v0, np.float64).copy()
    bd = np.zeros(B) if bond_damage is None else _c(bond_damage, np.float64).copy()
    a = np.zeros((N, 3)); th = np.zeros(N); f = np.zeros((N, 3)); dmg = np.zeros(N)
    ptr, nodes, coord, c0, c1 = _bc_arrays(bcs)
    rc = baseline_lib().bl_explicit_run(
        C.c_int(threads), C.c_int(N), _p(x), _p(vol), _p(rho), _p(nlist), _p(m), C.c_double(K),
        C.c_double(mu), C.c_double(sc), C.c_double(dt), C.c_double(t0), C.c_int(nsteps),
        C.c_int(len(bcs)), _p(ptr), _p(nodes), _p(coord), _p(c0), _p(c1), _p(u), _p(v), _p(a),
        _p(bd), _p(dmg), _p(th), _p(f))
    if rc != 0:
        raise FloatingPointError("baseline explicit run: non-finite force")
    return dict(u=u, v=v, a=a, bond_damage=bd, damage=dmg, theta=th, force=f)


def baseline_threads():
    return int(baseline_lib().bl_max_threads())


def compute_force_pmb(x, y, vol, bond_damage, nlist, K, horizon, planar=0, height=0.0):
    """Elastic bond-based force density (src/materials/elastic_bond_based.cxx:55-130)."""
    x = _c(x, np.float64); y = _c(y, np.float64); vol = _c(vol, np.float64)
    bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    f = np.zeros((x.shape[0], 3))
    lib().orc_compute_force_pmb(C.c_int(x.shape[0]), _p(x), _p(y), _p(vol), _p(bd), _p(nlist),
                                C.c_double(K), C.c_double(horizon), C.c_int(planar),
                                C.c_double(height), _p(f))
    return f


def contact_search(y, R, nlist):
    """Contact neighbour list (flat [n, j...]) of the configuration y: points
    within R, bonded neighbours and the point itself excluded."""
    y = _c(y, np.float64); nlist = _c(nlist, np.int32)
    N = y.shape[0]
    sz = C.c_longlong(0)
    lib().orc_contact_search(C.c_int(N), _p(y), C.c_double(R), _p(nlist), None, C.c_longlong(0),
                             C.byref(sz))
    out = np.zeros(sz.value, np.int32)
    lib().orc_contact_search(C.c_int(N), _p(y), C.c_double(R), _p(nlist), _p(out),
                             C.c_longlong(sz.value), C.byref(sz))
    return out


def contact_force(y, v, vol, clist, rc, ks, mu_f, horizon):
    """Short-range contact force density for the contact list clist."""
    y = _c(y, np.float64); v = _c(v, np.float64); vol = _c(vol, np.float64)
    clist = _c(clist, np.int32)
    f = np.zeros((y.shape[0], 3))
    lib().orc_contact_force(C.c_int(y.shape[0]), _p(y), _p(v), _p(vol), _p(clist),
                            C.c_double(rc), C.c_double(ks), C.c_double(mu_f), C.c_double(horizon),
                            _p(f))
    return f


def explicit_run_ext(x, vol, rho, nlist, K, mu, horizon, sc, dt, nsteps, material=0, planar=0,
                     height=0.0, sc_eval=None, u0=None, v0=None, bond_damage=None, bcs=(),
                     t0=0.0, infl=0, contact=None):
    """explicit_run with a material model (0 elastic LPS, 1 elastic bond
    based), optional per-evaluation critical stretches sc_eval (nsteps+1) and
    optional short-range contact: contact = dict(search_radius,
    search_frequency, contact_radius, spring_constant, friction)."""
    x = _c(x, np.float64); vol = _c(vol, np.float64)
    N = x.shape[0]
    rho = _c(np.broadcast_to(rho, (N,)), np.float64)
    nlist = _c(nlist, np.int32)
    B = num_bonds(nlist, N)
    u = np.zeros((N, 3)) if u0 is None else _c(u0, np.float64).copy()
    v = np.zeros((N, 3)) if v0 is None else _c(v0, np.float64).copy()
    bd = np.zeros(B) if bond_damage is None else _c(bond_damage, np.float64).copy()
    a = np.zeros((N, 3)); th = np.zeros(N); f = np.zeros((N, 3)); dmg = np.zeros(N); m = np.zeros(N)
    fc = np.zeros((N, 3))
    ptr, nodes, coord, c0, c1 = _bc_arrays(bcs)
    sce = None if sc_eval is None else _c(sc_eval, np.float64)
    if sce is not None:
        assert sce.size == nsteps + 1
    ct = contact or {}
    rc = lib().orc_explicit_run_ext(
        C.c_int(N), _p(x), _p(vol), _p(rho), _p(nlist), C.c_int(material), C.c_int(planar),
        C.c_double(height), C.c_double(K), C.c_double(mu), C.c_double(horizon), C.c_int(infl),
        C.c_double(sc), None if sce is None else _p(sce), C.c_double(dt), C.c_double(t0),
        C.c_int(nsteps), C.c_int(len(bcs)), _p(ptr), _p(nodes), _p(coord), _p(c0), _p(c1),
        C.c_double(ct.get("search_radius", 0.0)), C.c_int(ct.get("search_frequency", 0)),
        C.c_double(ct.get("contact_radius", 0.0)), C.c_double(ct.get("spring_constant", 0.0)),
        C.c_double(ct.get("friction", 0.0)),
        _p(u), _p(v), _p(a), _p(bd), _p(dmg), _p(th), _p(f), _p(fc), _p(m))
    if rc != 0:
        raise FloatingPointError("oracle explicit run: non-finite force")
    return dict(u=u, v=v, a=a, bond_damage=bd, damage=dmg, theta=th, force=f, m=m,
                contact_force=fc)


def force_at(x, u, m, vol, bond_damage, nlist, K, mu, horizon, infl=0):
    x = _c(x, np.float64); u = _c(u, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    f = np.zeros((x.shape[0], 3))
    lib().orc_force_at(C.c_int(x.shape[0]), _p(x), _p(u), _p(m), _p(vol), _p(bd), _p(nlist),
                       C.c_double(K), C.c_double(mu), C.c_double(horizon), C.c_int(infl), _p(f))
    return f


def force_directional(x, u, w, m, vol, bond_damage, nlist, K, mu, horizon, infl=0):
    """(f(u), J(u) w) by forward-mode dual numbers."""
    x = _c(x, np.float64); u = _c(u, np.float64); w = _c(w, np.float64); m = _c(m, np.float64)
    vol = _c(vol, np.float64); bd = _c(bond_damage, np.float64); nlist = _c(nlist, np.int32)
    N = x.shape[0]
    f = np.zeros((N, 3)); jw = np.zeros((N, 3))
    lib().orc_force_directional(C.c_int(N), _p(x), _p(u), _p(w), _p(m), _p(vol), _p(bd),
                                _p(nlist), C.c_double(K), C.c_double(mu), C.c_double(horizon),
                                C.c_int(infl), _p(f), _p(jw))
    return f, jw


# --------------------------------------------------------------------- gnn
def gnn_forward(row_ptr, col, ew, dims, params, feat):
    row_ptr = _c(row_ptr, np.int32); col = _c(col, np.int32); ew = _c(ew, np.float64)
    dims = _c(dims, np.int32); params = _c(params, np.float64); feat = _c(feat, np.float64)
    N = row_ptr.shape[0] - 1
    L = dims.shape[0] - 1
    acts = np.zeros(int(sum(N * d for d in dims)))
    lib().orc_gnn_forward(C.c_int(N), _p(row_ptr), _p(col), _p(ew), C.c_int(L), _p(dims),
                          _p(params), _p(feat), _p(acts))
    out = acts[-N * int(dims[-1]):].reshape(N, int(dims[-1]))
    return out, acts


def gnn_backward(row_ptr, col, ew, dims, params, acts, gout):
    row_ptr = _c(row_ptr, np.int32); col = _c(col, np.int32); ew = _c(ew, np.float64)
    dims = _c(dims, np.int32); params = _c(params, np.float64); acts = _c(acts, np.float64)
    gout = _c(gout, np.float64)
    N = row_ptr.shape[0] - 1
    dp = np.zeros_like(params)
    lib().orc_gnn_backward(C.c_int(N), _p(row_ptr), _p(col), _p(ew), C.c_int(dims.shape[0] - 1),
                           _p(dims), _p(params), _p(acts), _p(gout), _p(dp))
    return dp


# ------------------------------------------------------------------ helpers
def list_to_csr(nlist, N):
    """Flattened [n, j...] list -> (row_ptr[N+1], col[B])."""
    nlist = np.asarray(nlist, dtype=np.int64)
    counts = np.zeros(N, np.int64)
    heads = np.zeros(N, np.int64)
    li = 0
    for i in range(N):
        heads[i] = li
        counts[i] = nlist[li]
        li += int(nlist[li]) + 1
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(counts)
    idx = np.concatenate([np.arange(h + 1, h + 1 + c) for h, c in zip(heads, counts)]) if N else np.zeros(0, np.int64)
    col = nlist[idx] if idx.size else np.zeros(0, np.int64)
    return row_ptr.astype(np.int32), col.astype(np.int32)
