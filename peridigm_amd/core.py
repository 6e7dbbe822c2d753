"""Host-side handle on a pdgpu context (one device, one stream).

Thin numpy-facing wrapper of the C ABI: every array is passed in Peridigm's
own layout (FP64, xyz-interleaved node vectors, bond fields in
neighbourhood-list order); device mirrors live in the context.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PdgError, pdg_bc


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """pdg_ctx on HIP device `device`."""

    def __init__(self, device: int = 0, lanes: int = 4, variant: int = 0, tile_nodes: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.pdg_ctx_create(device, C.byref(h))
        if rc != 0:
            raise PdgError(rc, "pdg_ctx_create failed (no HIP device?)")
        self.h = h
        self.N = 0
        self.B = 0
        self._keep = []
        self.check(self.lib.pdg_set_lanes_per_node(self.h, lanes))
        self.check(self.lib.pdg_set_kernel_variant(self.h, variant))
        self.check(self.lib.pdg_set_tile_nodes(self.h, tile_nodes))

    # ------------------------------------------------------------ plumbing
    def check(self, rc):
        if rc != 0:
            msg = self.lib.pdg_last_error(self.h)
            raise PdgError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.pdg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ----------------------------------------------------------- topology
    def set_topology(self, neighborhood_list, num_owned=None, owned_ids=None):
        nl = _i32(neighborhood_list)
        if num_owned is None:
            num_owned = count_points(nl)
        ids = None if owned_ids is None else _i32(owned_ids)
        self.check(self.lib.pdg_set_topology(self.h, int(num_owned), _p(ids), _p(nl), nl.size))
        n = C.c_int(); b = C.c_longlong()
        self.check(self.lib.pdg_get_sizes(self.h, C.byref(n), C.byref(b)))
        self.N, self.B = n.value, b.value

    def set_constants(self, x, volume, K, mu, horizon, influence=0, critical_stretch=0.0):
        x = _f64(x, (self.N, 3)); v = _f64(volume, (self.N,))
        if isinstance(influence, str):
            influence = _lib.INFLUENCE[influence]
        self.check(self.lib.pdg_set_constants(self.h, _p(x), _p(v), float(K), float(mu),
                                              float(horizon), int(influence),
                                              float(critical_stretch)))

    def set_critical_stretch(self, sc):
        """Per-node critical stretch (multi-specimen runs)."""
        v = _f64(np.broadcast_to(np.asarray(sc, float), (self.N,)))
        self.check(self.lib.pdg_set_critical_stretch(self.h, _p(v)))

    def set_material_model(self, model, planar=0, height=0.0):
        """model: "Elastic" (0), "Elastic Bond Based" (1) or "Linear LPS" (2,
        the linearized LPS of "Linear LPS Partial Volume"); planar: 0 3D,
        1 plane strain, 2 plane stress (bond-based only). Call before
        set_constants."""
        codes = {"Elastic": 0, "Elastic Bond Based": 1, "Linear LPS": 2,
                 "Linear LPS Partial Volume": 2, "Elastic Partial Volume": 0}
        m = codes.get(model, model)
        self.check(self.lib.pdg_set_material_model(self.h, int(m), int(planar), float(height)))

    def set_critical_stretch_value(self, sc):
        """Uniform critical stretch for damage-enabled nodes (<= 0: per-node values)."""
        self.check(self.lib.pdg_set_critical_stretch_value(self.h, float(sc)))

    def set_critical_stretch_function(self, times, values):
        """Critical stretch as a function of time for explicit runs
        (piecewise linear through the points; empty removes it)."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._sc_keep = (t, v)
        self.check(self.lib.pdg_set_critical_stretch_function(self.h, t.size, _p(t), _p(v)))

    def set_density(self, rho):
        r = _f64(np.broadcast_to(np.asarray(rho, float), (self.N,)))
        self.check(self.lib.pdg_set_density(self.h, _p(r)))

    # -------------------------------------------------------- drop-in calls
    def initialize(self):
        m = np.zeros(self.N)
        self.check(self.lib.pdg_initialize(self.h, _p(m)))
        return m

    def compute_damage(self, y, bond_damage_n=None):
        y = _f64(y, (self.N, 3))
        bdn = None if bond_damage_n is None else _f64(bond_damage_n, (self.B,))
        out = np.zeros(self.B); dmg = np.zeros(self.N); near = C.c_longlong()
        self.check(self.lib.pdg_compute_damage(self.h, _p(y), _p(bdn), _p(out), _p(dmg),
                                               C.byref(near)))
        return out, dmg, near.value

    def compute_force(self, y, bond_damage=None, dt=0.0):
        y = _f64(y, (self.N, 3))
        bd = None if bond_damage is None else _f64(bond_damage, (self.B,))
        th = np.zeros(self.N); f = np.zeros((self.N, 3))
        self.check(self.lib.pdg_compute_force(self.h, float(dt), _p(y), _p(bd), _p(th), _p(f)))
        return th, f

    # ------------------------------------------------------------- explicit
    def upload_state(self, u=None, v=None, a=None, bond_damage=None):
        u = None if u is None else _f64(u, (self.N, 3))
        v = None if v is None else _f64(v, (self.N, 3))
        a = None if a is None else _f64(a, (self.N, 3))
        bd = None if bond_damage is None else _f64(bond_damage, (self.B,))
        self.check(self.lib.pdg_state_upload(self.h, _p(u), _p(v), _p(a), _p(bd)))

    def download_state(self):
        u = np.zeros((self.N, 3)); v = np.zeros((self.N, 3)); a = np.zeros((self.N, 3))
        bd = np.zeros(self.B); dmg = np.zeros(self.N); th = np.zeros(self.N)
        f = np.zeros((self.N, 3))
        self.check(self.lib.pdg_state_download(self.h, _p(u), _p(v), _p(a), _p(bd), _p(dmg),
                                               _p(th), _p(f)))
        return dict(u=u, v=v, a=a, bond_damage=bd, damage=dmg, theta=th, force=f)

    # ------------------------------------------------------------- thermal
    def set_thermal_expansion(self, coefficient):
        """"Thermal Expansion Coefficient" of the elastic model (0: off)."""
        self.check(self.lib.pdg_set_thermal_expansion(self.h, float(coefficient)))

    def set_temperature_change(self, dT0=None, dT1=None):
        """Temperature_Change per node, dT(t) = dT1 t + dT0 (None: zero)."""
        a = None if dT0 is None else _f64(np.broadcast_to(dT0, (self.N,)), (self.N,))
        b = None if dT1 is None else _f64(np.broadcast_to(dT1, (self.N,)), (self.N,))
        if a is None and b is not None:
            a = np.zeros(self.N)
        self.check(self.lib.pdg_set_temperature_change(self.h, _p(a), _p(b)))

    def set_temperature_time(self, t):
        self.check(self.lib.pdg_set_temperature_time(self.h, float(t)))

    def compute_partial_stress(self, y, bond_damage=None):
        """Partial_Stress (N, 3, 3) at configuration y (FP64)."""
        y = _f64(y, (self.N, 3))
        bd = None if bond_damage is None else _f64(bond_damage, (self.B,))
        ps = np.zeros((self.N, 3, 3))
        self.check(self.lib.pdg_compute_partial_stress(self.h, _p(y), _p(bd), _p(ps)))
        return ps

    # -------------------------------------------------------------- contact
    def set_contact(self, search_radius, search_frequency, contact_radius, spring_constant,
                    friction=0.0):
        """ContactManager + ShortRangeForceContactModel parameters
        (src/core/Peridigm_ContactManager.cpp, src/contact/
        Peridigm_ShortRangeForceContactModel.cpp); contact_radius <= 0 removes
        the model. verlet_run then searches and adds the contact force."""
        self.check(self.lib.pdg_set_contact(self.h, float(search_radius), int(search_frequency),
                                            float(contact_radius), float(spring_constant),
                                            float(friction)))

    def contact_search(self, y):
        """ContactManager::contactSearch of configuration y: flat [n, j...] list."""
        y = _f64(y, (self.N, 3))
        size = C.c_longlong(0)
        self.check(self.lib.pdg_contact_search(self.h, _p(y), None, C.byref(size)))
        out = np.zeros(max(size.value, 1), np.int32)
        self.check(self.lib.pdg_contact_search(self.h, _p(y), _p(out), C.byref(size)))
        return out[:size.value]

    def contact_compute_force(self, y, v, contact_list, dt=0.0):
        """ShortRangeForceContactModel::computeForce: contact force density (N, 3)."""
        y = _f64(y, (self.N, 3)); v = _f64(v, (self.N, 3))
        cl = _i32(contact_list)
        f = np.zeros((self.N, 3))
        self.check(self.lib.pdg_contact_compute_force(self.h, float(dt), _p(y), _p(v), _p(cl),
                                                      int(cl.size), _p(f)))
        return f

    def contact_force_density(self):
        """Contact force density (N, 3) of the last verlet_run evaluation."""
        f = np.zeros((self.N, 3))
        self.check(self.lib.pdg_contact_force_density(self.h, _p(f)))
        return f

    def bc_array(self, bcs):
        """ctypes pdg_bc array of (node_ids, coord, c0, c1): u_coord(t) = c1*t + c0;
        the node arrays are kept alive on the context."""
        arr = (pdg_bc * max(1, len(bcs)))()
        self._keep = []
        for k, (ids, coord, c0, c1) in enumerate(bcs):
            ids = _i32(ids)
            self._keep.append(ids)
            arr[k].nodes = ids.ctypes.data_as(C.POINTER(C.c_int))
            arr[k].numNodes = ids.size
            arr[k].coord = int(coord)
            arr[k].c0 = float(c0)
            arr[k].c1 = float(c1)
        return arr

    def set_bcs(self, bcs):
        """bcs: iterable of (node_ids, coord, c0, c1): u_coord(t) = c1*t + c0."""
        self.check(self.lib.pdg_set_bcs(self.h, self.bc_array(bcs), len(bcs)))

    def verlet_run(self, nsteps, dt, t0=0.0, initial_eval=True):
        self.check(self.lib.pdg_verlet_run(self.h, int(nsteps), float(dt), float(t0),
                                           int(bool(initial_eval))))

    def verlet_run_persistent(self, nsteps, dt, t0=0.0, initial_eval=True):
        """All steps in one persistent cooperative launch; returns the grid size."""
        g = C.c_int()
        self.check(self.lib.pdg_verlet_run_persistent(self.h, int(nsteps), float(dt), float(t0),
                                                      int(bool(initial_eval)), C.byref(g)))
        return g.value

    def set_kernel_variant(self, variant):
        """0: LDS-staged Hilbert tiles (default), 1: untiled gathers."""
        self.check(self.lib.pdg_set_kernel_variant(self.h, int(variant)))

    def tiling(self):
        n = C.c_int(); t = C.c_longlong()
        self.check(self.lib.pdg_get_tiling(self.h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def lds_schedule(self):
        """(16-lane group iterations, LDS conflict cost in neighbour order,
        cost in the scheduled order); cost = groups when conflict-free."""
        g = C.c_longlong(); b = C.c_longlong(); a = C.c_longlong()
        self.check(self.lib.pdg_lds_schedule(self.h, C.byref(g), C.byref(b), C.byref(a)))
        return g.value, b.value, a.value

    def tile_config(self):
        """(nodes per tile, lanes per node) the bond kernels run with; the
        tile size is 0 when the untiled kernels run."""
        t = C.c_int(); g = C.c_int()
        self.check(self.lib.pdg_tile_config(self.h, C.byref(t), C.byref(g)))
        return t.value, g.value

    def set_per_pass_timing(self, on=True):
        self.check(self.lib.pdg_set_per_pass_timing(self.h, int(bool(on))))

    def last_timing(self):
        t = C.c_double(); a = C.c_double(); b = C.c_double()
        self.check(self.lib.pdg_last_timing(self.h, C.byref(t), C.byref(a), C.byref(b)))
        return t.value, a.value, b.value

    def near_critical_count(self):
        c = C.c_longlong()
        self.check(self.lib.pdg_near_critical_count(self.h, C.byref(c)))
        return c.value

    def critical_time_step(self, density):
        d = C.c_double()
        self.check(self.lib.pdg_critical_time_step(self.h, float(density), C.byref(d)))
        return d.value


def count_points(nlist):
    """Number of points described by a flattened [n, j...] list."""
    nl = np.asarray(nlist)
    li = 0
    n = 0
    while li < nl.size:
        li += int(nl[li]) + 1
        n += 1
    return n


def list_to_csr(nlist, N):
    """Flattened [n, j...] list -> (row_ptr[N+1], col[B]) (vectorised)."""
    nl = np.asarray(nlist, dtype=np.int64)
    heads = np.zeros(N, np.int64)
    counts = np.zeros(N, np.int64)
    li = 0
    for i in range(N):
        heads[i] = li
        c = int(nl[li])
        counts[i] = c
        li += c + 1
    rp = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    mask = np.ones(nl.size, bool)
    mask[heads] = False
    col = nl[mask]
    return rp, col.astype(np.int32)


def csr_to_list(row_ptr, col):
    N = len(row_ptr) - 1
    deg = np.diff(row_ptr)
    out = np.empty(N + len(col), np.int32)
    heads = row_ptr[:-1] + np.arange(N)
    out[heads] = deg
    mask = np.ones(out.size, bool)
    mask[heads] = False
    out[mask] = col
    return out
