"""Discretizations and neighbour lists on the device.

``quickgrid`` mirrors PdQuickGridDiscretization's TensorProduct3DMeshGenerator
(src/io/discretization/Peridigm_PdQuickGridDiscretization.cpp:139-232,
src/io/mesh_input/quick_grid/QuickGrid.cxx:510-647) and ``radius_search``
the TextFile discretization's fixed-radius search plus bond filters
(src/io/discretization/Peridigm_TextFileDiscretization.cpp:282-286,
src/io/pdneigh/NeighborhoodList.cxx:621-761). Both run on the GPU through the
C ABI; the lists are returned in Peridigm's flattened [n, j...] layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PdgError


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def quickgrid(n, origin, length, horizon, spherical=True):
    """Returns (x[N,3], volume[N], neighborhood_list)."""
    lib = _lib.load()
    nx, ny, nz = (int(v) for v in n)
    o = np.ascontiguousarray(origin, np.float64)
    L = np.ascontiguousarray(length, np.float64)
    size = C.c_longlong(0)
    rc = lib.pdg_quickgrid(nx, ny, nz, _p(o), _p(L), float(horizon), int(bool(spherical)),
                           None, None, None, C.byref(size))
    if rc:
        raise PdgError(rc, "pdg_quickgrid (size)")
    N = nx * ny * nz
    x = np.zeros((N, 3)); vol = np.zeros(N); nl = np.zeros(size.value, np.int32)
    rc = lib.pdg_quickgrid(nx, ny, nz, _p(o), _p(L), float(horizon), int(bool(spherical)),
                           _p(x), _p(vol), _p(nl), C.byref(size))
    if rc:
        raise PdgError(rc, "pdg_quickgrid")
    return x, vol, nl


def radius_search(x, radius, planes=None):
    """|x_j - x_i|^2 <= R^2 neighbours (sorted ids), minus bonds cut by planes.

    planes: array (P, 11) of FinitePlane parameters (normal, lower-left corner,
    bottom unit vector, bottom length, side length)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    N = x.shape[0]
    pl = np.zeros((0, 11)) if planes is None else np.ascontiguousarray(planes, np.float64).reshape(-1, 11)
    size = C.c_longlong(0)
    rc = lib.pdg_radius_search(N, _p(x), float(radius), pl.shape[0], _p(pl) if pl.size else None,
                               None, C.byref(size))
    if rc:
        raise PdgError(rc, "pdg_radius_search (size)")
    nl = np.zeros(size.value, np.int32)
    rc = lib.pdg_radius_search(N, _p(x), float(radius), pl.shape[0], _p(pl) if pl.size else None,
                               _p(nl), C.byref(size))
    if rc:
        raise PdgError(rc, "pdg_radius_search")
    return nl


def lattice_centres(n, origin, length):
    """QuickGrid cell centres by accumulation (p += cellSize), gid order i + j nx + k nx ny."""
    cs = [length[d] / n[d] for d in range(3)]
    c = []
    for d in range(3):
        v = np.empty(n[d])
        p = origin[d] + cs[d] / 2.0
        for i in range(n[d]):
            v[i] = p
            p += cs[d]
        c.append(v)
    X = np.empty((n[0] * n[1] * n[2], 3))
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    X[:, 0] = c[0][i.ravel()]
    X[:, 1] = c[1][j.ravel()]
    X[:, 2] = c[2][k.ravel()]
    vol = np.full(X.shape[0], cs[0] * cs[1] * cs[2])
    return X, vol
