"""Synthetic workloads of the explicit and quasi-static benchmarks.

Each builder returns a dict with the geometry (x, volume), the material and
damage constants, the initial velocity, the boundary conditions and the run
length. Neighbour lists are left to the caller (``neighbours`` below builds
them on the device), so the same geometry can feed the GPU path and a CPU
check.

  bar      100x10x10 LPS elastic bar, h = 1 mm, horizon 3.015h, no damage,
           v_x = 1 m/s on x < 5h, dt = 1e-8 s, 1000 steps.
  plate    400x400x4 notched plate, h = 0.5 mm, critical stretch 5e-4, a
           finite-plane notch 40h long from the x = 0 edge on the mid plane
           y = 200h (bonds crossing it removed, as a Rectangular_Plane bond
           filter does), top/bottom 3-node layers pulled at +-2 m/s as a
           prescribed displacement, 2000 steps.
  cylinder cubic lattice at h = 1 mm restricted to 40h <= r <= 50h,
           0 <= z < 200h, positions perturbed by U[-0.05h, 0.05h] (seeded),
           radial velocity 200 (1 + U[-0.01, 0.01]) m/s, critical stretch
           1.5e-3, 500 steps.
  plate_contact
           the plate with short-range contact on every node: search radius
           = horizon every 100 steps, contact radius 0.9h, spring constant
           15 K (the ratio of the reference's Contact_2x1x1 deck).
  qs_bar   60x20x20 lattice, h = 1 mm, quasi-static tension to 1e-3 strain in
           10 load steps with prescribed x displacement on 3-node end layers.
  qs_bar_200k
           the same tension on a 125x40x40 lattice (200,000 nodes), the
           quasi-static variant of the batched-sweep configuration.
"""
from __future__ import annotations

import numpy as np

from .discretization import lattice_centres, quickgrid, radius_search

STEEL = dict(density=7800.0, K=130.0e9, mu=78.0e9)


def bar(nx=100, ny=10, nz=10, h=1.0e-3):
    n = (nx, ny, nz)
    L = [nx * h, ny * h, nz * h]
    x, vol = lattice_centres(n, [0.0, 0.0, 0.0], L)
    v0 = np.zeros_like(x)
    v0[x[:, 0] < 5 * h, 0] = 1.0
    return dict(name="bar", n=n, h=h, origin=[0.0, 0.0, 0.0], length=L, x=x, volume=vol,
                horizon=3.015 * h, critical_stretch=0.0, dt=1.0e-8, nsteps=1000, v0=v0, bcs=[],
                search="quickgrid", planes=None, **STEEL)


def plate_notch_plane(h, nx, ny, nz, notch_cells=40):
    """FinitePlane parameters of an edge notch on y = (ny/2) h from x = 0.

    normal (0,1,0), bottom edge along +x, side along cross(ub, n) = +z. The
    tip sits 0.03h past notch_cells*h so no bond midpoint ties with it."""
    y0 = 0.5 * ny * h
    r0 = [-h, y0, -h]
    return np.array([0.0, 1.0, 0.0, r0[0], r0[1], r0[2], 1.0, 0.0, 0.0,
                     (notch_cells + 1.03) * h, (nz + 2) * h])


def plate(nx=400, ny=400, nz=4, h=5.0e-4, sc=5.0e-4, notch_cells=40, dt=2.5e-8, nsteps=2000,
          pull=2.0):
    n = (nx, ny, nz)
    L = [nx * h, ny * h, nz * h]
    x, vol = lattice_centres(n, [0.0, 0.0, 0.0], L)
    j = np.rint(x[:, 1] / h - 0.5).astype(int)
    top = np.nonzero(j >= ny - 3)[0]
    bot = np.nonzero(j <= 2)[0]
    bcs = [(top, 1, 0.0, pull), (bot, 1, 0.0, -pull)]
    return dict(name="plate", n=n, h=h, origin=[0.0, 0.0, 0.0], length=L, x=x, volume=vol,
                horizon=3.015 * h, critical_stretch=sc, dt=dt, nsteps=nsteps,
                v0=np.zeros_like(x), bcs=bcs, search="radius",
                planes=plate_notch_plane(h, nx, ny, nz, notch_cells)[None, :], **STEEL)


def plate_contact(**kw):
    p = plate(**kw)
    p["name"] = "plate_contact"
    p["contact"] = dict(search_radius=p["horizon"], search_frequency=100,
                        contact_radius=0.9 * p["h"], spring_constant=15.0 * p["K"], friction=0.0)
    return p


def cylinder(h=1.0e-3, r_in=40, r_out=50, length=200, seed=1234, sc=1.5e-3, dt=5.0e-8, nsteps=500):
    rng = np.random.default_rng(seed)
    m = r_out + 1
    ii = np.arange(-m, m)
    zz = np.arange(0, length)
    X, Y, Z = np.meshgrid(ii, ii, zz, indexing="ij")
    cx = (X.ravel() + 0.5) * h
    cy = (Y.ravel() + 0.5) * h
    cz = (Z.ravel() + 0.5) * h
    r = np.hypot(cx, cy)
    keep = (r >= r_in * h) & (r <= r_out * h)
    x = np.stack([cx[keep], cy[keep], cz[keep]], axis=1)
    x = x + rng.uniform(-0.05 * h, 0.05 * h, size=x.shape)
    vol = np.full(x.shape[0], h ** 3)
    rr = np.hypot(x[:, 0], x[:, 1])
    speed = 200.0 * (1.0 + rng.uniform(-0.01, 0.01, size=x.shape[0]))
    v0 = np.zeros_like(x)
    v0[:, 0] = speed * x[:, 0] / rr
    v0[:, 1] = speed * x[:, 1] / rr
    return dict(name="cylinder", n=None, h=h, x=x, volume=vol, horizon=3.015 * h,
                critical_stretch=sc, dt=dt, nsteps=nsteps, v0=v0, bcs=[], search="radius",
                planes=None, seed=seed, **STEEL)


def qs_bar(nx=60, ny=20, nz=20, h=1.0e-3, strain=1.0e-3, load_steps=10):
    n = (nx, ny, nz)
    L = [nx * h, ny * h, nz * h]
    x, vol = lattice_centres(n, [0.0, 0.0, 0.0], L)
    i = np.rint(x[:, 0] / h - 0.5).astype(int)
    left = np.nonzero(i <= 2)[0]
    right = np.nonzero(i >= nx - 3)[0]
    return dict(name="qs_bar", n=n, h=h, origin=[0.0, 0.0, 0.0], length=L, x=x, volume=vol,
                horizon=3.015 * h, critical_stretch=0.0, left=left, right=right,
                strain=strain, load_steps=load_steps, search="quickgrid", planes=None, **STEEL)


def qs_bar_200k(**kw):
    p = qs_bar(nx=125, ny=40, nz=40, **kw)
    p["name"] = "qs_bar_200k"
    return p


def batched_plates(nspec=64, seed=2024, h=5.0e-4, nz=4, nodes=(5000, 60000), sc=(3e-4, 1e-3),
                   notch=(10, 60), dt=2.5e-8, nsteps=1000, pull=2.0):
    """Parameter sweep: nspec notched plates of uneven size laid side by side
    (gap of 10h, so no bond crosses specimens) in one node/bond space.

    Per specimen (seeded): nx x ny x nz lattice with N in [5k, 60k], critical
    stretch U[3e-4, 1e-3], notch length U[10h, 60h] from the x = 0 edge on
    the mid plane, top/bottom 3-node layers pulled at +-pull."""
    rng = np.random.default_rng(seed)
    xs, vols, scs, planes, bcs, specs = [], [], [], [], [], []
    off = 0.0
    base = 0
    for s in range(nspec):
        target = int(rng.integers(nodes[0], nodes[1] + 1))
        per_layer = max(1, target // nz)
        aspect = float(rng.uniform(0.5, 2.0))
        ny = max(8, int(round(np.sqrt(per_layer / aspect))))
        nx = max(8, int(round(per_layer / ny)))
        n = (nx, ny, nz)
        L = [nx * h, ny * h, nz * h]
        x, vol = lattice_centres(n, [off, 0.0, 0.0], L)
        j = np.rint(x[:, 1] / h - 0.5).astype(int)
        top = base + np.nonzero(j >= ny - 3)[0]
        bot = base + np.nonzero(j <= 2)[0]
        bcs += [(top, 1, 0.0, pull), (bot, 1, 0.0, -pull)]
        ncells = float(rng.uniform(notch[0], notch[1]))
        ncells = min(ncells, 0.8 * nx)
        pl = plate_notch_plane(h, nx, ny, nz, 0)
        pl[3] = off - h                       # lower-left corner x
        pl[9] = (ncells + 1.03) * h           # bottom length
        planes.append(pl)
        s_c = float(rng.uniform(sc[0], sc[1]))
        xs.append(x); vols.append(vol); scs.append(np.full(len(vol), s_c))
        specs.append(dict(n=n, nodes=len(vol), critical_stretch=s_c, notch_cells=ncells, first=base))
        base += len(vol)
        off += L[0] + 10 * h
    x = np.concatenate(xs); vol = np.concatenate(vols)
    return dict(name="batched", h=h, x=x, volume=vol, horizon=3.015 * h,
                critical_stretch=np.concatenate(scs), dt=dt, nsteps=nsteps,
                v0=np.zeros_like(x), bcs=bcs, search="radius", planes=np.array(planes),
                specimens=specs, seed=seed, **STEEL)


def neighbours(p):
    """Neighbour list of a problem, built on the device."""
    if p["search"] == "quickgrid":
        x, vol, nl = quickgrid(p["n"], p["origin"], p["length"], p["horizon"], spherical=True)
        assert np.array_equal(x, p["x"])
        return nl
    return radius_search(p["x"], p["horizon"], p["planes"])
