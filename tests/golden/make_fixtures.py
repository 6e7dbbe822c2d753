"""Writes the golden fixtures under tests/golden/ from the reference's own tests.

Run once in a container that has the reference checkout:

    python tests/golden/make_fixtures.py /root/reference

Sources (paths relative to the reference root):
  * src/materials/unit_test/utPeridigm_ElasticMaterial.cpp  testTwoPts (:78-186),
    testEightPts (:189-452), testThreePts (:456-659): inputs and the asserted
    weighted volume, dilatation and force values with their tolerances.
  * src/io/pdneigh/unit_test/ut_neighborhood_list.cxx:68-152 (1x1x11 bar,
    horizon 3.1 cells, self included, expected sets and sizes).
  * src/io/discretization/unit_test/utPeridigm_PdQuickGridDiscretization.cpp:70-201
    (2x2x2 lattice, horizon 0.501, spherical, 3 neighbours, centres, volume).
  * src/io/pdneigh/unit_test/utFinitePlane.cxx:52-104 (plane/bond intersection).
  * test/verification/ElasticBondBased/ElasticBondBased3D/ElasticBondBased3D_gold.e
    and test/verification/BondBreakingInitialVelocity_TimeDependentCS/
    BondBreakingInitialVelocity_gold.e: their output frames (element variables
    of all blocks concatenated), decks and meshes copied to decks/.
  * test/regression/PrecrackedPlate/PrecrackedPlate_gold.e: the two output
    frames of the quasi-static run, with the .comp tolerances; the WaveInBar
    and PrecrackedPlate decks and meshes are copied to decks/ unchanged.
  * test/verification/BondBreakingInitialVelocity/BondBreakingInitialVelocity.xml
    and _gold.e (netCDF-3 read with scipy): the deck parameters and the nine
    output frames of displacement, velocity, dilatation, weighted volume and
    damage, with the .comp tolerances.
Only data (inputs and expected outputs) is written; nothing of the reference's
source text.
"""
import json
import os
import sys

import numpy as np


def elastic_unit_tests():
    out = {}
    out["params"] = dict(density=7800.0, K=130.0e9, mu=78.0e9, horizon=10.0)
    # testTwoPts
    out["two"] = dict(
        x=[[0, 0, 0], [1, 0, 0]], y=[[0, 0, 0], [2, 0, 0]], vol=[1.0, 1.0],
        nlist=[1, 1, 1, 0],
        m=[1.0, 1.0], m_tol=1e-15, theta=[3.0, 3.0], theta_tol=1e-15,
        fx=[2.34e12, -2.34e12], f_tol=1e-2)
    # testEightPts
    x8 = [[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1], [1, 0, 0], [1, 0, 1], [1, 1, 0], [1, 1, 1]]
    y8 = [[a, b, 0.98 if c == 1 else 0.0] for a, b, c in x8]
    nl8 = []
    for i in range(8):
        nl8.append(7)
        nl8.extend(j for j in range(8) if j != i)
    # hand solution (pairwise t = (|Y|-|X|)*9.75e10, factor 2 for equal and
    # opposite pair forces), force on cell 0; the others differ by sign.
    f = np.zeros(3)
    def add(t, uv):
        nonlocal f
        f += 2.0 * np.asarray(uv) * t
    t = (0.98 - 1.0) * 9.75e10; add(t, [0, 0, 0.98 / 0.98])
    vm = np.sqrt(1.0 + 0.98 * 0.98); t = (vm - np.sqrt(2.0)) * 9.75e10; add(t, [0, 1 / vm, 0.98 / vm])
    add(t, [1 / vm, 0, 0.98 / vm])
    vm = np.sqrt(2.0 + 0.98 * 0.98); t = (vm - np.sqrt(3.0)) * 9.75e10; add(t, [1 / vm, 1 / vm, 0.98 / vm])
    signs = [[1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1], [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]]
    out["eight"] = dict(x=x8, y=y8, vol=[1.0] * 8, nlist=nl8, m=12.0, m_tol=1e-12,
                        theta=-0.01991593994643333, theta_tol=1e-12,
                        f0=f.tolist(), signs=signs, f_tol=1e-11)
    # testThreePts
    vol3 = [0.9, 1.1, 0.8]
    nl3 = [2, 1, 2, 2, 0, 2, 2, 0, 1]
    r = {}
    r[0] = np.array([-2753550531.144094, -1510011581.595148, -621769474.7744727]) * vol3[1] \
        - np.array([3398655528.6806289, 1863778838.308732, 767438345.18594845]) * vol3[1] \
        + np.array([7736514797.0571524, 15402697641.41379, -11956431959.088315]) * vol3[2] \
        - np.array([-8687228655.850972, -17295482505.73966, 13425717013.587864]) * vol3[2]
    r[1] = np.array([3398655528.6806289, 1863778838.308732, 767438345.18594845]) * vol3[0] \
        - np.array([-2753550531.144094, -1510011581.595148, -621769474.7744727]) * vol3[0] \
        + np.array([5110873051.7924152, -1252163897.689138, 6133047662.15088]) * vol3[2] \
        - np.array([-4649613866.523322, 1139155397.2982139, -5579536639.827986]) * vol3[2]
    r[2] = np.array([-8687228655.850972, -17295482505.73966, 13425717013.587864]) * vol3[0] \
        - np.array([7736514797.0571524, 15402697641.41379, -11956431959.088315]) * vol3[0] \
        + np.array([-4649613866.523322, 1139155397.2982139, -5579536639.827986]) * vol3[1] \
        - np.array([5110873051.7924152, -1252163897.689138, 6133047662.15088]) * vol3[1]
    out["three"] = dict(
        x=[[1.1, 2.6, -0.1], [-2.0, 0.9, -0.3], [0.0, 0.01, 1.8]],
        y=[[1.2, 2.4, -0.1], [-1.9, 0.7, -0.8], [0.1, 0.21, 1.6]], vol=vol3, nlist=nl3,
        m=[23.016479999999931, 18.64767999999997, 20.497599999999963], m_tol=1e-12,
        theta=[-0.114127034572639, 0.08257537372985, -0.12166177890553], theta_tol=1e-11,
        f=[r[0].tolist(), r[1].tolist(), r[2].tolist()],
        f_tol=[[1e-11, 1e-11, 1e-11], [1e-11, 1e-10, 1e-11], [1e-11, 1e-11, 1e-11]])
    return out


def neighborhood_list_bar():
    sets = [[0, 1, 2, 3], [0, 1, 2, 3, 4], [0, 1, 2, 3, 4, 5], [0, 1, 2, 3, 4, 5, 6],
            [1, 2, 3, 4, 5, 6, 7], [2, 3, 4, 5, 6, 7, 8], [3, 4, 5, 6, 7, 8, 9],
            [4, 5, 6, 7, 8, 9, 10], [5, 6, 7, 8, 9, 10], [6, 7, 8, 9, 10], [7, 8, 9, 10]]
    return dict(nx=1, ny=1, nz=11, origin=[0.0, 0.0, 0.0], length=[1.0, 1.0, 1.0],
                scale=3.1, include_self=True, sizes=[4, 5, 6, 7, 7, 7, 7, 7, 6, 5, 4],
                total=65, sets=sets)


def pdquickgrid_2x2x2():
    return dict(n=[2, 2, 2], origin=[0.0, 0.0, 0.0], length=[1.0, 1.0, 1.0], horizon=0.501,
                num_bonds=24, neighbors_each=3,
                x=[[0.25, 0.25, 0.25], [0.75, 0.25, 0.25], [0.25, 0.75, 0.25], [0.75, 0.75, 0.25],
                   [0.25, 0.25, 0.75], [0.75, 0.25, 0.75], [0.25, 0.75, 0.75], [0.75, 0.75, 0.75]],
                x_tol=1e-16, volume=0.125)


def finite_plane():
    # normal, lower-left corner, bottom unit vector, bottom length, side length
    plane = [1, 0, 0, 0, 1, 0, 0, -1, 0, 1.0, 1.0]
    return dict(plane=plane,
                cases=[dict(p0=[-.5, .5, .5], p1=[.5, .5, .5], cuts=True, t=0.5, x=[0.0, 0.5, 0.5]),
                       dict(p0=[-.5, .5, .5], p1=[-.5, 1.5, .5], cuts=False)])


def bond_breaking(ref):
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/BondBreakingInitialVelocity")
    f = netcdf_file(os.path.join(d, "BondBreakingInitialVelocity_gold.e"), "r", mmap=False)

    def names(k):
        return [b"".join(r).strip(b"\x00").decode() for r in f.variables[k][:]]

    nod = names("name_nod_var")
    ele = names("name_elem_var")
    out = dict(
        deck=dict(n=[2, 2, 1], origin=[0.0, 0.0, 0.0], length=[2.1, 2.0, 1.0], horizon=0.51,
                  spherical=False, density=7800.0, K=130.0e9, mu=60.0e9, critical_stretch=0.02,
                  dt=1.0e-8, final_time=0.00008, output_every=1000,
                  initial_velocity_y=[-244.939, -244.959, 244.939, 244.959]),
        tolerances=dict(displacement=1e-9, velocity=1e-9, weighted_volume=1e-12,
                        dilatation=1e-12, damage=1e-3),
        time=np.asarray(f.variables["time_whole"][:]).tolist(),
        coordx=np.asarray(f.variables["coordx"][:]).tolist(),
        coordy=np.asarray(f.variables["coordy"][:]).tolist(),
        coordz=np.asarray(f.variables["coordz"][:]).tolist())
    for i, nme in enumerate(nod):
        out[nme] = np.asarray(f.variables["vals_nod_var%d" % (i + 1)][:]).tolist()
    for i, nme in enumerate(ele):
        out[nme] = np.asarray(f.variables["vals_elem_var%deb1" % (i + 1)][:]).tolist()
    f.close()
    return out


def _exodus_frames(f, frames):
    """Nodal and element variables of an Exodus (netCDF-3) output file at the
    given frame indices, keyed by variable name."""
    def names(k):
        return [b"".join(r).strip(b"\x00").decode() for r in f.variables[k][:]]
    out = {}
    for i, nme in enumerate(names("name_nod_var")):
        out[nme] = np.asarray(f.variables["vals_nod_var%d" % (i + 1)][:])[frames]
    for i, nme in enumerate(names("name_elem_var")):
        out[nme] = np.asarray(f.variables["vals_elem_var%deb1" % (i + 1)][:])[frames]
    return out


def compression_qs_explicit(ref):
    """Compression_QS_Explicit_3x2x2: 20 quasi-static load steps (frames
    0..20), then 500 Verlet steps of dt 1e-7 from t = 5e-5 with the same
    prescribed displacements (frame 21 is the explicit solver's initial
    output, frame 21+k the state after step k)."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/regression/Compression_QS_Explicit_3x2x2")
    f = netcdf_file(os.path.join(d, "Compression_QS_Explicit_3x2x2_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    frames = list(range(0, 24)) + [26, 31, 71, 121, 221, 321, 421, 521]
    vals = _exodus_frames(f, frames)
    out = dict(
        deck=dict(n=[3, 2, 2], origin=[-1.5, -1.0, -1.0], length=[3.0, 2.0, 2.0], horizon=1.75,
                  spherical=True, density=7800.0, K=130.0e9, mu=78.0e9,
                  qs=dict(t0=0.0, t1=5.0e-5, load_steps=20, absolute_tolerance=1.0e-2,
                          max_iterations=10),
                  explicit=dict(t0=5.0e-5, t1=1.0e-4, dt=1.0e-7),
                  # "Prescribed Displacement" node sets (1-based ids in the deck:
                  # "1 4 7 10", "3 6 9 12", "1 4", "1 7"), coordinate, value(t)
                  bcs=[dict(nodes=[0, 3, 6, 9], coord=0, c0=0.0, c1=0.0),
                       dict(nodes=[2, 5, 8, 11], coord=0, c0=0.0, c1=-0.1 / 0.00005),
                       dict(nodes=[0, 3], coord=2, c0=0.0, c1=0.0),
                       dict(nodes=[0, 6], coord=1, c0=0.0, c1=0.0)]),
        tolerances=dict(displacement=1e-9, velocity=5e-8, force_density=1.0,
                        weighted_volume=1e-12, dilatation=1e-12, coordinates=1e-12),
        frames=frames, time=time[frames].tolist(),
        coordx=np.asarray(f.variables["coordx"][:]).tolist(),
        coordy=np.asarray(f.variables["coordy"][:]).tolist(),
        coordz=np.asarray(f.variables["coordz"][:]).tolist())
    for k, v in vals.items():
        out[k] = v.tolist()
    f.close()
    return out


def wave_in_bar(ref):
    """WaveInBar: Exodus hex mesh (525 HEX8 elements) turned into one point
    per element, explicit LPS with an initial velocity of -100 m/s on the
    elements that touch node set 1; 12 output frames (every 20 steps)."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/regression/WaveInBar")
    g = netcdf_file(os.path.join(d, "WaveInBar.g"), "r", mmap=False)
    X = np.stack([np.asarray(g.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1)
    conn = np.asarray(g.variables["connect1"][:]) - 1
    ns = np.asarray(g.variables["node_ns1"][:]) - 1
    hexes = X[conn]                                   # (525, 8, 3)
    # the mesh is made of axis-aligned boxes: volume = product of the extents
    lo, hi = hexes.min(1), hexes.max(1)
    assert all(np.allclose(np.sort(np.unique(np.round(h[:, c], 12))), [lo[e, c], hi[e, c]])
               for e, h in enumerate(hexes) for c in range(3))
    vol = np.prod(hi - lo, axis=1)
    in_set = np.isin(conn, ns).any(axis=1)
    g.close()
    f = netcdf_file(os.path.join(d, "WaveInBar_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    frames = list(range(len(time)))
    vals = _exodus_frames(f, frames)
    out = dict(
        hex_nodes=hexes, volume=vol, initial_velocity_elements=np.nonzero(in_set)[0],
        coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
        time=time, density=np.array(2200.0), K=np.array(14.90e9), mu=np.array(8.94e9),
        horizon=np.array(0.00601), dt=np.array(9.0659e-08), output_every=np.array(20),
        initial_velocity_x=np.array(-100.0),
        tol_displacement=np.array(5e-14), tol_velocity=np.array(1e-8),
        tol_force_density=np.array(5.0), tol_weighted_volume=np.array(1e-15),
        tol_dilatation=np.array(5e-13), tol_coordinates=np.array(1e-12))
    for k, v in vals.items():
        out[k] = v
    f.close()
    return out


def precracked_plate(ref):
    """PrecrackedPlate: Exodus hex mesh (640 elements) with a rectangular-plane
    bond filter, one quasi-static load step to 1% strain (absolute tolerance
    1.0); the gold frames at t = 0 and t = 1 with the .comp tolerances."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/regression/PrecrackedPlate")
    f = netcdf_file(os.path.join(d, "PrecrackedPlate_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-12), tol_force=np.array(1.0),
               tol_number_of_neighbors=np.array(1e-15), tol_dilatation=np.array(1e-13),
               tol_default=np.array(1e-9))
    for k, v in _exodus_frames(f, list(range(len(time)))).items():
        out[k] = v
    f.close()
    return out


def _exodus_all_blocks(f, frames):
    """Like _exodus_frames, with element variables concatenated over blocks."""
    def names(k):
        return [b"".join(r).strip(b"\x00").decode() for r in f.variables[k][:]]
    out = {}
    for i, nme in enumerate(names("name_nod_var")):
        out[nme] = np.asarray(f.variables["vals_nod_var%d" % (i + 1)][:])[frames]
    nblk = f.dimensions["num_el_blk"]
    if "name_elem_var" in f.variables:
        for i, nme in enumerate(names("name_elem_var")):
            out[nme] = np.concatenate(
                [np.asarray(f.variables["vals_elem_var%deb%d" % (i + 1, b + 1)][:])[frames]
                 for b in range(nblk)], axis=1)
    return out


def elastic_bond_based_3d(ref):
    """ElasticBondBased3D (verification): 384 hex points in three blocks of
    the elastic bond-based material, every point moved by the prescribed
    displacement u_y = 0.001 y t / 1e-8; five Verlet steps of 2e-9, six frames."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/ElasticBondBased/ElasticBondBased3D")
    f = netcdf_file(os.path.join(d, "ElasticBondBased3D_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-9), tol_velocity=np.array(1e-9), tol_force=np.array(1e-3))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def bond_breaking_time_dependent(ref):
    """BondBreakingInitialVelocity_TimeDependentCS: the four-point deck with a
    critical stretch of 50 until t = 7e-5 and 0.015 after; 11 frames."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/BondBreakingInitialVelocity_TimeDependentCS")
    f = netcdf_file(os.path.join(d, "BondBreakingInitialVelocity_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-9), tol_velocity=np.array(1e-9), tol_damage=np.array(1e-3),
               tol_weighted_volume=np.array(1e-12), tol_dilatation=np.array(1e-12))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def contact_2x1x1(ref):
    """Contact_2x1x1 (verification): two unbonded points 2 mm apart at the
    contact radius closing at 1 mm/ms each; short-range contact force only;
    1000 Verlet steps, 11 frames."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/Contact_2x1x1")
    f = netcdf_file(os.path.join(d, "Contact_2x1x1_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-13), tol_velocity=np.array(1e-13))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def contact_friction(ref):
    """Contact_Friction (verification): two blocks of an elastic material
    without bonds between them, driven together at 10 m/s; short-range
    contact with friction coefficient 0.3."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/Contact_Friction")
    f = netcdf_file(os.path.join(d, "Contact_Friction_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-13), tol_velocity=np.array(1e-13),
               tol_contact_force=np.array(5e-6))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def linear_lps_bar(ref):
    """LinearLPSBar (verification): two disjoint hex bars (blocks 1-3 and 4-6,
    500 points each) of "Linear LPS Partial Volume" and "Elastic Partial
    Volume" without partial volumes, every point moved by the prescribed
    displacement u = (0.1 x t, 0, 0); quasi-static, five load steps, six
    frames, the .comp tolerances."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/LinearLPSBar")
    f = netcdf_file(os.path.join(d, "LinearLPSBar_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-9), tol_force=np.array(1e-3),
               tol_weighted_volume=np.array(1e-14), tol_dilatation=np.array(1e-10),
               tol_volume=np.array(1e-12))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def thermal_expansion_cube(ref):
    """ThermalExpansionCube (verification): 5x5x5 PdQuickGrid cube, elastic
    LPS with thermal expansion coefficient 1e-5, the whole domain's
    temperature prescribed as 273 + 1000 t / 3600, rigid motions held by
    zero displacements on three axis node sets; quasi-static, ten load
    steps, eleven frames, the .comp tolerances."""
    from scipy.io import netcdf_file
    d = os.path.join(ref, "test/verification/ThermalExpansionCube")
    f = netcdf_file(os.path.join(d, "ThermalExpansionCube_gold.e"), "r", mmap=False)
    time = np.asarray(f.variables["time_whole"][:])
    out = dict(time=time,
               coords=np.stack([np.asarray(f.variables[k][:]) for k in ("coordx", "coordy", "coordz")], 1),
               tol_displacement=np.array(1e-9), tol_velocity=np.array(1e-9), tol_force=np.array(1e-3),
               tol_temperature=np.array(1e-9), tol_weighted_volume=np.array(1e-12),
               tol_dilatation=np.array(1e-12))
    out.update(_exodus_all_blocks(f, list(range(len(time)))))
    f.close()
    return out


def copy_decks(ref, here):
    """The input decks and meshes of the WaveInBar and PrecrackedPlate
    regression tests (data files, copied unchanged)."""
    import shutil
    dst = os.path.join(here, "decks")
    os.makedirs(dst, exist_ok=True)
    for rel, name in (("test/regression/WaveInBar/WaveInBar.xml", None),
                      ("test/regression/WaveInBar/WaveInBar.g", None),
                      ("test/regression/PrecrackedPlate/PrecrackedPlate.xml", None),
                      ("test/regression/PrecrackedPlate/Plate.g", None),
                      ("test/verification/ElasticBondBased/ElasticBondBased3D/ElasticBondBased3D.yaml", None),
                      ("test/verification/ElasticBondBased/ElasticBondBased3D/ElasticBondBased3D.g", None),
                      ("test/verification/BondBreakingInitialVelocity_TimeDependentCS/"
                       "BondBreakingInitialVelocity.xml", "BondBreakingInitialVelocity_TimeDependentCS.xml"),
                      ("test/verification/Contact_2x1x1/Contact_2x1x1.xml", None),
                      ("test/verification/Contact_Friction/Contact_Friction.xml", None),
                      ("test/verification/Contact_Friction/Contact_Friction.g", None),
                      ("test/verification/LinearLPSBar/LinearLPSBar.yaml", None),
                      ("test/verification/LinearLPSBar/LinearLPSBar.g", None),
                      ("test/verification/ThermalExpansionCube/ThermalExpansionCube.xml", None)):
        shutil.copyfile(os.path.join(ref, rel), os.path.join(dst, name or os.path.basename(rel)))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "wave_in_bar.npz"), **wave_in_bar(ref))
    print("wrote wave_in_bar.npz")
    np.savez_compressed(os.path.join(here, "precracked_plate.npz"), **precracked_plate(ref))
    print("wrote precracked_plate.npz")
    np.savez_compressed(os.path.join(here, "elastic_bond_based_3d.npz"), **elastic_bond_based_3d(ref))
    np.savez_compressed(os.path.join(here, "bond_breaking_time_dependent.npz"),
                        **bond_breaking_time_dependent(ref))
    print("wrote elastic_bond_based_3d.npz, bond_breaking_time_dependent.npz")
    np.savez_compressed(os.path.join(here, "contact_2x1x1.npz"), **contact_2x1x1(ref))
    np.savez_compressed(os.path.join(here, "contact_friction.npz"), **contact_friction(ref))
    print("wrote contact_2x1x1.npz, contact_friction.npz")
    np.savez_compressed(os.path.join(here, "linear_lps_bar.npz"), **linear_lps_bar(ref))
    print("wrote linear_lps_bar.npz")
    np.savez_compressed(os.path.join(here, "thermal_expansion_cube.npz"), **thermal_expansion_cube(ref))
    print("wrote thermal_expansion_cube.npz")
    copy_decks(ref, here)
    print("copied decks")
    fx = {
        "elastic_unit_tests.json": elastic_unit_tests(),
        "neighborhood_list_bar.json": neighborhood_list_bar(),
        "pdquickgrid_2x2x2.json": pdquickgrid_2x2x2(),
        "finite_plane.json": finite_plane(),
        "bond_breaking_initial_velocity.json": bond_breaking(ref),
        "compression_qs_explicit_3x2x2.json": compression_qs_explicit(ref),
    }
    for name, obj in fx.items():
        with open(os.path.join(here, name), "w") as fh:
            json.dump(obj, fh, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
