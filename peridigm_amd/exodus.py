"""Exodus II (netCDF-3) meshes in, Exodus II results out.

Reading follows Peridigm's ExodusDiscretization
(src/io/discretization/Peridigm_ExodusDiscretization.cpp:240-570): every
element becomes one peridynamic point at its centroid with its volume (hex:
24 tetrahedra on the face barycentres and the hex barycentre,
src/io/discretization/Peridigm_GeometryUtils.cpp:69-339; tet: its own
centroid and volume; sphere elements: the node and the volume attribute), and
a node set becomes the set of points whose element contains one of its nodes
(:519-569). Points are numbered block by block in element order.

Writing follows the layout of Peridigm's ExodusII output
(src/io/Peridigm_OutputManager_ExodusII.cpp:210-470): one SPHERE element
per point, nodal variables (vector fields as X/Y/Z components) and element
variables per output frame, written here with scipy's netCDF-3 writer.
"""
from __future__ import annotations

import numpy as np


def _names(var):
    return [b"".join(r).decode("latin-1").strip("\x00 ") for r in var.data]


def read_mesh(path):
    """Returns dict(coords (n,3), blocks [(name, elem_type, conn (ne,k) 0-based)],
    node_sets {name: 0-based node ids in file order})."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        V, D = nc.variables, nc.dimensions
        if "coordx" in V:
            coords = np.stack([V["coordx"].data, V["coordy"].data,
                               V["coordz"].data if "coordz" in V else 0 * V["coordx"].data], 1)
        else:
            c = np.asarray(V["coord"].data, float)
            coords = np.vstack([c, np.zeros((3 - c.shape[0], c.shape[1]))]).T
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        nblk = int(D.get("num_el_blk", 0))
        bnames = _names(V["eb_names"]) if "eb_names" in V else [""] * nblk
        bids = V["eb_prop1"].data if "eb_prop1" in V else np.arange(1, nblk + 1)
        blocks = []
        for b in range(nblk):
            var = V["connect%d" % (b + 1)]
            et = getattr(var, "elem_type", b"HEX8")
            et = et.decode() if isinstance(et, bytes) else str(et)
            attr = V.get("attrib%d" % (b + 1))
            name = bnames[b] or "block_%d" % int(bids[b])
            blocks.append((name, et.upper(), np.asarray(var.data, np.int64) - 1,
                           None if attr is None else np.asarray(attr.data, float)))
        node_sets = {}
        nns = int(D.get("num_node_sets", 0))
        if nns:
            ids = V["ns_prop1"].data
            nnames = _names(V["ns_names"]) if "ns_names" in V else [""] * nns
            for k in range(nns):
                name = nnames[k] or "nodelist_%d" % int(ids[k])
                node_sets[name] = np.asarray(V["node_ns%d" % (k + 1)].data, np.int64) - 1
    return dict(coords=coords, blocks=blocks, node_sets=node_sets)


def _tet_volume(p0, p1, p2, p3):
    # |a . (b x c)| / 6 with a, b, c the edges from the first vertex
    a, b, c = p1 - p0, p2 - p0, p3 - p0
    return (a[..., 0] * (b[..., 1] * c[..., 2] - b[..., 2] * c[..., 1]) +
            a[..., 1] * (b[..., 2] * c[..., 0] - b[..., 0] * c[..., 2]) +
            a[..., 2] * (b[..., 0] * c[..., 1] - b[..., 1] * c[..., 0])) / 6.0


def _tet_centroid(p0, p1, p2, p3):
    return (p0 + p1 + p2 + p3) * 0.25


# the four tetrahedra on each hex face (face barycentre f, hex barycentre c):
# (first corner, second corner, face) in the face's own cyclic order
_HEX_TETS = [(1, 0, 0), (5, 1, 0), (4, 5, 0), (0, 4, 0),
             (2, 1, 1), (6, 2, 1), (5, 6, 1), (1, 5, 1),
             (3, 2, 2), (7, 3, 2), (6, 7, 2), (2, 6, 2),
             (0, 3, 3), (3, 7, 3), (7, 4, 3), (4, 0, 3),
             (0, 1, 4), (1, 2, 4), (2, 3, 4), (3, 0, 4),
             (5, 4, 5), (4, 7, 5), (7, 6, 5), (6, 5, 5)]
_HEX_FACES = [(0, 1, 4, 5), (1, 2, 5, 6), (2, 3, 6, 7), (0, 3, 4, 7), (0, 1, 2, 3), (4, 5, 6, 7)]


def hex_centroid_volume(nodes):
    """nodes (ne, 8, 3) -> centroids (ne, 3), volumes (ne): the hex split
    into 24 tetrahedra, volume-weighted tet centroids."""
    n = np.asarray(nodes, float)
    f = [0.25 * (n[:, a] + n[:, b] + n[:, c] + n[:, d]) for a, b, c, d in _HEX_FACES]
    cen = 0.5 * (f[4] + f[5])
    vol = np.zeros(n.shape[0])
    acc = np.zeros((n.shape[0], 3))
    for a, b, k in _HEX_TETS:
        tv = _tet_volume(n[:, a], n[:, b], f[k], cen)
        vol = vol + tv
        acc = acc + _tet_centroid(n[:, a], n[:, b], f[k], cen) * tv[:, None]
    if np.any(vol <= 0.0):
        raise ValueError("hexahedron with non-positive volume (inverted or misnumbered)")
    return acc / vol[:, None], vol


def discretize(mesh):
    """Peridynamic points of an Exodus mesh: (x (n,3), volume (n), block index
    (n), point sets {node-set name: sorted-by-insertion point ids})."""
    xs, vs, bs, owner = [], [], [], []
    base = 0
    node_elems = [[] for _ in range(mesh["coords"].shape[0])]
    for bi, (name, et, conn, attr) in enumerate(mesh["blocks"]):
        pts = mesh["coords"][conn]
        if et.startswith("HEX"):
            c, v = hex_centroid_volume(pts[:, :8])
        elif et.startswith("TET"):
            p = pts[:, :4]
            v = _tet_volume(p[:, 0], p[:, 1], p[:, 2], p[:, 3])
            if np.any(v < 0.0):
                raise ValueError("tetrahedron with negative volume")
            c = _tet_centroid(p[:, 0], p[:, 1], p[:, 2], p[:, 3])
        elif et.startswith("SPHERE"):
            c = pts[:, 0]
            v = attr[:, 1] if attr is not None and attr.shape[1] > 1 else attr[:, 0]
        else:
            raise ValueError("unsupported element type %s" % et)
        xs.append(c); vs.append(v); bs.append(np.full(len(v), bi))
        for e in range(conn.shape[0]):
            for nd in conn[e]:
                node_elems[nd].append(base + e)
        base += conn.shape[0]
    sets = {}
    for name, nodes in mesh["node_sets"].items():
        seen, out = set(), []
        for nd in nodes:
            for e in node_elems[nd]:
                if e not in seen:
                    seen.add(e)
                    out.append(e)
        sets[name] = np.asarray(out, np.int64)
    return (np.ascontiguousarray(np.concatenate(xs)), np.concatenate(vs), np.concatenate(bs), sets)


def write_results(path, x, times, nodal, element, title="peridigm_amd"):
    """Exodus II results: points x (n,3) as SPHERE elements of one block;
    nodal / element: {name: array (frames, n)}; vector fields are passed as
    their X, Y, Z component names."""
    from scipy.io import netcdf_file
    n = x.shape[0]
    nf = len(times)
    with netcdf_file(path, "w", version=2) as nc:
        nc.title = title.encode()
        nc.api_version = np.float32(5.22)
        nc.version = np.float32(5.22)
        nc.floating_point_word_size = np.int32(8)
        nc.file_size = np.int32(1)
        nc.createDimension("time_step", None)  # the record dimension comes first
        nc.createDimension("len_string", 33)
        nc.createDimension("len_line", 81)
        nc.createDimension("four", 4)
        nc.createDimension("len_name", 33)
        nc.createDimension("num_dim", 3)
        nc.createDimension("num_nodes", n)
        nc.createDimension("num_elem", n)
        nc.createDimension("num_el_blk", 1)
        nc.createDimension("num_el_in_blk1", n)
        nc.createDimension("num_nod_per_el1", 1)

        def names_var(vname, dim, names):
            v = nc.createVariable(vname, "S1", (dim, "len_name"))
            arr = np.zeros((len(names), 33), "S1")
            for i, s in enumerate(names):
                b = s.encode()[:32]
                arr[i, :len(b)] = np.frombuffer(b, "S1")
            v[:] = arr

        t = nc.createVariable("time_whole", "d", ("time_step",))
        t[:nf] = np.asarray(times, float)
        for i, c in enumerate("xyz"):
            v = nc.createVariable("coord" + c, "d", ("num_nodes",))
            v[:] = x[:, i]
        names_var("coor_names", "num_dim", ["X", "Y", "Z"])
        v = nc.createVariable("eb_status", "i", ("num_el_blk",)); v[:] = [1]
        v = nc.createVariable("eb_prop1", "i", ("num_el_blk",)); v[:] = [1]
        v.name = b"ID"
        names_var("eb_names", "num_el_blk", [""])
        conn = nc.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        conn.elem_type = b"SPHERE"
        conn[:] = np.arange(1, n + 1)[:, None]
        v = nc.createVariable("node_num_map", "i", ("num_nodes",)); v[:] = np.arange(1, n + 1)
        v = nc.createVariable("elem_num_map", "i", ("num_elem",)); v[:] = np.arange(1, n + 1)
        if nodal:
            nc.createDimension("num_nod_var", len(nodal))
            names_var("name_nod_var", "num_nod_var", list(nodal))
            for k, (name, val) in enumerate(nodal.items()):
                v = nc.createVariable("vals_nod_var%d" % (k + 1), "d", ("time_step", "num_nodes"))
                v[:nf] = np.asarray(val, float)
        if element:
            nc.createDimension("num_elem_var", len(element))
            names_var("name_elem_var", "num_elem_var", list(element))
            tab = nc.createVariable("elem_var_tab", "i", ("num_el_blk", "num_elem_var"))
            tab[:] = np.ones((1, len(element)), np.int32)
            for k, (name, val) in enumerate(element.items()):
                v = nc.createVariable("vals_elem_var%deb1" % (k + 1), "d",
                                      ("time_step", "num_el_in_blk1"))
                v[:nf] = np.asarray(val, float)


def read_results(path):
    """(times, x, {nodal and element variable name: (frames, n)}) of an
    Exodus II results file (one block)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        V = nc.variables
        times = np.array(V["time_whole"].data, float)
        x = np.stack([V["coordx"].data, V["coordy"].data, V["coordz"].data], 1).astype(float)
        out = {}
        if "name_nod_var" in V:
            for k, nm in enumerate(_names(V["name_nod_var"])):
                out[nm] = np.array(V["vals_nod_var%d" % (k + 1)].data, float)
        if "name_elem_var" in V:
            for k, nm in enumerate(_names(V["name_elem_var"])):
                out[nm] = np.array(V["vals_elem_var%deb1" % (k + 1)].data, float)
    return times, x, out
