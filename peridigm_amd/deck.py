"""Input decks: Peridigm's XML parameter lists run on the device.

`run_deck(path)` reads a deck in Peridigm's Teuchos XML format and runs it
through the C ABI: the discretization (Exodus mesh with bond filters, or a
PdQuickGrid tensor-product lattice), one block of the "Elastic" (LPS)
material with an optional "Critical Stretch" damage model, the initial and
boundary conditions, the "Verlet" or "QuasiStatic" solver, and the ExodusII
output frames (src/core/Peridigm.cpp:1290-1593 and :2470-3043 for the
solvers, src/io/Peridigm_OutputManager_ExodusII.cpp:210-470 for the output).

Supported subset (anything else raises DeckError):
  - materials "Elastic" and "Elastic Bond Based" (3D, plane strain, plane
    stress); damage models "Critical Stretch" and "Time Dependent Critical
    Stretch";
  - boundary-condition and damage values are C-like statements of x, y, z
    and t (`value = expr;`, if / else if / else with comparisons, && || !;
    expressions with numbers, + - * /, parentheses, and sin cos tan exp log
    sqrt fabs pow);
    prescribed displacements must be linear in t (u = c1(x) t + c0(x)), as the
    device integrator applies them;
  - elastic constants: any two of bulk, shear, Young's modulus and Poisson's
    ratio (src/materials/Peridigm_Material.cpp:335-433);
  - block groups that together hold all the mesh's blocks and share one
    horizon; groups of different materials may share no bond (each runs in
    a context of its own); the influence function ("One", "Parabolic Decay"
    or "Gaussian") from the Discretization list; "Omit Bonds Between Blocks"
    "All" or "None" (src/io/discretization/Peridigm_ExodusDiscretization.cpp:709-760);
  - short-range contact ("Contact": Search Radius, Search Frequency and one
    "Short Range Force" model, every node taking part).
"""
from __future__ import annotations

import ast
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import exodus as EX
from ._lib import INFLUENCE
from .core import Context
from .discretization import quickgrid, radius_search


class DeckError(ValueError):
    pass


# ------------------------------------------------------------------ parsing
def _convert(typ, val):
    typ = (typ or "string").lower()
    if typ in ("double", "float"):
        return float(val)
    if typ == "int":
        return int(val)
    if typ == "bool":
        return val.strip().lower() in ("true", "1", "yes")
    return val


def _plist(elem):
    out = {}
    for ch in elem:
        if ch.tag == "ParameterList":
            out[ch.get("name")] = _plist(ch)
        elif ch.tag == "Parameter":
            out[ch.get("name")] = _convert(ch.get("type"), ch.get("value"))
    return out


def _numbers(d):
    """YAML 1.1 reads 2.0e9 (no exponent sign) as a string: numbers back."""
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            v = _numbers(v)
        elif isinstance(v, str):
            try:
                v = float(v) if any(ch in v for ch in ".eE") else int(v)
            except ValueError:
                pass
        out[k] = v
    return out


def parse_deck(path):
    """The deck as nested dicts: an XML ParameterList (ParameterList -> dict,
    Parameter -> typed value) or a YAML deck (its "Peridigm" mapping)."""
    if path.endswith((".yaml", ".yml")):
        import yaml
        with open(path) as fh:
            d = yaml.safe_load(fh)
        return _numbers(d.get("Peridigm", d))
    return _plist(ET.parse(path).getroot())


# -------------------------------------------------------------- expressions
_FUNCS = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "log": np.log,
          "sqrt": np.sqrt, "fabs": np.abs, "abs": np.abs, "pow": np.power}
_BIN = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.divide}


def evaluate(expr, x, y, z, t):
    """Value of a boundary-condition expression at points (x, y, z) and time t.
    Parsed with Python's ast and evaluated node by node: only numbers, the
    variables x y z t, the four operators, unary minus and the listed
    functions are accepted."""
    s = expr.strip().rstrip(";")
    if "=" in s:
        lhs, s = s.split("=", 1)
        if lhs.strip() != "value":
            raise DeckError("expression must assign 'value': %r" % expr)
    try:
        tree = ast.parse(s.strip(), mode="eval")
    except SyntaxError as e:
        raise DeckError("cannot parse expression %r: %s" % (expr, e))
    env = {"x": x, "y": y, "z": z, "t": t, "pi": math.pi, "M_PI": math.pi}

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name) and n.id in env:
            return env[n.id]
        if isinstance(n, ast.BinOp) and type(n.op) in _BIN:
            return _BIN[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS \
                and not n.keywords:
            return _FUNCS[n.func.id](*[ev(a) for a in n.args])
        raise DeckError("unsupported expression element in %r" % expr)

    return np.broadcast_to(np.asarray(ev(tree), dtype=np.float64), np.shape(x)).copy()


def _tokens(code):
    import re
    tok = re.findall(r"\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|[A-Za-z_]\w*|"
                     r"&&|\|\||<=|>=|==|!=|[-+*/(){};=<>!,]", code)
    return tok


def evaluate_program(code, x, y, z, t):
    """Value of a boundary-condition or damage function written as C-like
    statements: `value = expr;` and `if (cond) { ... } else { ... }` (else if
    chains included), with comparisons, && || ! in the conditions. Each
    expression and condition goes through the ast-based evaluator."""
    s = code.replace("\\{", "{").replace("\\}", "}").replace("\\", "")
    if "if" not in _tokens(s):
        return evaluate(s, x, y, z, t)
    toks = _tokens(s)
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else None

    def take(v=None):
        tk = peek()
        if v is not None and tk != v:
            raise DeckError("expected %r in %r" % (v, code))
        pos[0] += 1
        return tk

    def until(stop):
        depth, out = 0, []
        while True:
            tk = peek()
            if tk is None:
                raise DeckError("unterminated statement in %r" % code)
            if depth == 0 and tk in stop:
                return " ".join(out)
            if tk == "(":
                depth += 1
            elif tk == ")":
                depth -= 1
            out.append(tk)
            take()

    state = {"value": np.zeros(np.shape(x))}

    def cond_value(txt):
        py = txt.replace("&&", " and ").replace("||", " or ").replace("!", " not ").replace(" not =", "!=")
        return _eval_condition(py, x, y, z, t)

    def stmt(active):
        tk = peek()
        if tk == "if":
            take("if"); take("(")
            c = cond_value(until(")"))
            take(")")
            block(active & c)
            if peek() == "else":
                take("else")
                if peek() == "if":
                    stmt(active & ~c)
                else:
                    block(active & ~c)
        elif tk == "value":
            take("value"); take("=")
            v = evaluate(until(";"), x, y, z, t)
            take(";")
            state["value"] = np.where(active, v, state["value"])
        elif tk == ";":
            take(";")
        else:
            raise DeckError("unsupported statement in %r" % code)

    def block(active):
        if peek() == "{":
            take("{")
            while peek() != "}":
                stmt(active)
            take("}")
        else:
            stmt(active)

    active = np.ones(np.shape(x), bool)
    while peek() is not None:
        stmt(active)
    return state["value"]


_CMP = {ast.Lt: np.less, ast.LtE: np.less_equal, ast.Gt: np.greater, ast.GtE: np.greater_equal,
        ast.Eq: np.equal, ast.NotEq: np.not_equal}


def _eval_condition(txt, x, y, z, t):
    try:
        tree = ast.parse(txt.strip(), mode="eval")
    except SyntaxError as e:
        raise DeckError("cannot parse condition %r: %s" % (txt, e))

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.BoolOp):
            vals = [ev(v) for v in n.values]
            out = vals[0]
            for v in vals[1:]:
                out = (out & v) if isinstance(n.op, ast.And) else (out | v)
            return out
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.Not):
            return ~ev(n.operand)
        if isinstance(n, ast.Compare) and len(n.ops) == 1 and type(n.ops[0]) in _CMP:
            a = evaluate(ast.unparse(n.left), x, y, z, t)
            b = evaluate(ast.unparse(n.comparators[0]), x, y, z, t)
            return _CMP[type(n.ops[0])](a, b)
        raise DeckError("unsupported condition %r" % txt)

    return np.broadcast_to(ev(tree), np.shape(x)).copy()


def linear_in_t(expr, x, y, z):
    """(c0, c1) with value = c1 t + c0 at every point, checked at t = 0.5, 2."""
    c0 = evaluate_program(expr, x, y, z, 0.0)
    c1 = evaluate_program(expr, x, y, z, 1.0) - c0
    for t in (0.5, 2.0):
        v = evaluate_program(expr, x, y, z, t)
        if not np.allclose(v, c0 + c1 * t, rtol=1e-12, atol=1e-300 + 1e-14 * np.abs(c0 + c1 * t).max()):
            raise DeckError("prescribed displacement %r is not linear in t" % expr)
    return c0, c1


# ---------------------------------------------------------------- material
def elastic_constants(mat):
    """(K, mu) from exactly two of bulk modulus, shear modulus, Young's
    modulus and Poisson's ratio, by the reference's formulas
    (src/materials/Peridigm_Material.cpp:335-433)."""
    K = mat.get("Bulk Modulus")
    G = mat.get("Shear Modulus")
    E = mat.get("Young's Modulus")
    nu = mat.get("Poisson's Ratio")
    if sum(v is not None for v in (K, G, E, nu)) != 2:
        raise DeckError("exactly two elastic constants must be given")
    if K is not None:
        k = K
    elif E is not None and G is not None:
        k = (E * G) / (3.0 * (3.0 * G - E))
    elif E is not None and nu is not None:
        if mat.get("Plane Strain", False):
            k = E / (2.0 * (1.0 - nu - 2.0 * nu * nu))
        elif mat.get("Plane Stress", False):
            k = E / (2.0 * (1.0 - nu))
        else:
            k = E / (3.0 * (1.0 - 2.0 * nu))
    else:
        k = (2.0 * G * (1.0 + nu)) / (3.0 * (1.0 - 2.0 * nu))
    if G is not None:
        g = G
    elif K is not None and E is not None:
        g = (3.0 * K * E) / (9.0 * K - E)
    elif K is not None and nu is not None:
        g = (3.0 * K * (1.0 - 2.0 * nu)) / (2.0 * (1.0 + nu))
    else:
        g = E / (2.0 * (1.0 + nu))
    return float(k), float(g)


# ----------------------------------------------------------- discretization
_PLANE_KEYS = ["Normal_X", "Normal_Y", "Normal_Z", "Lower_Left_Corner_X", "Lower_Left_Corner_Y",
               "Lower_Left_Corner_Z", "Bottom_Unit_Vector_X", "Bottom_Unit_Vector_Y",
               "Bottom_Unit_Vector_Z", "Bottom_Length", "Side_Length"]


def discretize(deck, base_dir, horizon):
    """(x, volume, neighbour list, point sets, block names, block index of
    every point) of the deck's discretization."""
    d = deck["Discretization"]
    typ = d.get("Type")
    planes = []
    for flt in d.get("Bond Filters", {}).values():
        if flt.get("Type") != "Rectangular_Plane":
            raise DeckError("bond filter type %r not supported" % flt.get("Type"))
        planes.append([float(flt[k]) for k in _PLANE_KEYS])
    planes = np.array(planes) if planes else None
    if typ == "Exodus":
        mesh = EX.read_mesh(os.path.join(base_dir, d["Input Mesh File"]))
        x, vol, blk, sets = EX.discretize(mesh)
        nl = radius_search(x, horizon, planes)
        omit = d.get("Omit Bonds Between Blocks", "None")
        if omit == "All":
            nl = omit_bonds_between_blocks(nl, blk)
        elif omit != "None":
            raise DeckError("Omit Bonds Between Blocks: All or None, not %r" % omit)
        return x, vol, nl, sets, [b[0] for b in mesh["blocks"]], blk
    if typ == "PdQuickGrid":
        g = d["TensorProduct3DMeshGenerator"]
        n = [g["Number Points X"], g["Number Points Y"], g["Number Points Z"]]
        o = [g["X Origin"], g["Y Origin"], g["Z Origin"]]
        L = [g["X Length"], g["Y Length"], g["Z Length"]]
        if planes is not None:
            raise DeckError("PdQuickGrid does not take bond filters (as in the reference)")
        sph = d.get("NeighborhoodType", "Parallelepiped") == "Spherical"
        x, vol, nl = quickgrid(n, o, L, horizon, spherical=sph)
        return x, vol, nl, {}, ["block_1"], np.zeros(len(vol), np.int64)
    raise DeckError("discretization type %r not supported" % typ)


def omit_bonds_between_blocks(nl, blk):
    """The neighbour list without the bonds whose ends lie in different
    blocks (ExodusDiscretization::filterBonds with "All")."""
    nl = np.asarray(nl)
    out = []
    p = 0
    for i in range(len(blk)):
        n = int(nl[p])
        row = nl[p + 1:p + 1 + n]
        keep = row[blk[row] == blk[i]]
        out.append(len(keep))
        out.extend(int(j) for j in keep)
        p += n + 1
    return np.array(out, np.int32)


def contact_params(deck):
    """Contact parameters of the deck (None without a Contact list): search
    radius and frequency and the first model's constants, as ContactManager
    uses them (src/core/Peridigm_ContactManager.cpp:228-253)."""
    c = deck.get("Contact")
    if not c:
        return None
    models = list(c.get("Models", {}).values())
    if not models:
        raise DeckError("Contact needs a model")
    m = models[0]
    if m.get("Contact Model") != "Short Range Force":
        raise DeckError("contact model %r not supported" % m.get("Contact Model"))
    return dict(search_radius=float(c["Search Radius"]),
                search_frequency=int(c["Search Frequency"]),
                contact_radius=float(m["Contact Radius"]),
                spring_constant=float(m["Spring Constant"]),
                friction=float(m.get("Friction Coefficient", 0.0)))


# ---------------------------------------------------------------- the run
_COORD = {"x": 0, "y": 1, "z": 2}


def _node_sets(deck, sets):
    """The mesh's point sets plus those listed in the deck itself (any plain
    parameter whose name holds "Node Set",
    src/core/Peridigm_BoundaryAndInitialConditionManager.cpp:80-86: 1-based
    ids, space separated)."""
    sets = dict(sets)
    for name, val in deck.get("Boundary Conditions", {}).items():
        if not isinstance(val, dict) and "NODE SET" in name.upper():
            sets[name] = np.array([int(tk) - 1 for tk in str(val).split()], np.int64)
    return sets


def _temperature(deck, x, sets):
    """Prescribed temperatures as T(t) = c1 t + c0 per point, (c0, c1), or
    None without any: "Prescribed Temperature" sets the temperature and its
    change from the initial value, T(t) - T(0), which drives the thermal
    strains (src/core/Peridigm_BoundaryAndInitialConditionManager.cpp:148-159)."""
    sets = _node_sets(deck, sets)
    N = x.shape[0]
    c0, c1, any_ = np.zeros(N), np.zeros(N), False
    for name, bc in deck.get("Boundary Conditions", {}).items():
        if isinstance(bc, dict) and bc.get("Type") == "Prescribed Temperature":
            ids = sets[bc["Node Set"]]
            a0, a1 = linear_in_t(str(bc["Value"]), x[ids, 0], x[ids, 1], x[ids, 2])
            c0[ids], c1[ids] = a0, a1
            any_ = True
    return (c0, c1) if any_ else None


def _bcs(deck, x, sets):
    """(v0, u0, prescribed [(ids, coord, c0, c1)]) from the deck's conditions."""
    N = x.shape[0]
    v0 = np.zeros((N, 3))
    u0 = np.zeros((N, 3))
    presc = []
    sets = _node_sets(deck, sets)
    bcl = deck.get("Boundary Conditions", {})
    for name, bc in bcl.items():
        if not isinstance(bc, dict):
            continue
        typ = bc.get("Type")
        if typ == "Prescribed Temperature":
            continue  # _temperature
        ns = bc.get("Node Set")
        if ns not in sets:
            raise DeckError("node set %r of %r not found" % (ns, name))
        ids = sets[ns]
        c = _COORD[bc["Coordinate"].strip().lower()]
        px, py, pz = x[ids, 0], x[ids, 1], x[ids, 2]
        if typ == "Initial Velocity":
            v0[ids, c] = evaluate_program(str(bc["Value"]), px, py, pz, 0.0)
        elif typ == "Initial Displacement":
            u0[ids, c] = evaluate_program(str(bc["Value"]), px, py, pz, 0.0)
        elif typ == "Prescribed Displacement":
            c0, c1 = linear_in_t(str(bc["Value"]), px, py, pz)
            pairs = np.stack([c0, c1], 1)
            uniq, inv = np.unique(pairs, axis=0, return_inverse=True)
            for k, (a0, a1) in enumerate(uniq):
                presc.append((ids[inv.ravel() == k], c, float(a0), float(a1)))
        else:
            raise DeckError("boundary condition type %r not supported" % typ)
    return v0, u0, presc


def _material(deck, grp):
    """Material, damage and influence constants of a block group."""
    mat = deck["Materials"][grp["Material"]]
    model = mat.get("Material Model")
    planar, height = 0, 0.0
    if model in ("Elastic", "Elastic Partial Volume", "Linear LPS Partial Volume"):
        # the partial-volume models without partial volumes (the element-horizon
        # intersections are not computed): node volumes, as the reference uses
        # them then
        if mat.get("Use Partial Volume", False) or mat.get("Use Analytic Weighted Volume", False):
            raise DeckError("partial volumes and analytic weighted volumes are not supported")
        K, mu = elastic_constants(mat)
    elif model == "Elastic Bond Based":
        # one constant; the second is Poisson's ratio 1/4 (1/3 in plane
        # stress), src/materials/Peridigm_ElasticBondBasedMaterial.cpp:59-82
        planar = 1 if mat.get("Plane Strain", False) else (2 if mat.get("Plane Stress", False) else 0)
        height = float(mat.get("Height", 0.0))
        one = {k: mat[k] for k in ("Bulk Modulus", "Young's Modulus") if k in mat}
        if len(one) != 1 or "Shear Modulus" in mat or "Poisson's Ratio" in mat:
            raise DeckError("the bond-based model takes only the bulk or Young's modulus")
        one["Poisson's Ratio"] = 1.0 / 3.0 if planar == 2 else 0.25
        one["Plane Stress"] = planar == 2
        K, mu = elastic_constants(one)
    else:
        raise DeckError("material model %r not supported" % model)
    sc, sc_fn = 0.0, None
    if "Damage Model" in grp:
        dm = deck["Damage Models"][grp["Damage Model"]]
        if dm.get("Damage Model") == "Critical Stretch":
            sc = float(dm["Critical Stretch"])
        elif dm.get("Damage Model") == "Time Dependent Critical Stretch":
            sc_fn = str(dm["Time Dependent Critical Stretch"])
            sc = 1.0  # enables the damage model; the function sets the value
        else:
            raise DeckError("damage model %r not supported" % dm.get("Damage Model"))
    return dict(model=model, K=K, mu=mu, planar=planar, height=height,
                rho=float(mat["Density"]),
                # one influence function for the run, from the discretization
                # list (src/core/Peridigm.cpp:184-185)
                infl=_influence(deck["Discretization"].get("Influence Function", "One")),
                sc=sc, sc_fn=sc_fn, alpha_t=float(mat.get("Thermal Expansion Coefficient", 0.0)))


def _influence(name):
    if name not in INFLUENCE:
        raise DeckError("influence function %r not supported (One, Parabolic Decay, Gaussian)" % name)
    return INFLUENCE[name]


def _subset(idx, N, nl, v0, u0, presc):
    """The points idx (ascending) as a problem of their own: renumbered
    neighbour list (no bond may leave the subset) and conditions."""
    new = np.full(N, -1, np.int64)
    new[idx] = np.arange(len(idx))
    nl = np.asarray(nl)
    heads = np.zeros(N, np.int64)
    p = 0
    for i in range(N):
        heads[i] = p
        p += int(nl[p]) + 1
    out = []
    for i in idx:
        h = heads[i]
        n = int(nl[h])
        row = new[nl[h + 1:h + 1 + n]]
        if n and row.min() < 0:
            raise DeckError("bonds join block groups of different materials")
        out.append(n)
        out.extend(int(j) for j in row)
    sub_presc = []
    for ids, c, c0, c1 in presc:
        keep = new[np.asarray(ids, np.int64)]
        keep = keep[keep >= 0]
        if keep.size:
            sub_presc.append((keep, c, c0, c1))
    return np.array(out, np.int32), v0[idx], u0[idx], sub_presc


def run_deck(path, output=None, lanes=4):
    """Runs a deck. Returns dict(times, frames {variable: (frames, N)}, x,
    volume, neighbour list, info); with `output` a file name, also writes the
    frames there in ExodusII. Block groups of different materials must share
    one horizon and no bonds; each runs in a context of its own."""
    deck = parse_deck(path)
    base = os.path.dirname(os.path.abspath(path))
    groups = list(deck["Blocks"].values())
    horizons = {float(g["Horizon"]) for g in groups}
    if len(horizons) != 1:
        raise DeckError("block groups must share one horizon")
    horizon = horizons.pop()
    x, vol, nl, sets, mesh_blocks, blk = discretize(deck, base, horizon)
    N = x.shape[0]
    owner = {}
    for gi, g in enumerate(groups):
        for nm in str(g.get("Block Names", "")).split():
            owner[nm] = gi
    if owner and set(owner) != set(mesh_blocks):
        raise DeckError("the block groups must hold every block of the mesh (%s)" % mesh_blocks)
    point_group = (np.array([owner[mesh_blocks[b]] for b in blk], np.int64) if owner
                   else np.zeros(N, np.int64))
    # groups of one material, damage model and influence function run together
    sig = [(g.get("Material"), g.get("Damage Model")) for g in groups]
    run_of = {}
    for s_ in sig:
        run_of.setdefault(s_, len(run_of))
    point_run = np.array([run_of[sig[g]] for g in point_group], np.int64)
    v0, u0, presc = _bcs(deck, x, sets)
    temp = _temperature(deck, x, sets)
    contact = contact_params(deck)
    if contact is not None and len(run_of) > 1:
        raise DeckError("contact between block groups of different materials is not supported")
    solver = deck["Solver"]
    if len(run_of) > 1 and "Verlet" in solver and "Fixed dt" not in solver["Verlet"]:
        raise DeckError("several materials need a fixed time step")
    out = deck.get("Output", deck.get("Output1", {}))
    want = {k for k, v in out.get("Output Variables", {}).items() if v}
    times, frames, info = None, {}, {}
    for s_, r in run_of.items():
        idx = np.nonzero(point_run == r)[0]
        if idx.size == 0:
            continue
        grp = groups[sig.index(s_)]
        if len(run_of) == 1:
            sub = (nl, v0, u0, presc)
        else:
            sub = _subset(idx, N, nl, v0, u0, presc)
        t_r, f_r, info = _run_points(deck, _material(deck, grp), horizon, x[idx], vol[idx], sub,
                                     contact, lanes,
                                     None if temp is None else (temp[0][idx], temp[1][idx]))
        if times is None:
            times = t_r
        elif len(t_r) != len(times) or np.abs(np.asarray(t_r) - times).max() > 0:
            raise DeckError("block groups ran to different output times")
        for name, arr in f_r.items():
            if name not in frames:
                frames[name] = np.zeros((arr.shape[0], N))
            frames[name][:, idx] = arr
    for name in ("Element_Id",):
        if name in want:
            frames[name] = np.tile(np.arange(1, N + 1, dtype=float), (len(times), 1))
    res = dict(times=np.array(times), frames=frames, x=x, volume=vol, nlist=nl, info=info)
    if output:
        nodal = {k: v for k, v in res["frames"].items()
                 if k[:-1] in ("Displacement", "Velocity", "Force_Density", "Force",
                               "Contact_Force_Density")}
        elem = {k: v for k, v in res["frames"].items() if k not in nodal}
        EX.write_results(output, x, res["times"], nodal, elem)
    return res


def _run_points(deck, mat, horizon, x, vol, sub, contact, lanes, temp=None):
    """Runs the deck's solver on one set of points of one material: returns
    (times, frames {variable: (frames, n)}, info). temp: prescribed
    temperatures T(t) = c1 t + c0 as (c0, c1), or None."""
    nl, v0, u0, presc = sub
    N = x.shape[0]
    solver = deck["Solver"]
    t0 = float(solver.get("Initial Time", 0.0))
    t1 = float(solver.get("Final Time", 1.0))
    out = deck.get("Output", deck.get("Output1", {}))
    freq = int(out.get("Output Frequency", 1))
    want = {k for k, v in out.get("Output Variables", {}).items() if v}
    model, rho, sc, sc_fn = mat["model"], mat["rho"], mat["sc"], mat["sc_fn"]
    ctx = Context(0, lanes=lanes)
    try:
        ctx.set_topology(nl, N)
        if model == "Elastic Bond Based":
            ctx.set_material_model(model, mat["planar"], mat["height"])
        elif model == "Linear LPS Partial Volume":
            ctx.set_material_model(model)
        ctx.set_constants(x, vol, mat["K"], mat["mu"], horizon, mat["infl"], sc)
        if contact is not None:
            ctx.set_contact(**contact)
        if mat["alpha_t"] != 0.0:
            ctx.set_thermal_expansion(mat["alpha_t"])
        if temp is not None:
            # the change from the initial value, T(t) - T(0) = c1 t
            ctx.set_temperature_change(np.zeros(N), temp[1])
        m = ctx.initialize()
        deg = _degrees(nl, N)
        frames = {}
        times = []

        def record(t, st):
            times.append(t)
            f = st["force"]
            fields = {
                "Displacement": st["u"], "Velocity": st["v"], "Force_Density": f,
                "Force": f * vol[:, None],
            }
            if "contact_force" in st:
                fields["Contact_Force_Density"] = st["contact_force"]
            for name, arr in fields.items():
                if name in want:
                    for c, ax in enumerate("XYZ"):
                        frames.setdefault(name + ax, []).append(arr[:, c].copy())
            if temp is not None:
                st = dict(st, temperature=temp[1] * t + temp[0], temperature_change=temp[1] * t)
            scal = {"Dilatation": st["dilatation"], "Weighted_Volume": m,
                    "Damage": st["damage"], "Number_Of_Neighbors": deg.astype(float),
                    "Volume": vol, "Element_Id": np.arange(1, N + 1, dtype=float),
                    "Proc_Num": np.zeros(N)}
            if temp is not None:
                scal["Temperature"] = st["temperature"]
                scal["Temperature_Change"] = st["temperature_change"]
            for name, arr in scal.items():
                if name in want:
                    frames.setdefault(name, []).append(np.asarray(arr, float).copy())

        info = {}
        if "Verlet" in solver:
            vp = solver["Verlet"]
            dt = vp["Fixed dt"] if "Fixed dt" in vp else ctx.critical_time_step(rho)
            dt *= vp.get("Safety Factor", 1.0)
            nsteps = int(math.floor((t1 - t0) / dt))
            if sc_fn is not None:
                # the damage function at every evaluation time t0 + k dt
                ts = np.array([t0 + (k * dt) for k in range(nsteps + 1)])
                vals = np.array([float(evaluate_program(sc_fn, np.zeros(1), np.zeros(1),
                                                        np.zeros(1), tk)[0]) for tk in ts])
                ctx.set_critical_stretch_function(ts, vals)
            ctx.set_density(np.full(N, rho))
            ctx.set_bcs(presc)
            ctx.upload_state(u=u0, v=v0)
            done = 0
            ctx.verlet_run(0, dt, t0=t0, initial_eval=True)
            record(t0, _explicit_state(ctx, contact is not None))
            while done < nsteps:
                k = min(freq, nsteps - done)
                ctx.verlet_run(k, dt, t0=t0 + done * dt, initial_eval=False)
                done += k
                if done % freq == 0 or done == nsteps:
                    record(t0 + done * dt, _explicit_state(ctx, contact is not None))
            info = dict(solver="Verlet", dt=dt, steps=nsteps)
        elif "QuasiStatic" in solver:
            from . import quasistatic as Q
            qp = solver["QuasiStatic"]
            nls = int(qp["Number of Load Steps"])
            ts = [t0] + [t0 + (i + 1) * (t1 - t0) / nls for i in range(nls)]
            prm = dict(maxSolverIterations=int(qp.get("Maximum Solver Iterations", 10)),
                       tolerance=float(qp.get("Relative Tolerance", 1e-6)),
                       disableHeuristics=int(bool(solver.get("Disable Heuristics", False))),
                       linearTolerance=float(qp.get("Belos Relative Tolerance", 1e-4)))
            if "Absolute Tolerance" in qp:
                prm.update(absoluteTolerance=1, tolerance=float(qp["Absolute Tolerance"]))
            record(t0, _qs_state(ctx, u0, Q, t0, np.zeros_like(u0)))
            # one run; a frame every `freq` accepted load steps, with the
            # velocity of the step's increment (src/core/Peridigm.cpp:2683-2687)
            u, steps = Q.run(ctx, presc, ts, u0=u0, record=True, **prm)
            u_prev = u0
            for k, st in enumerate(steps):
                if (k + 1) % freq == 0 or k == len(steps) - 1:
                    v = (st["u"] - u_prev) / (st["t1"] - st["t0"])
                    record(st["t1"], _qs_state(ctx, st["u"], Q, st["t1"], v))
                u_prev = st["u"]
            info = dict(solver="QuasiStatic", steps=[{k: v for k, v in st.items() if k != "u"}
                                                      for st in steps])
        else:
            raise DeckError("solver not supported (Verlet or QuasiStatic)")
    finally:
        ctx.close()
    return times, {k: np.array(v) for k, v in frames.items()}, info


def _degrees(nl, N):
    out = np.zeros(N, np.int64)
    p = 0
    for i in range(N):
        out[i] = nl[p]
        p += nl[p] + 1
    return out


def _explicit_state(ctx, contact=False):
    st = ctx.download_state()
    out = dict(u=st["u"], v=st["v"], force=st["force"], dilatation=st["theta"],
               damage=st["damage"])
    if contact:
        out["contact_force"] = ctx.contact_force_density()
    return out


def _qs_state(ctx, u, Q, t, v):
    """FP64 state of a quasi-static frame at time t: displacement, velocity,
    force density and dilatation at u (the quasi-static kernels, the
    temperature change at t)."""
    ctx.set_temperature_time(t)
    f, th = Q.force_and_dilatation(ctx, u)
    N = ctx.N
    return dict(u=np.asarray(u, float).reshape(N, 3), v=np.asarray(v, float).reshape(N, 3),
                force=f, dilatation=th, damage=np.zeros(N))
