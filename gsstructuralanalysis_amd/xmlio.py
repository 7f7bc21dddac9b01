"""gsFileData XML problem input and point-evaluation / CSV output for the shell solvers.

Input mirrors the reading section of the reference's XML drivers
(solvers/static_shell_XML.cpp:90-206, solvers/quasistatic_shell_XML.cpp, the files under
filedata/pde/): the geometry (ID 0: a MultiPatch naming one TensorBSpline2, or the geometry
itself), the material (ID 10: a MaterialMatrix of type Linear / Linear3, the St. Venant-Kirchhoff
law; or the older Function IDs 10-13 for thickness, E, nu, density), boundary conditions (ID 20),
the surface load (ID 21), the follower pressure (ID 22), point loads (IDs 30-32), reference points
and values (IDs 50-52) and the option lists (IDs 90-95).  The geometry is degree-elevated and uniformly refined as the drivers'
-e / -r switches do, and handed to the problem builder as an interpolating control net of the
refined space (which contains the given surface, so the surface is reproduced exactly).

What the GPU problem class cannot represent raises XmlError naming the item: several patches,
rational (NURBS) geometry, non-uniform or repeated interior knots together with degree elevation,
unequal degrees in the two directions, other material laws, non-zero Dirichlet values,
"Collapsed" / weak conditions and non-constant functions.  A constant pressure (ID 22) is the
follower pressure of ksh_problem_set_pressure.

Output mirrors constructDisplacement + patch evaluation at reference points
(solvers/static_shell_XML.cpp:320-330) and gsStructuralAnalysisOutput's CSV layout
(src/gsStructuralAnalysisTools/gsStructuralAnalysisUtils.h:21-160), and the step files of
benchmarks/benchmark_Roof.cpp:540-624.
"""
from __future__ import annotations

import ast
import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import geometry as g

# gismo's boxSide numbering (1 west, 2 east, 3 south, 4 north) and the side names in the files
_SIDE_OF_INDEX = {1: "west", 2: "east", 3: "south", 4: "north"}
_SIDE_NAMES = {"west": "west", "east": "east", "south": "south", "north": "north", "left": "west", "right": "east",
               "bottom": "south", "top": "north"}
# boxCorner numbering: 1 southwest, 2 southeast, 3 northwest, 4 northeast -> (row, column) of the net
_CORNERS = {1: (0, 0), 2: (0, -1), 3: (-1, 0), 4: (-1, -1)}


class XmlError(ValueError):
    """The file holds something the shell problem class cannot represent (or is malformed)."""


# ---- FunctionExpr values ------------------------------------------------------------------------
_FUNCS = {"sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "log": math.log,
          "abs": abs, "atan": math.atan, "asin": math.asin, "acos": math.acos}
_CONSTS = {"pi": math.pi, "e": math.e}


def constant_expr(text: str) -> float:
    """Value of a constant gsFunctionExpr expression (numbers, + - * / ^, parentheses, pi, e and
    the elementary functions); an expression of x, y, z or t raises XmlError."""
    src = " ".join(str(text).split()).replace("^", "**")
    if not src:
        raise XmlError("empty function expression")
    try:
        tree = ast.parse(src, mode="eval")
    except SyntaxError as exc:
        raise XmlError(f"cannot read function expression {text!r}") from exc

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)):
            a, b = ev(node.left), ev(node.right)
            if isinstance(node.op, ast.Add):
                return a + b
            if isinstance(node.op, ast.Sub):
                return a - b
            if isinstance(node.op, ast.Mult):
                return a * b
            if isinstance(node.op, ast.Div):
                if b == 0:
                    raise XmlError(f"division by zero in {text!r}")
                return a / b
            return a ** b
        if isinstance(node, ast.Name):
            if node.id in _CONSTS:
                return _CONSTS[node.id]
            raise XmlError(f"function {text!r} depends on {node.id!r}: only constant functions are supported")
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS \
                and len(node.args) == 1 and not node.keywords:
            return float(_FUNCS[node.func.id](ev(node.args[0])))
        raise XmlError(f"unsupported construct in function expression {text!r}")

    return float(ev(tree))


def function_value(el: ET.Element, dim: int | None = None) -> np.ndarray:
    """Constant value of a <Function type="FunctionExpr"> element: its <c> components, or its text
    (a scalar; with dim > 1 and one expression, every component takes it)."""
    if el.get("type", "FunctionExpr") != "FunctionExpr":
        raise XmlError(f"function type {el.get('type')!r} is not supported (FunctionExpr only)")
    comps = el.findall("c")
    if comps:
        return np.array([constant_expr(c.text or "") for c in comps])
    v = constant_expr(el.text or "")
    d = int(el.get("dim", dim or 1)) if dim is None else dim
    return np.full(max(d, 1), v)


def _numbers(text: str | None) -> np.ndarray:
    return np.array([float(v) for v in (text or "").split()], dtype=np.float64)


# ---- the file --------------------------------------------------------------------------------
@dataclass
class ShellXml:
    """A problem read from a gsFileData file: the patch for ShellProblem, the reference points
    (parametric, 2 x k), their patch indices and reference values, and the option lists by ID
    (label -> typed value)."""
    patch: g.Patch
    ref_points: np.ndarray = field(default_factory=lambda: np.zeros((2, 0)))
    ref_patches: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    ref_values: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    options: dict = field(default_factory=dict)
    source: str = ""


def _by_id(root: ET.Element) -> dict:
    out = {}
    for el in root:
        if el.get("id") is not None:
            out.setdefault(int(el.get("id")), el)
    return out


def _matrix(el: ET.Element | None, rows: int | None = None) -> np.ndarray:
    if el is None:
        return np.zeros((rows or 0, 0))
    vals = _numbers(el.text)
    r = int(el.get("rows", rows or 1))
    c = int(el.get("cols", 0))
    if r == 0:
        return np.zeros((0, 0))
    if vals.size == 0:
        return np.zeros((r, 0))
    if c == 0 or vals.size != r * c:
        c = vals.size // r
    if vals.size != r * c:
        raise XmlError(f"matrix id={el.get('id')}: {vals.size} values do not fill {r} rows")
    return vals.reshape(r, c)


def option_list(el: ET.Element) -> dict:
    """<OptionList>: label -> int / float / bool / str."""
    out = {}
    for o in el:
        label, val = o.get("label"), o.get("value")
        if label is None:
            continue
        if o.tag == "int":
            out[label] = int(float(val))
        elif o.tag == "real":
            out[label] = float(val)
        elif o.tag in ("bool", "switch"):
            out[label] = str(val).strip().lower() in ("1", "true", "yes", "on")
        else:
            out[label] = val
    return out


def _knots(el: ET.Element) -> tuple[int, np.ndarray]:
    kv = el.find("KnotVector")
    if kv is None:
        raise XmlError("B-spline basis without a KnotVector")
    return int(kv.get("degree")), _numbers(kv.text)


def _uniform_elements(p: int, U: np.ndarray, direction: str, elevate: int) -> int:
    """Element count of an open knot vector with uniformly spaced simple interior knots."""
    if U.size < 2 * (p + 1):
        raise XmlError(f"knot vector in {direction} is too short for degree {p}")
    a, b = U[0], U[-1]
    if not (np.all(U[:p + 1] == a) and np.all(U[-(p + 1):] == b)):
        raise XmlError(f"knot vector in {direction} is not open")
    inner = U[p + 1:-(p + 1)]
    breaks = np.concatenate([[a], inner, [b]])
    if np.any(np.diff(breaks) <= 0):
        raise XmlError(f"knot vector in {direction} has repeated interior knots")
    n = breaks.size - 1
    if n > 1 and elevate > 0:
        raise XmlError(f"degree elevation of a knot vector with interior knots ({direction}) lowers the continuity "
                       f"at those knots, which the uniform C^(p-1) spaces here cannot represent")
    if not np.allclose(np.diff(breaks), (b - a) / n, rtol=1e-12, atol=0):
        raise XmlError(f"knot vector in {direction} is not uniform")
    return n


class _BSplineSurface:
    """Evaluation of a tensor B-spline surface given by knot vectors and its control net."""

    def __init__(self, pu, Uu, pv, Uv, coefs):
        self.pu, self.Uu, self.pv, self.Uv = pu, Uu, pv, Uv
        self.mu, self.mv = Uu.size - pu - 1, Uv.size - pv - 1
        if coefs.shape != (self.mu * self.mv, 3):
            raise XmlError(f"geometry has {coefs.shape[0]} control points, the basis {self.mu * self.mv}")
        self.C = coefs.reshape(self.mv, self.mu, 3)

    @staticmethod
    def _basis(p, U, x):
        x = np.clip(np.asarray(x, dtype=np.float64), U[0], U[-1])
        m = U.size - p - 1
        B = np.zeros((x.size, U.size - 1))
        for k, xv in enumerate(x):
            span = int(np.searchsorted(U, xv, side="right") - 1)
            span = min(max(span, p), m - 1)
            B[k, span] = 1.0
        for q in range(1, p + 1):
            Bn = np.zeros((x.size, U.size - 1 - q))
            for i in range(U.size - 1 - q):
                d1, d2 = U[i + q] - U[i], U[i + q + 1] - U[i + 1]
                if d1 > 0:
                    Bn[:, i] += (x - U[i]) / d1 * B[:, i]
                if d2 > 0:
                    Bn[:, i] += (U[i + q + 1] - x) / d2 * B[:, i + 1]
            B = Bn
        return B[:, :m]

    def __call__(self, xi, eta):
        shape = np.shape(xi)
        xs = self.Uu[0] + np.ravel(xi) * (self.Uu[-1] - self.Uu[0])
        ys = self.Uv[0] + np.ravel(eta) * (self.Uv[-1] - self.Uv[0])
        Bu, Bv = self._basis(self.pu, self.Uu, xs), self._basis(self.pv, self.Uv, ys)
        out = np.einsum("ki,kj,jic->kc", Bu, Bv, self.C)
        return out.reshape(shape + (3,))


def _geometry(ids: dict, root: ET.Element, geometry_root: ET.Element | None):
    pool = _by_id(geometry_root) if geometry_root is not None else {}  # the geometry file first
    for k, v in ids.items():
        pool.setdefault(k, v)
    g0 = pool.get(0)
    if g0 is None:
        raise XmlError("no geometry (ID 0) in the file; pass the geometry file")
    names = {}
    if g0.tag == "MultiPatch":
        patches = g0.find("patches")
        if patches is None:
            raise XmlError("MultiPatch without patches")
        if patches.get("type") == "id_range":
            lo, hi = (int(v) for v in patches.text.split())
            pids = list(range(lo, hi + 1))
        else:
            pids = [int(v) for v in patches.text.split()]
        if len(pids) != 1:
            raise XmlError(f"{len(pids)} patches: the shell problem class holds a single patch")
        for b in g0.findall("boundary"):
            vals = [int(v) for v in (b.text or "").split()]
            if b.get("name"):
                names[b.get("name")] = [_SIDE_OF_INDEX.get(s) for s in vals[1::2]]
        geo = pool.get(pids[0])
        if geo is None:
            raise XmlError(f"geometry id {pids[0]} named by the MultiPatch is not in the file")
    else:
        geo = g0
    if geo.get("type") != "TensorBSpline2":
        raise XmlError(f"geometry type {geo.get('type')!r}: only TensorBSpline2 (no NURBS) is supported")
    basis = geo.find("Basis")
    subs = {int(b.get("index", i)): b for i, b in enumerate(basis.findall("Basis"))}
    (pu, Uu), (pv, Uv) = _knots(subs[0]), _knots(subs[1])
    coefs = _numbers(geo.find("coefs").text)
    gd = int(geo.find("coefs").get("geoDim", 3))
    if gd != 3:
        raise XmlError(f"geometry of dimension {gd}: a surface in 3D is needed")
    return (pu, Uu, pv, Uv, coefs.reshape(-1, 3)), names


def _material(ids: dict):
    el = ids.get(10)
    if el is None:
        raise XmlError("no material (ID 10)")
    if el.tag == "MaterialMatrix":
        kind = el.get("type", "")
        if kind not in ("Linear", "Linear3", "Linear2"):
            raise XmlError(f"material {kind!r}: only the St. Venant-Kirchhoff law (Linear) is implemented")

        def one(tag, idx=0):
            parent = el.find(tag)
            if parent is None:
                raise XmlError(f"material without {tag}")
            fs = {int(f.get("index", i)): f for i, f in enumerate(parent.findall("Function"))}
            return float(function_value(fs[idx], 1)[0])

        t = one("Thickness")
        rho = one("Density") if el.find("Density") is not None else 0.0
        E, nu = one("Parameters", 0), one("Parameters", 1)
        return t, E, nu, rho
    # the older form: Function IDs 10 (thickness), 11 (E), 12 (nu), 13 (density)
    vals = []
    for k in (10, 11, 12, 13):
        f = ids.get(k)
        vals.append(float(function_value(f, 1)[0]) if f is not None else 0.0)
    return vals[0], vals[1], vals[2], vals[3]


def read_problem(path: str, geometry: str | None = None, elevate: int = 0, refine: int = 0) -> ShellXml:
    """Read a shell problem file (and optionally its geometry from another file), degree-elevate
    `elevate` times and refine uniformly `refine` times (the drivers' -e and -r)."""
    root = ET.parse(path).getroot()
    groot = ET.parse(geometry).getroot() if geometry else None
    ids = _by_id(root)
    (pu, Uu, pv, Uv, coefs), side_names = _geometry(ids, root, groot)
    if pu != pv:
        raise XmlError(f"degrees {pu} and {pv} differ: the problem class takes one degree in both directions")
    p = pu + elevate
    nx = _uniform_elements(pu, Uu, "u", elevate) * 2 ** refine
    ny = _uniform_elements(pv, Uv, "v", elevate) * 2 ** refine
    surf = _BSplineSurface(pu, Uu, pv, Uv, coefs)
    t, E, nu, rho = _material(ids)
    pt = g.Patch(p, nx, ny, g.interpolate_surface(p, nx, ny, surf), t, E, nu, rho)
    # boundary conditions (ID 20)
    bcs = ids.get(20)
    if bcs is not None:
        funcs = {int(f.get("index", i)): f for i, f in enumerate(bcs.findall("Function"))}
        for bc in bcs.findall("bc"):
            kind = bc.get("type", "")
            sides = []
            if bc.get("name"):
                nm = bc.get("name")
                sides = side_names.get(nm) or [_SIDE_NAMES.get(nm)]
                if None in sides:
                    raise XmlError(f"boundary name {nm!r} is unknown")
            vals = [int(v) for v in (bc.text or "").split()]
            for patch_i, s in zip(vals[0::2], vals[1::2]):
                if patch_i != 0:
                    raise XmlError(f"boundary condition on patch {patch_i} of a single-patch problem")
                if s not in _SIDE_OF_INDEX:
                    raise XmlError(f"side {s} is not an edge of a surface patch")
                sides.append(_SIDE_OF_INDEX[s])
            if not sides:
                continue  # placeholder without sides
            comp = int(bc.get("component", -1))
            comps = (0, 1, 2) if comp < 0 else (comp,)
            fidx = int(bc.get("function", 0))
            if kind in ("Dirichlet", "Clamped"):
                if fidx in funcs and np.any(function_value(funcs[fidx], 3) != 0):
                    raise XmlError(f"{kind} condition with a non-zero value: only homogeneous conditions are supported")
                for s in sides:
                    pt.fix(s, comps) if kind == "Dirichlet" else pt.clamp(s, comps)
            elif kind == "Neumann":
                q = function_value(funcs[fidx], 3)
                if q.size != 3:
                    raise XmlError("Neumann function must have 3 components")
                for s in sides:
                    g.add_edge_load(pt, s, q)
            else:
                raise XmlError(f"boundary condition type {kind!r} is not supported")
        for cv in bcs.findall("cv"):
            if float(constant_expr(cv.text or "0")) != 0:
                raise XmlError("corner value other than zero")
            if int(cv.get("patch", 0)) != 0:
                raise XmlError("corner value on a patch other than 0")
            r, c = _CORNERS[int(cv.get("corner"))]
            comp = int(cv.get("component", -1))
            net = np.arange(pt.mx * pt.my).reshape(pt.my, pt.mx)
            cp = int(net[r, c])
            for k in ((0, 1, 2) if comp < 0 else (comp,)):
                pt.fixed[3 * cp + k] = 1
    # loads
    if 21 in ids:
        pt.fsurf = tuple(function_value(ids[21], 3))
    if 22 in ids:  # follower pressure (static_shell_XML.cpp:145-149, 236-237: setPressure)
        pt.pressure = float(function_value(ids[22], 1)[0])
    pts, loads = _matrix(ids.get(30), 2), _matrix(ids.get(31), 3)
    if pts.shape[1] != loads.shape[1]:
        raise XmlError("point load locations and vectors differ in number")
    if 32 in ids and np.any(_matrix(ids[32], 1) != 0):
        raise XmlError("point load on a patch other than 0")
    for k in range(pts.shape[1]):
        pt.add_point_load(pts[0, k], pts[1, k], loads[:, k])
    out = ShellXml(pt, source=path)
    refp = _matrix(ids.get(50), 2)
    if refp.shape[0] == 3:
        raise XmlError("reference points in physical coordinates: give them in parametric coordinates")
    out.ref_points = refp if refp.size else np.zeros((2, 0))
    out.ref_patches = _matrix(ids.get(51), 1).ravel().astype(np.int64) if 51 in ids else np.zeros(0, dtype=np.int64)
    out.ref_values = _matrix(ids.get(52)) if 52 in ids else np.zeros((0, 0))
    out.options = {k: option_list(v) for k, v in ids.items() if v.tag == "OptionList"}
    return out


# ---- output ---------------------------------------------------------------------------------------
def basis_at(pt: g.Patch, params) -> np.ndarray:
    """Tensor basis values N_cp(xi, eta) at parametric points (2 x k) -> (k, ncp)."""
    params = np.asarray(params, dtype=np.float64).reshape(2, -1)
    Bx = g._basis_matrix(pt.p, pt.nx, params[0])
    By = g._basis_matrix(pt.p, pt.ny, params[1])
    return np.einsum("ki,kj->kji", Bx, By).reshape(params.shape[1], -1)


def displacement_at(pt: g.Patch, u_free, params) -> np.ndarray:
    """constructDisplacement(u) evaluated at parametric points: (3 x k)."""
    d = pt.to_full(np.asarray(u_free, dtype=np.float64)).reshape(-1, 3)
    return (basis_at(pt, params) @ d).T


def geometry_at(pt: g.Patch, params) -> np.ndarray:
    """The undeformed surface at parametric points: (3 x k)."""
    return (basis_at(pt, params) @ pt.ctrl).T


class StructuralAnalysisOutput:
    """gsStructuralAnalysisOutput's CSV: a header of "point<p>-<h>," for every point and point
    header, then the other headers; one row per add() with the point values (point-major) and the
    other values, each followed by a comma, at `precision` significant digits."""

    def __init__(self, fname: str, points=None, precision: int = 5):
        self.fname = fname
        self.points = np.zeros((2, 0)) if points is None else np.asarray(points, dtype=np.float64).reshape(2, -1)
        self.precision = precision
        self.nph = self.noh = 0

    def _fmt(self, v):
        return f"{float(v):.{self.precision}g}"

    def init(self, point_headers, other_headers=()):
        self.nph, self.noh = len(point_headers), len(other_headers)
        with open(self.fname, "w") as f:
            f.write("".join(f"point{p}-{h}," for p in range(self.points.shape[1]) for h in point_headers))
            f.write("".join(f"{h}," for h in other_headers))
            f.write("\n")

    def add(self, point_solutions, other_data=()):
        ps = np.asarray(point_solutions, dtype=np.float64).reshape(self.nph, -1)
        if len(other_data) != self.noh:
            raise ValueError("number of other data differs from the other headers")
        with open(self.fname, "a") as f:
            f.write("".join(self._fmt(ps[c, p]) + "," for p in range(ps.shape[1]) for c in range(ps.shape[0])))
            f.write("".join(self._fmt(v) + "," for v in other_data))
            f.write("\n")


def write_step_csv(fname: str, pt: g.Patch, points, unorm, U_steps, lambdas, indicators):
    """benchmark_Roof's step file (initStepOutput / writeStepOutput, extreme = -1): ||U|| , the
    displacement (x, y, z) at each point, the load factor and the stability indicator per step."""
    points = np.asarray(points, dtype=np.float64).reshape(2, -1)
    with open(fname, "w") as f:
        f.write("Deformation norm,")
        f.write("".join(f"point {k} - x,point {k} - y,point {k} - z," for k in range(points.shape[1])))
        f.write("Lambda,Indicator\n")
        for un, u, lam, ind in zip(unorm, U_steps, lambdas, indicators):
            d = displacement_at(pt, u, points)
            f.write(f"{un:.6g},")
            f.write("".join(f"{d[0, k]:.6g},{d[1, k]:.6g},{d[2, k]:.6g}," for k in range(points.shape[1])))
            f.write(f"{lam:.6g},{ind:.6g},\n")


# ---- writing a problem ---------------------------------------------------------------------------
def write_problem(pt: g.Patch, fname: str, ref_points=None, options: dict | None = None):
    """A gsFileData file of a problem built here: the geometry, the Linear material, the fixed
    dofs as Dirichlet / Clamped rows where they are whole sides (otherwise as corner-free
    per-control-point fixes are not expressible and ValueError is raised), the surface load and
    the point loads.  Edge loads go out as the point loads they are."""
    root = ET.Element("xml")
    mp = ET.SubElement(root, "MultiPatch", parDim="2", id="0")
    ET.SubElement(mp, "patches", type="id_range").text = "1000 1000"
    geo = ET.SubElement(root, "Geometry", type="TensorBSpline2", id="1000")
    basis = ET.SubElement(geo, "Basis", type="TensorBSplineBasis2")
    for i, n in enumerate((pt.nx, pt.ny)):
        b = ET.SubElement(basis, "Basis", type="BSplineBasis", index=str(i))
        ET.SubElement(b, "KnotVector", degree=str(pt.p)).text = " ".join(repr(float(v)) for v in g.uniform_knots(pt.p, n))
    ET.SubElement(geo, "coefs", geoDim="3").text = "\n" + "\n".join(" ".join(repr(float(v)) for v in row)
                                                                  for row in pt.ctrl) + "\n"
    mat = ET.SubElement(root, "MaterialMatrix", type="Linear3", id="10")
    for tag, vals in (("Thickness", [pt.t]), ("Density", [pt.rho]), ("Parameters", [pt.E, pt.nu])):
        el = ET.SubElement(mat, tag)
        for i, v in enumerate(vals):
            ET.SubElement(el, "Function", type="FunctionExpr", dim="3", index=str(i)).text = repr(float(v))
    bcs = ET.SubElement(root, "boundaryConditions", id="20", multipatch="0")
    ET.SubElement(bcs, "Function", type="FunctionExpr", dim="3", index="0").text = "0"
    fixed = pt.fixed.reshape(-1, 3).copy()
    covered = np.zeros_like(fixed)
    side_index = {v: k for k, v in _SIDE_OF_INDEX.items()}
    for side in ("west", "east", "south", "north"):
        for depth, kind in ((0, "Dirichlet"), (1, "Clamped")):
            row = pt._row(side, depth)
            for c in range(3):
                if np.all(fixed[row, c] == 1):
                    bc = ET.SubElement(bcs, "bc", type=kind, function="0", unknown="0", component=str(c))
                    bc.text = f"0 {side_index[side]}"
                    covered[row, c] = 1
    if np.any(fixed.astype(bool) & ~covered.astype(bool)):
        raise ValueError("fixed dofs that are not whole boundary rows cannot be written as boundary conditions")
    f21 = ET.SubElement(root, "Function", type="FunctionExpr", id="21", tag="Loads", dim="3")
    for v in pt.fsurf:
        ET.SubElement(f21, "c").text = repr(float(v))
    ET.SubElement(root, "Function", type="FunctionExpr", id="22", tag="Loads", dim="3").text = repr(
        float(getattr(pt, "pressure", 0.0)))
    k = pt.point_xi.shape[0]
    ET.SubElement(root, "Matrix", rows="2", cols=str(k), id="30", tag="Loads").text = (
        " ".join(repr(float(v)) for v in pt.point_xi[:, 0]) + "\n" + " ".join(repr(float(v)) for v in pt.point_xi[:, 1]))
    ET.SubElement(root, "Matrix", rows="3", cols=str(k), id="31", tag="Loads").text = "\n".join(
        " ".join(repr(float(v)) for v in pt.point_F[:, c]) for c in range(3))
    if ref_points is not None:
        rp = np.asarray(ref_points, dtype=np.float64).reshape(2, -1)
        ET.SubElement(root, "Matrix", rows="2", cols=str(rp.shape[1]), id="50").text = "\n".join(
            " ".join(repr(float(v)) for v in rp[r]) for r in range(2))
        ET.SubElement(root, "Matrix", rows="1", cols=str(rp.shape[1]), id="51").text = " ".join("0" * rp.shape[1])
    for oid, opts in (options or {}).items():
        ol = ET.SubElement(root, "OptionList", id=str(oid))
        for label, val in opts.items():
            tag = "bool" if isinstance(val, bool) else "int" if isinstance(val, int) else \
                "real" if isinstance(val, float) else "string"
            ET.SubElement(ol, tag, label=label, desc="", value=str(int(val) if isinstance(val, bool) else val))
    ET.ElementTree(root).write(fname, encoding="UTF-8", xml_declaration=True)


# ---- option lists -> solver keyword arguments -----------------------------------------------------
def dr_kwargs(opts: dict) -> dict:
    """gsStaticDR / gsStaticBase option list (ID 90) -> ShellProblem.run_dr keywords."""
    m = {"maxIt": "max_iter", "tol": "tol", "tolF": "tol_f", "tolU": "tol_u", "damping": "damping", "alpha": "alpha",
         "tolE": "tol_e", "ResetIt": "reset_it"}
    return {m[k]: v for k, v in opts.items() if k in m}


def newton_kwargs(opts: dict) -> dict:
    """gsStaticNewton option list (ID 91) -> ShellProblem.run_static keywords."""
    m = {"maxIt": "max_iter", "tol": "tol"}
    return {m[k]: v for k, v in opts.items() if k in m}


def alm_kwargs(opts: dict) -> dict:
    """gsALMBase option list (ID 93, solvers/quasistatic_shell_XML.cpp) -> ShellProblem.run_alm
    keywords ("Method" chooses Crisfield / Riks; "Length" is the arc length)."""
    m = {"MaxIter": "max_iter", "Tol": "tol", "TolF": "tol_f", "TolU": "tol_u", "Scaling": "scaling",
         "AngleMethod": "angle_method", "BifurcationMethod": "bifurcation_method", "Relaxation": "relaxation",
         "Quasi": "quasi", "QuasiIterations": "quasi_iterations", "AdaptiveLength": "adaptive",
         "AdaptiveIterations": "adaptive_iterations", "SingularPointFailure": "sp_fail",
         "SingularPointTestTol": "sp_test_tol", "SingularPointTestIt": "sp_test_it",
         "SingularPointComputeTolE": "sp_tol_e", "SingularPointComputeTolB": "sp_tol_b", "Perturbation": "perturbation",
         "Length2": "length_after_sp"}
    out = {m[k]: (int(v) if isinstance(v, bool) else v) for k, v in opts.items() if k in m}
    method = str(opts.get("Method", "Crisfield"))
    if method not in ("Crisfield", "Riks"):
        raise XmlError(f"arc-length method {method!r}: Crisfield or Riks")
    out["method"] = 0 if method == "Crisfield" else 1
    # BifurcationMethod 1 (eigenvalues of K) counts the same negatives as the pivots of its LDL^T
    # (Sylvester's law of inertia), so the stability and its changes are those of method 0 here
    if out.get("bifurcation_method", 0) == 1:
        out["bifurcation_method"] = 0
    return out


# ---- drivers --------------------------------------------------------------------------------------
def solve_static(x: ShellXml, method: str = "NR", csv: str | None = None, **overrides) -> dict:
    """solvers/static_shell_XML.cpp:255-335 on the device: Newton (option list 91) or dynamic
    relaxation (option list 90) at the full load, then the displacement at the reference points,
    written with gsStructuralAnalysisOutput's layout when `csv` is given."""
    from .problem import ShellProblem
    sp = ShellProblem(x.patch)
    try:
        if method == "NR":
            kw = newton_kwargs(x.options.get(91, {}))
            kw.update(overrides)
            r = sp.run_static(1, 1.0, keep_U=True, **kw)
            U = r["U"][-1] if r["steps_done"] else np.zeros(sp.nfree)
            converged = r["steps_done"] == 1
        elif method == "DR":
            kw = dr_kwargs(x.options.get(90, {}))
            kw.update(overrides)
            r = sp.run_dr(**kw)
            U, converged = r["U"], r["converged"]
        else:
            raise ValueError("method: 'NR' or 'DR'")
    finally:
        sp.close()
    pts = displacement_at(x.patch, U, x.ref_points)
    if csv is not None:
        w = StructuralAnalysisOutput(csv, x.ref_points)
        w.init(["x", "y", "z"])
        w.add(pts)
    return {"U": U, "converged": converged, "points": pts, "result": r}


def solve_arclength(x: ShellXml, nsteps: int, csv: str | None = None, **overrides) -> dict:
    """solvers/quasistatic_shell_XML.cpp's step loop on the device with the arc-length option list
    (ID 93, "Length" the arc length), and benchmark_Roof's step file at the reference points."""
    from .problem import ShellProblem
    opts = x.options.get(93, {})
    kw = alm_kwargs(opts)
    kw.update(overrides)
    length = kw.pop("length", opts.get("Length", 1e-2))
    sp = ShellProblem(x.patch)
    try:
        r = sp.run_alm(nsteps, length, keep_U=True, **kw)
    finally:
        sp.close()
    if csv is not None:
        write_step_csv(csv, x.patch, x.ref_points, r["unorm"], r["U"], r["lambda"], r["indicator"])
    r["points"] = [displacement_at(x.patch, u, x.ref_points) for u in r["U"]]
    return r
