"""Single-patch tensor B-spline shell geometries, boundary conditions and loads.

These builders produce the *input data* of the shell hot path: an undeformed
control net on a uniform open knot vector of degree ``p`` with ``n`` elements
per direction, the Dirichlet flags of the dofs and the loads.  They play the
role of the reference drivers' geometry set-up:

* unit plate, degree-elevated and refined  -- ``filedata/planar/unitplate.xml``,
  ``tutorials/nonlinear_shell_static.cpp``;
* shallow cylindrical roof (parabolic section through the arc's end points
  with the middle control point on the tangent intersection)
  -- ``filedata/surface/scordelis_lo_roof_shallow.xml:28-52``,
  ``benchmarks/benchmark_Roof.cpp:146-156`` (degree elevation + refinement);
* hinged straight edges / free curved edges and the central point load
  -- ``benchmarks/benchmark_Roof.cpp:198-236``;
* clamped edges as "boundary row fixed + second-row normal component fixed",
  the homogeneous form of gsKLShell's ``condition_type::clamped``
  (``benchmarks/benchmark_Roof.cpp:214-220``).

A polynomial geometry of degree <= p in each variable lies in the spline space,
so interpolation at the Greville abscissae reproduces it exactly (up to
rounding); that is how refined/elevated control nets are made here instead of
knot insertion + degree elevation, with the same result.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SIDES = ("west", "east", "south", "north")


def uniform_knots(p: int, n: int) -> np.ndarray:
    """Open uniform knot vector on [0, 1] with n elements."""
    return np.concatenate([np.zeros(p + 1), np.arange(1, n) / n, np.ones(p + 1)])


def greville(p: int, n: int) -> np.ndarray:
    """Greville abscissae of the n+p basis functions."""
    U = uniform_knots(p, n)
    m = n + p
    return np.array([U[i + 1:i + p + 1].mean() for i in range(m)])


def _basis_matrix(p: int, n: int, x: np.ndarray) -> np.ndarray:
    """Collocation matrix B[k, i] = N_i(x_k) (Cox-de Boor recursion)."""
    U = uniform_knots(p, n)
    m = n + p
    x = np.asarray(x, dtype=np.float64)
    # degree-0 functions on non-empty spans; x == 1 belongs to the last span
    B = np.zeros((x.size, len(U) - 1))
    for k, xv in enumerate(x):
        e = min(int(np.floor(xv * n)), n - 1)
        B[k, e + p] = 1.0
    for q in range(1, p + 1):
        Bn = np.zeros((x.size, len(U) - 1 - q))
        for i in range(len(U) - 1 - q):
            d1 = U[i + q] - U[i]
            d2 = U[i + q + 1] - U[i + 1]
            if d1 > 0:
                Bn[:, i] += (x - U[i]) / d1 * B[:, i]
            if d2 > 0:
                Bn[:, i] += (U[i + q + 1] - x) / d2 * B[:, i + 1]
        B = Bn
    return B[:, :m]


def interpolate_surface(p: int, nx: int, ny: int, fun) -> np.ndarray:
    """Control net (mx*my, 3) interpolating fun(xi, eta) -> (.., 3) at Greville points."""
    gx, gy = greville(p, nx), greville(p, ny)
    Bx, By = _basis_matrix(p, nx, gx), _basis_matrix(p, ny, gy)
    XI, ETA = np.meshgrid(gx, gy, indexing="xy")  # [j, i]
    F = np.asarray(fun(XI, ETA), dtype=np.float64)  # (my, mx, 3)
    C = np.empty_like(F)
    for c in range(3):
        # F = By C Bx^T
        C[:, :, c] = np.linalg.solve(By, np.linalg.solve(Bx, F[:, :, c].T).T)
    return C.reshape(-1, 3)  # cp = i + mx*j


@dataclass
class Patch:
    """Tensor B-spline shell patch plus material, BCs and loads."""

    p: int
    nx: int
    ny: int
    ctrl: np.ndarray                      # (mx*my, 3)
    t: float
    E: float
    nu: float
    rho: float = 1.0
    fixed: np.ndarray | None = None       # (3*mx*my,) int32
    point_xi: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    point_F: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    fsurf: tuple = (0.0, 0.0, 0.0)
    pressure: float = 0.0                 # follower pressure along a1 x a2 (gsThinShellAssembler::setPressure)

    @property
    def mx(self) -> int:
        return self.nx + self.p

    @property
    def my(self) -> int:
        return self.ny + self.p

    @property
    def ndof(self) -> int:
        return 3 * self.mx * self.my

    def __post_init__(self):
        self.ctrl = np.ascontiguousarray(self.ctrl, dtype=np.float64).reshape(-1, 3)
        assert self.ctrl.shape[0] == self.mx * self.my
        if self.fixed is None:
            self.fixed = np.zeros(self.ndof, dtype=np.int32)
        self.fixed = np.ascontiguousarray(self.fixed, dtype=np.int32)
        self.point_xi = np.ascontiguousarray(np.reshape(self.point_xi, (-1, 2)), dtype=np.float64)
        self.point_F = np.ascontiguousarray(np.reshape(self.point_F, (-1, 3)), dtype=np.float64)

    # -- boundary conditions -------------------------------------------------
    def _row(self, side: str, depth: int):
        mx, my = self.mx, self.my
        if side == "west":
            return [depth + mx * j for j in range(my)]
        if side == "east":
            return [mx - 1 - depth + mx * j for j in range(my)]
        if side == "south":
            return [i + mx * depth for i in range(mx)]
        if side == "north":
            return [i + mx * (my - 1 - depth) for i in range(mx)]
        raise ValueError(side)

    def fix(self, side: str, comps=(0, 1, 2)):
        """Homogeneous Dirichlet on the boundary row of ``side``."""
        for cp in self._row(side, 0):
            for c in comps:
                self.fixed[3 * cp + c] = 1
        return self

    def clamp(self, side: str, comps=(2,)):
        """Clamped: the second row copies the (zero) boundary row in ``comps``."""
        for cp in self._row(side, 1):
            for c in comps:
                self.fixed[3 * cp + c] = 1
        return self

    def add_point_load(self, xi, eta, F):
        self.point_xi = np.vstack([self.point_xi, [[xi, eta]]])
        self.point_F = np.vstack([self.point_F, [F]])
        return self

    def copy(self) -> "Patch":
        return Patch(self.p, self.nx, self.ny, self.ctrl.copy(), self.t, self.E, self.nu, self.rho,
                     self.fixed.copy(), self.point_xi.copy(), self.point_F.copy(), tuple(self.fsurf))

    # -- dof helpers -----------------------------------------------------------
    def free_map(self) -> np.ndarray:
        idx = -np.ones(self.ndof, dtype=np.int64)
        free = np.flatnonzero(self.fixed == 0)
        idx[free] = np.arange(free.size)
        return idx

    @property
    def nfree(self) -> int:
        return int(np.count_nonzero(self.fixed == 0))

    def to_full(self, u_free: np.ndarray) -> np.ndarray:
        u = np.zeros(self.ndof)
        u[self.fixed == 0] = u_free
        return u

    def to_free(self, v_full: np.ndarray) -> np.ndarray:
        return np.asarray(v_full)[self.fixed == 0]


# ---------------------------------------------------------------------------
# builders
# ---------------------------------------------------------------------------
def plate(p: int, n: int, L: float = 1.0, t: float = 0.01, E: float = 1e7, nu: float = 0.3,
          rho: float = 1.0, z0=None, ny: int | None = None) -> Patch:
    """Flat square plate [0,L]^2 (optionally with an initial deflection z0(x, y))."""
    ny = n if ny is None else ny

    def fun(xi, eta):
        x, y = L * xi, L * eta
        z = np.zeros_like(x) if z0 is None else z0(x, y)
        return np.stack([x, y, z], axis=-1)

    return Patch(p, n, ny, interpolate_surface(p, n, ny, fun), t, E, nu, rho)


def shallow_roof(p: int, n: int, R: float = 2.5, theta: float = 0.1, L: float = 0.5,
                 t: float = 0.0127, E: float = 3.1e3, nu: float = 0.3, rho: float = 1.0,
                 ny: int | None = None) -> Patch:
    """Shallow cylindrical panel: x along the straight axis (length L), eta along the arc.

    Section: quadratic through (0,0), (-2R sin th, 0) with the middle control point at
    the intersection of the arc's end tangents (-R sin th, R sin th tan th); unit
    weights, i.e. the parabolic form of the circular arc used by the reference's
    shallow Scordelis-Lo roof file.
    """
    ny = n if ny is None else ny
    s, h = R * np.sin(theta), R * np.sin(theta) * np.tan(theta)

    def fun(xi, eta):
        return np.stack([L * xi, -2.0 * s * eta, 2.0 * h * eta * (1.0 - eta)], axis=-1)

    return Patch(p, n, ny, interpolate_surface(p, n, ny, fun), t, E, nu, rho)


def clamped_plate_problem(p=2, n=16, q=1e2, L=1.0, t=0.01, E=1e7, nu=0.3, rho=1.0, z0=None):
    """Plate with all four edges clamped under a uniform dead surface load q (downwards)."""
    pt = plate(p, n, L=L, t=t, E=E, nu=nu, rho=rho, z0=z0)
    for side in SIDES:
        pt.fix(side)
        pt.clamp(side)
    pt.fsurf = (0.0, 0.0, -q)
    return pt


def simply_supported_plate_problem(p=3, n=8, q=1.0, L=1.0, t=0.01, E=1e7, nu=0.3, rho=1.0, z0=None):
    """Plate with all four edges hinged (boundary row fixed) under uniform dead load q."""
    pt = plate(p, n, L=L, t=t, E=E, nu=nu, rho=rho, z0=z0)
    for side in SIDES:
        pt.fix(side)
    pt.fsurf = (0.0, 0.0, -q)
    return pt


def rectangle(p: int, nx: int, ny: int, a: float, b: float, t: float, E: float, nu: float, rho: float = 1.0) -> Patch:
    """Flat rectangle [0, a] x [0, b] with nx x ny elements."""

    def fun(xi, eta):
        return np.stack([a * xi, b * eta, np.zeros_like(xi)], axis=-1)

    return Patch(p, nx, ny, interpolate_surface(p, nx, ny, fun), t, E, nu, rho)


def add_edge_load(pt: Patch, side: str, q) -> Patch:
    """Uniform dead line load q (force per unit undeformed length) on one straight edge, as point
    loads at the edge's Gauss points ((p + 1) per element, weights times the edge length): the
    consistent nodal forces of the line load, exactly."""
    gx, gw = np.polynomial.legendre.leggauss(pt.p + 1)
    ne = pt.ny if side in ("west", "east") else pt.nx
    x0 = pt.ctrl.reshape(pt.my, pt.mx, 3)
    length = np.linalg.norm(x0[-1, 0] - x0[0, 0]) if side in ("west", "east") else np.linalg.norm(x0[0, -1] - x0[0, 0])
    q = np.asarray(q, dtype=np.float64)
    for e in range(ne):
        for s, w in zip(gx, gw):
            t = (e + 0.5 * (s + 1.0)) / ne
            wt = 0.5 * w / ne * length
            xi, eta = {"west": (0.0, t), "east": (1.0, t), "south": (t, 0.0), "north": (t, 1.0)}[side]
            pt.add_point_load(xi, eta, wt * q)
    return pt


# the clamped strip of examples/gsThinShell_ArcLength.cpp:186-195,455-487 (test case 1): EI = 1
STRIP_T = 0.005313292845913 * 2


def compressed_strip_problem(p=3, nx=16, ny=1, a=1.0, b=0.1, t=STRIP_T, E=1e8, nu=0.0) -> Patch:
    """Strip a x b, both ends clamped, compressed by the dead line load E b t 1e-6 on the west
    edge (x free there), y held on every edge; perfect and flat, so the compressed state loses
    stability at a symmetric bifurcation, the clamped Euler load 4 pi^2 E I / a^2 (total force;
    E I = E b t^3 / 12 = 1 for the default sizes)."""
    pt = rectangle(p, nx, ny, a, b, t, E, nu)
    pt.fix("west", (1, 2))
    pt.clamp("west", (2,))
    pt.fix("east", (0, 1, 2))
    pt.clamp("east", (2,))
    pt.fix("south", (1,))
    pt.fix("north", (1,))
    add_edge_load(pt, "west", (E * b * t * 1e-6, 0.0, 0.0))
    return pt


def roof_problem(p=3, n=32, P=-10.0, **kw) -> Patch:
    """Hinged straight edges (south/north), free curved edges, central point load P in z."""
    pt = shallow_roof(p, n, **kw)
    pt.fix("south")
    pt.fix("north")
    pt.add_point_load(0.5, 0.5, (0.0, 0.0, P))
    return pt


def gaussian_random_field(points: np.ndarray, corr_len: float, rng: np.random.Generator,
                          jitter: float = 1e-10) -> np.ndarray:
    """Unit-variance squared-exponential GRF sampled at points (k, 2) via Cholesky."""
    d2 = ((points[:, None, :] - points[None, :, :]) ** 2).sum(-1)
    C = np.exp(-0.5 * d2 / corr_len ** 2) + jitter * np.eye(len(points))
    Lc = np.linalg.cholesky(C)
    return Lc @ rng.standard_normal(len(points))


# Arc length of the shallow-roof benchmark.  The reference runs the roof in millimetres with
# Scaling = 0 (a displacement-only arc length) and dL = 0.5 on its default mesh: degree 2 elevated
# once and refined once, 2 x 2 elements of degree 3, 45 free dofs (benchmarks/benchmark_Roof.cpp:
# 44-76, 365).  Here lengths are in metres, and the displacement norm over n free dofs of the same
# physical motion grows like sqrt(n): at one fixed arc length on the imperfect-roof batch the
# median step of the first negative pivot is 15, 24, 30, 36, 45 for 16, 24, 32, 40, 48 elements a
# side, in proportion to sqrt(n_free).  So the same physical step is 0.5e-3 * sqrt(n / 45);
# `kappa` scales it.
REF_ARC_LENGTH_M = 0.5e-3
REF_FREE_DOFS = 45


def roof_arc_length(nfree: int, kappa: float = 1.0) -> float:
    return kappa * REF_ARC_LENGTH_M * float(np.sqrt(nfree / REF_FREE_DOFS))


def imperfect_roof_instance(i: int, sizes=(16, 24, 32, 40, 48), p: int = 3, seed0: int = 20261202,
                            corr_len: float = 0.1, amp_frac: float = 0.5, **kw) -> Patch:
    """Instance i of the batched imperfection sweep on the shallow-roof panel family.

    Mesh n ~ U(sizes); z-perturbation of the control net = a * GRF(corr_len) with
    a ~ U[0, amp_frac * t]; generator PCG64(seed0 + i).
    """
    rng = np.random.Generator(np.random.PCG64(seed0 + i))
    n = int(rng.choice(sizes))
    amp = rng.uniform(0.0, amp_frac) * kw.get("t", 0.0127)
    pt = roof_problem(p=p, n=n, **kw)
    xy = pt.ctrl[:, :2]
    field_ = gaussian_random_field(xy, corr_len, rng)
    pt.ctrl[:, 2] += amp * field_
    return pt
