"""Quadrature rules and Lagrange simplex shape tables, as Fluidity's
element_type carries them (the caller side of flgpu_field_mesh_bind).

Restates femtools/Quadrature.F90 (template selection :304-321, expansion
:572-605, tet rules :649-735, tri rules :909-993, interval rules
:1140-1230) and
femtools/Shape_Functions.F90:128-366 with femtools/Elements.F90:615-717
(derivatives with respect to the first dim barycentric coordinates, the last
one dependent).  Layouts are Fortran's: n(loc, ngi), dn(loc, ngi, dim).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# permutation columns: per point, 1-based indices into the generator coords
_TET7 = [[1, 1, 1, 1]]
_TET8 = [[1, 1, 1, 2], [1, 1, 2, 1], [1, 2, 1, 1], [2, 1, 1, 1]]
_TET9 = [[1, 1, 2, 2], [1, 2, 1, 2], [1, 2, 2, 1], [2, 1, 1, 2], [2, 1, 2, 1], [2, 2, 1, 1]]
_TRI4 = [[1, 1, 1]]
_TRI5 = [[1, 1, 2], [1, 2, 1], [2, 1, 1]]
_INT1 = [[1, 1]]
_INT2 = [[1, 2], [2, 1]]


def _expand(perm, coords, weight, pts, wts):
    for col in perm:
        pts.append([coords[j - 1] if j else 0.0 for j in col])
        wts.append(weight)


def quadrature(dim: int, degree: int) -> tuple[np.ndarray, np.ndarray]:
    """Barycentric points l(ngi, dim+1) and weights w(ngi) of the lowest
    degree >= `degree` Fluidity rule for the simplex of dimension dim."""
    pts: list = []
    wts: list = []
    if dim == 3:
        if degree <= 1:
            _expand(_TET7, [0.25], 0.166666666666666666666666666666666, pts, wts)
        elif degree == 2:
            a = 0.138196601125010515179541316563436
            _expand(_TET8, [a, 1.0 - 3.0 * a], 0.041666666666666666666666666666666, pts, wts)
        elif degree == 3:
            _expand(_TET7, [0.25], -0.133333333333333333333333333333333, pts, wts)
            a = 0.166666666666666666666666666666666
            _expand(_TET8, [a, 1.0 - 3.0 * a], 0.075, pts, wts)
        elif degree == 4:
            _expand(_TET7, [0.25], -0.013155555555555555555555555555555, pts, wts)
            a = 0.0714285714285714285714285714285714
            _expand(_TET8, [a, 1.0 - 3.0 * a], 7.62222222222222222222222222222222e-3, pts, wts)
            a = 0.399403576166799204996102147461640
            _expand(_TET9, [a, 0.5 - a], 0.0248888888888888888888888888888888, pts, wts)
        else:
            raise ValueError(f"no tetrahedron rule of degree {degree}")
    elif dim == 2:
        if degree <= 1:
            _expand(_TRI4, [0.333333333333333333333333333333333], 0.5, pts, wts)
        elif degree == 2:
            a = 0.166666666666666666666666666666666
            _expand(_TRI5, [a, 1.0 - 2.0 * a], 0.166666666666666666666666666666666, pts, wts)
        elif degree == 3:
            _expand(_TRI4, [0.333333333333333333333333333333333], -0.28125, pts, wts)
            _expand(_TRI5, [0.2, 1.0 - 2.0 * 0.2], 0.260416666666666666666666666666666, pts, wts)
        elif degree == 4:
            a = 0.0915762135097707434595714634022015
            _expand(_TRI5, [a, 1.0 - 2.0 * a], 0.0549758718276609338191631624501052, pts, wts)
            a = 0.445948490915964886318329253883051
            _expand(_TRI5, [a, 1.0 - 2.0 * a], 0.111690794839005732847503504216561, pts, wts)
        else:
            raise ValueError(f"no triangle rule of degree {degree}")
    elif dim == 1:
        # interval rules (Quadrature.F90:1140-1230), the surface rules of 2-D meshes
        if degree <= 1:
            _expand(_INT1, [0.5], 1.0, pts, wts)
        elif degree == 2:
            a = 0.788675134594813
            _expand(_INT2, [a, 1.0 - a], 0.5, pts, wts)
        elif degree == 3:
            a = 0.887298334620742
            _expand(_INT2, [a, 1.0 - a], 0.277777777777777, pts, wts)
            _expand(_INT1, [0.5], 0.444444444444444, pts, wts)
        elif degree == 4:
            a = 0.9305681557970262
            _expand(_INT2, [a, 1.0 - a], 0.173927422568727, pts, wts)
            a = 0.6699905217924281
            _expand(_INT2, [a, 1.0 - a], 0.326072577431273, pts, wts)
        else:
            raise ValueError(f"no interval rule of degree {degree}")
    else:
        raise ValueError(f"dim {dim}")
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64)


def simplex_counts(dim: int, degree: int) -> np.ndarray:
    """number2count of the Lagrange simplex numbering: per local node the
    count of each barycentric coordinate (P2 order v1,e12,v2,e13,e23,v3,...)."""
    nv = dim + 1
    if degree == 1:
        return np.eye(nv, dtype=np.int64)
    if degree == 2:
        rows = []
        for b in range(nv):
            for a in range(b):
                c = [0] * nv
                c[a] = 1
                c[b] = 1
                rows.append(c)
            c = [0] * nv
            c[b] = 2
            rows.append(c)
        return np.array(rows, dtype=np.int64)
    raise ValueError(f"degree {degree}")


def _lag(c: int, dx: float, x: float) -> float:
    num = 1.0
    den = 1.0
    for i in range(c):
        num *= (x - i * dx)
        den *= (c * dx - i * dx)
    return num / den


def _dlag(c: int, dx: float, x: float) -> float:
    den = 1.0
    for i in range(c):
        den *= (c * dx - i * dx)
    s = 0.0
    for k in range(c):
        t = 1.0
        for i in range(c):
            if i != k:
                t *= (x - i * dx)
        s += t
    return s / den


def shape_tables(dim: int, degree: int, l: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """n(loc, ngi) and dn(loc, ngi, dim) at the barycentric points l."""
    counts = simplex_counts(dim, degree)
    loc, ngi, nv = counts.shape[0], l.shape[0], dim + 1
    dx = 1.0 / degree
    n = np.zeros((loc, ngi))
    dn = np.zeros((loc, ngi, dim))
    for i in range(loc):
        ci = [int(v) for v in counts[i]]
        for g in range(ngi):
            lg = [float(v) for v in l[g]]
            v = 1.0
            for j in range(nv):
                v *= _lag(ci[j], dx, lg[j])
            n[i, g] = v
            for d in range(dim):
                e = (_dlag(ci[d], dx, lg[d]) * _lag(ci[dim], dx, lg[dim])
                     + (-1.0) * _dlag(ci[dim], dx, lg[dim]) * _lag(ci[d], dx, lg[d]))
                for j in range(dim):
                    if j == d:
                        continue
                    e *= _lag(ci[j], dx, lg[j])
                dn[i, g, d] = e
    return n, dn


@dataclass
class Element:
    """The parts of Fluidity's element_type the C ABI takes."""
    dim: int
    degree: int
    quad_degree: int
    l: np.ndarray      # (ngi, dim+1) barycentric quadrature points
    w: np.ndarray      # (ngi,)
    n: np.ndarray      # (loc, ngi)
    dn: np.ndarray     # (loc, ngi, dim)

    @property
    def loc(self) -> int:
        return self.n.shape[0]

    @property
    def ngi(self) -> int:
        return self.n.shape[1]

    def fortran(self, a: np.ndarray) -> np.ndarray:
        """Column-major flat copy (what the Fortran arrays look like)."""
        return np.ascontiguousarray(a.ravel(order="F"))


def make_element(dim: int, degree: int, quad_degree: int) -> Element:
    l, w = quadrature(dim, quad_degree)
    n, dn = shape_tables(dim, degree, l)
    return Element(dim, degree, quad_degree, l, w, n, dn)


def make_face_element(dim: int, degree: int, quad_degree: int) -> Element:
    """face_shape of a dim-dimensional Lagrange simplex mesh: the element of
    dimension dim-1 and the same degree on the surface rule of the same
    degree as the volume rule; for an interval the face is a point: one
    node, one point of weight 1 (the detwei of a point facet)."""
    if dim == 1:
        return Element(0, degree, quad_degree, np.ones((1, 1)), np.ones(1), np.ones((1, 1)), np.zeros((1, 1, 0)))
    return make_element(dim - 1, degree, quad_degree)
