"""Synthetic meshes and fields of the benchmark configurations.

Everything is regenerated bit-identically from (configuration, seed) with a
portable splitmix64 counter RNG, so the GPU library, the tests and the CPU
baseline see the same inputs.  Meshes are Fluidity-shaped: 1-based,
element-major ndglno; coordinates dim x nodes column-major (stored here as
a C-ordered (nodes, dim) array, which is the same memory).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser of a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, ids: np.ndarray) -> np.ndarray:
    """Uniform [-1, 1) from splitmix64(seed ^ id)."""
    s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    z = splitmix64(np.asarray(ids, dtype=np.uint64) ^ s)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def uniform01(seed: int, ids: np.ndarray) -> np.ndarray:
    return 0.5 * (uniform_pm1(seed, ids) + 1.0)


def normal(seed: int, ids: np.ndarray) -> np.ndarray:
    """Standard normals by Box-Muller on two splitmix streams."""
    ids = np.asarray(ids, dtype=np.uint64)
    u1 = uniform01(seed, 2 * ids)
    u2 = uniform01(seed, 2 * ids + np.uint64(1))
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Mesh:
    dim: int
    nele: int
    # coordinate (linear) mesh
    nx: int
    ndglno_x: np.ndarray          # (nele*(dim+1),) int32, 1-based
    X: np.ndarray                 # (nx, dim) float64
    # field mesh (P1: same as the coordinate mesh)
    degree: int
    nnodes: int
    ndglno: np.ndarray            # (nele*loc,) int32, 1-based
    node_pos: np.ndarray          # (nnodes, dim) positions of the field nodes
    boundary: np.ndarray          # 1-based field nodes on the domain boundary
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def loc(self) -> int:
        return self.ndglno.size // self.nele

    @property
    def loc_x(self) -> int:
        return self.dim + 1


# Kuhn (Freudenthal) split of a hexahedron into 6 tetrahedra along the main
# diagonal: for each axis permutation (a, b, c) the path
# v000 -> +e_a -> +e_a+e_b -> v111.
_KUHN_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_corners():
    out = []
    for a, b, _ in _KUHN_PERMS:
        p0 = (0, 0, 0)
        p1 = [0, 0, 0]
        p1[a] = 1
        p2 = list(p1)
        p2[b] = 1
        out.append([p0, tuple(p1), tuple(p2), (1, 1, 1)])
    return np.array(out, dtype=np.int64)  # (6, 4, 3)


def cube_tets(n: int, jitter: float = 0.0, seed: int = 0, grading: float | None = None,
              degree: int = 1) -> Mesh:
    """Unit cube, n^3 hexahedra x 6 Kuhn tetrahedra.  Interior vertices are
    moved by jitter*h*U(-1,1) per coordinate; grading = size ratio of a tanh
    map of z concentrated at z = 0.5 (None: uniform).  degree 2 adds the
    edge nodes, numbered on the (2n+1)^3 half-spacing lattice."""
    m = n + 1
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    # node id = i + m*j + m*m*k (x fastest)
    ii, jj, kk = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    nx = m ** 3
    h = 1.0 / n
    xyz = np.stack([ii * h, jj * h, kk * h], axis=1).astype(np.float64)
    if jitter:
        interior = (ii > 0) & (ii < n) & (jj > 0) & (jj < n) & (kk > 0) & (kk < n)
        ids = np.arange(nx, dtype=np.uint64)
        for d in range(3):
            u = uniform_pm1(seed, ids * np.uint64(3) + np.uint64(d))
            xyz[:, d] = np.where(interior, xyz[:, d] + jitter * h * u, xyz[:, d])
    if grading is not None:
        s = np.arccosh(np.sqrt(grading))
        zeta = xyz[:, 2]
        xyz[:, 2] = 0.5 + 0.5 * np.tanh(s * (2.0 * zeta - 1.0)) / np.tanh(s)
    # elements
    hi, hj, hk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    hi, hj, hk = (a.transpose(2, 1, 0).ravel() for a in (hi, hj, hk))  # hex id = i + n j + n^2 k
    corners = _kuhn_corners()  # (6,4,3)
    nhex = n ** 3
    vi = hi[:, None, None] + corners[None, :, :, 0]
    vj = hj[:, None, None] + corners[None, :, :, 1]
    vk = hk[:, None, None] + corners[None, :, :, 2]
    conn_x = (vi + m * vj + m * m * vk).reshape(nhex * 6, 4)
    nele = nhex * 6
    ndglno_x = (conn_x + 1).astype(np.int32).ravel()
    bnd_x = np.nonzero((ii == 0) | (ii == n) | (jj == 0) | (jj == n) | (kk == 0) | (kk == n))[0] + 1
    if degree == 1:
        return Mesh(3, nele, nx, ndglno_x, xyz, 1, nx, ndglno_x, xyz, bnd_x.astype(np.int32),
                    name=f"cube{n}")
    # P2 on the half-spacing lattice
    mf = 2 * n + 1
    fi = 2 * vi
    fj = 2 * vj
    fk = 2 * vk
    fid = fi + mf * fj + mf * mf * fk  # (nhex, 6, 4) fine ids of the vertices
    fid = fid.reshape(nele, 4)
    F = np.stack([fi, fj, fk], axis=-1).reshape(nele, 4, 3)
    order = []  # v1,e12,v2,e13,e23,v3,e14,e24,e34,v4
    for b in range(4):
        for a in range(b):
            order.append((a, b))
        order.append((b, b))
    cols = []
    for a, b in order:
        g = (F[:, a, :] + F[:, b, :]) // 2
        cols.append(g[:, 0] + mf * g[:, 1] + mf * mf * g[:, 2])
    conn = np.stack(cols, axis=1)
    nnodes = mf ** 3
    fi3, fj3, fk3 = np.meshgrid(np.arange(mf), np.arange(mf), np.arange(mf), indexing="ij")
    fi3, fj3, fk3 = (a.transpose(2, 1, 0).ravel() for a in (fi3, fj3, fk3))
    # field-node positions: midpoints of the (jittered) coordinate edges
    pos = np.zeros((nnodes, 3))
    for q, (a, b) in enumerate(order):
        pos[conn[:, q]] = 0.5 * (xyz[conn_x[:, a]] + xyz[conn_x[:, b]])
    bnd = np.nonzero((fi3 == 0) | (fi3 == 2 * n) | (fj3 == 0) | (fj3 == 2 * n) | (fk3 == 0) | (fk3 == 2 * n))[0] + 1
    return Mesh(3, nele, nx, ndglno_x, xyz, 2, nnodes, (conn + 1).astype(np.int32).ravel(), pos,
                bnd.astype(np.int32), name=f"cube{n}p2")


def square_tris(n: int, jitter: float = 0.0, seed: int = 0, degree: int = 1) -> Mesh:
    """Unit square, n x n cells each split by the (i,j)-(i+1,j+1) diagonal
    into two P1 triangles; interior vertices jittered by jitter*h.  degree 2
    adds the edge nodes, numbered on the (2n+1)^2 half-spacing lattice, in
    the element order v1, e12, v2, e13, e23, v3."""
    m = n + 1
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    ii, jj = i.T.ravel(), j.T.ravel()  # node id = i + m j
    h = 1.0 / n
    xy = np.stack([ii * h, jj * h], axis=1).astype(np.float64)
    if jitter:
        interior = (ii > 0) & (ii < n) & (jj > 0) & (jj < n)
        ids = np.arange(m * m, dtype=np.uint64)
        for d in range(2):
            u = uniform_pm1(seed, ids * np.uint64(2) + np.uint64(d))
            xy[:, d] = np.where(interior, xy[:, d] + jitter * h * u, xy[:, d])
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ci, cj = ci.T.ravel(), cj.T.ravel()  # cell id = i + n j
    v00 = ci + m * cj
    v10 = v00 + 1
    v01 = v00 + m
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    conn = np.stack([t1, t2], axis=1).reshape(-1, 3)
    nd = (conn + 1).astype(np.int32).ravel()
    bnd = np.nonzero((ii == 0) | (ii == n) | (jj == 0) | (jj == n))[0] + 1
    if degree == 1:
        return Mesh(2, conn.shape[0], m * m, nd, xy, 1, m * m, nd, xy, bnd.astype(np.int32), name=f"square{n}")
    mf = 2 * n + 1
    F = np.stack([2 * (conn % m), 2 * (conn // m)], axis=-1)  # (nele, 3, 2) half-lattice coordinates
    order = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]
    cols = []
    for a, b in order:
        g = (F[:, a, :] + F[:, b, :]) // 2
        cols.append(g[:, 0] + mf * g[:, 1])
    conn2 = np.stack(cols, axis=1)
    nnodes = mf * mf
    pos = np.zeros((nnodes, 2))
    for q, (a, b) in enumerate(order):
        pos[conn2[:, q]] = 0.5 * (xy[conn[:, a]] + xy[conn[:, b]])
    fi, fj = np.arange(nnodes) % mf, np.arange(nnodes) // mf
    bnd2 = np.nonzero((fi == 0) | (fi == 2 * n) | (fj == 0) | (fj == 2 * n))[0] + 1
    return Mesh(2, conn.shape[0], m * m, nd, xy, 2, nnodes, (conn2 + 1).astype(np.int32).ravel(), pos,
                bnd2.astype(np.int32), name=f"square{n}p2")


def interval_mesh(n: int, left: float = 0.0, right: float = 1.0, jitter: float = 0.0, seed: int = 0,
                  degree: int = 1) -> Mesh:
    """The interval [left, right] in n elements (what Fluidity's `interval`
    script writes: vertices left + k dx, element k = (k, k+1)); interior
    vertices jittered by jitter*dx.  degree 2 adds the midpoints, numbered on
    the half-spacing lattice (vertex k = 2k, midpoint of element k = 2k+1),
    in the element order v1, mid, v2.  Boundary: the two end nodes (surface
    ids 1 at left, 2 at right in the interval script)."""
    h = (right - left) / n
    x = left + np.arange(n + 1) * h
    if jitter:
        u = uniform_pm1(seed, np.arange(n + 1, dtype=np.uint64))
        x[1:n] = x[1:n] + jitter * h * u[1:n]
    X = x[:, None].astype(np.float64)
    conn = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1)
    nd = (conn + 1).astype(np.int32).ravel()
    if degree == 1:
        return Mesh(1, n, n + 1, nd, X, 1, n + 1, nd, X, np.array([1, n + 1], np.int32), name=f"interval{n}")
    conn2 = np.stack([2 * conn[:, 0], 2 * conn[:, 0] + 1, 2 * conn[:, 1]], axis=1)
    pos = np.zeros((2 * n + 1, 1))
    pos[0::2, 0] = x
    pos[1::2, 0] = 0.5 * (x[:-1] + x[1:])
    return Mesh(1, n, n + 1, nd, X, 2, 2 * n + 1, (conn2 + 1).astype(np.int32).ravel(), pos,
                np.array([1, 2 * n + 1], np.int32), name=f"interval{n}p2")


def boundary_numbering(dim: int, degree: int, face: int) -> list[int]:
    """Local nodes (0-based) of face `face` (1-based; the face opposite
    vertex `face`) of a Lagrange simplex, in the order of
    femtools/Element_Numbering.F90:1564-1615: the element's nodes whose
    count of barycentric coordinate `face` is zero, in element order,
    reversed for even faces (in 3-D, row by row) so that every face is
    positively oriented with respect to its element."""
    from .elements import simplex_counts
    counts = simplex_counts(dim, degree)
    num = [i for i in range(counts.shape[0]) if counts[i, face - 1] == 0]
    if face % 2 == 0:
        if dim == 2:
            num = num[::-1]
        elif dim == 3:
            l, i = 1, len(num)
            while True:
                num[i - 1:i + l - 1] = num[i - 1:i + l - 1][::-1]
                l += 1
                i -= l
                if i < 1:
                    break
    return num


@dataclass
class Surface:
    """The surface elements of a mesh (its boundary facets), as the face
    loop of assemble_advection_diffusion_cg sees them: per face its T nodes
    (face_global_nodes(t, face)), its coordinate nodes
    (face_global_nodes(positions, face)) and its element (face_ele)."""
    nfaces: int
    floc: int
    face_nodes: np.ndarray     # (nfaces*floc,) int32, 1-based field nodes
    face_nodes_x: np.ndarray   # (nfaces*dim,) int32, 1-based coordinate nodes
    face_ele: np.ndarray       # (nfaces,) int32, 1-based
    centroid: np.ndarray       # (nfaces, dim) face centroids (for tagging)


def surface_of(mesh: Mesh) -> Surface:
    """Boundary facets (facets of exactly one element), ordered by element
    and then by local face number."""
    d = mesh.dim
    nv = d + 1
    cx = mesh.ndglno_x.reshape(mesh.nele, nv).astype(np.int64) - 1
    keys = []
    for f in range(nv):
        verts = np.sort(np.delete(cx, f, axis=1), axis=1)
        k = np.zeros(mesh.nele, np.int64)
        for c in range(d):
            k = k * np.int64(mesh.nx) + verts[:, c]
        keys.append(k)
    allk = np.stack(keys, axis=1).ravel()  # element-major, local face minor
    _, inv, cnt = np.unique(allk, return_inverse=True, return_counts=True)
    bnd = np.nonzero(cnt[inv] == 1)[0]
    ele, lf = bnd // nv, bnd % nv
    cf = mesh.ndglno.reshape(mesh.nele, mesh.loc)
    xnum = [boundary_numbering(d, 1, f + 1) for f in range(nv)]
    fnum = [boundary_numbering(d, mesh.degree, f + 1) for f in range(nv)]
    fx = np.stack([cx[ele, np.array(xnum)[lf, c]] for c in range(d)], axis=1) + 1
    fl = len(fnum[0])
    fn = np.stack([cf[ele, np.array(fnum)[lf, c]] for c in range(fl)], axis=1)
    cen = mesh.X[fx - 1].mean(axis=1)
    return Surface(int(bnd.size), fl, fn.astype(np.int32).ravel(), fx.astype(np.int32).ravel(),
                   (ele + 1).astype(np.int32), cen)


def surface_from_facets(mesh: Mesh, facets) -> Surface:
    """The surface elements a mesh file lists (vertex lists, 1-based), in
    file order, each matched to the element it bounds and numbered as that
    element's face (the face opposite the element vertex not on it; nodes in
    the element's boundary numbering), as Fluidity's surface mesh is."""
    d = mesh.dim
    nv = d + 1
    cx = mesh.ndglno_x.reshape(mesh.nele, nv)
    owner = {}
    for e in range(mesh.nele):
        for f in range(nv):
            owner.setdefault(tuple(sorted(int(v) for j, v in enumerate(cx[e]) if j != f)), []).append((e, f))
    cf = mesh.ndglno.reshape(mesh.nele, mesh.loc)
    fnum = [boundary_numbering(d, mesh.degree, f + 1) for f in range(nv)]
    xnum = [boundary_numbering(d, 1, f + 1) for f in range(nv)]
    fn, fx, fe = [], [], []
    for fac in facets:
        own = owner.get(tuple(sorted(int(v) for v in fac)))
        if not own or len(own) != 1:
            raise ValueError(f"surface element {fac} does not bound exactly one element")
        e, f = own[0]
        fn.append([cf[e, q] for q in fnum[f]])
        fx.append([cx[e, q] for q in xnum[f]])
        fe.append(e + 1)
    fx = np.array(fx, np.int32)
    return Surface(len(facets), len(fnum[0]), np.array(fn, np.int32).ravel(), fx.ravel(),
                   np.array(fe, np.int32), mesh.X[fx - 1].mean(axis=1))


BC_NONE, BC_NEUMANN, BC_WEAKDIRICHLET, BC_INTERNAL, BC_ROBIN = 0, 1, 2, 3, 4


def unit_box_bcs(mesh: Mesh, surf: Surface, kind: str = "neumann_robin"):
    """Boundary conditions on a box-shaped domain by side (x = 0 and x = 1
    stand for the smallest and largest x of the mesh).
    "neumann_robin": x = 0 Neumann with flux g = 1 + y/2, x = 1 Robin with
    g = 0.5 and coefficient h = 2 + y, every other side strong Dirichlet 0.
    "weak": x = 0 weak Dirichlet with value 1 - y, every other side none.
    Returns (bc_type, bc_value, bc_value2, dirichlet_nodes): per face its
    type, the bc fields at its nodes (nfaces*floc), and the nodes of the
    strong-Dirichlet sides (1-based)."""
    tol = 1e-12
    cen = surf.centroid
    xmin, xmax = mesh.X[:, 0].min(), mesh.X[:, 0].max()
    x0 = np.abs(cen[:, 0] - xmin) < tol * max(1.0, abs(xmin))
    x1 = np.abs(cen[:, 0] - xmax) < tol * max(1.0, abs(xmax))
    t = np.zeros(surf.nfaces, np.int32)
    pos = mesh.node_pos[surf.face_nodes.reshape(surf.nfaces, surf.floc) - 1]  # (nf, floc, dim)
    y = pos[:, :, 1] if mesh.dim > 1 else np.zeros(pos.shape[:2])
    v1 = np.zeros_like(y)
    v2 = np.zeros_like(y)
    if kind == "neumann_robin":
        t[x0] = BC_NEUMANN
        t[x1] = BC_ROBIN
        v1[x0] = 1.0 + 0.5 * y[x0]
        v1[x1] = 0.5
        v2[x1] = 2.0 + y[x1]
        strong = t == BC_NONE
    elif kind == "weak":
        t[x0] = BC_WEAKDIRICHLET
        v1[x0] = 1.0 - y[x0]
        strong = np.zeros(surf.nfaces, bool)
    else:
        raise ValueError(kind)
    dn = np.unique(surf.face_nodes.reshape(surf.nfaces, surf.floc)[strong])
    return t, np.ascontiguousarray(v1.ravel()), np.ascontiguousarray(v2.ravel()), dn.astype(np.int32)


def read_gmsh(path: str) -> Mesh:
    """P1 mesh from a gmsh 2.2 ASCII file (the format of Fluidity's
    read_mesh_files(..., format="gmsh")): triangles (type 2) in 2D or
    tetrahedra (type 4) in 3D, node numbers compacted to 1..N in file order,
    element vertex order kept.  Boundary nodes are those of facets that
    belong to one element only.  A 2D mesh is detected by z = 0 everywhere."""
    with open(path) as f:
        lines = f.read().split("\n")
    k = lines.index("$Nodes")
    nn = int(lines[k + 1])
    ids = np.empty(nn, np.int64)
    xyz = np.empty((nn, 3))
    for q in range(nn):
        t = lines[k + 2 + q].split()
        ids[q] = int(t[0])
        xyz[q] = [float(v) for v in t[1:4]]
    k = lines.index("$Elements")
    ne = int(lines[k + 1])
    lines_1d, tris, tets = [], [], []
    for q in range(ne):
        t = [int(v) for v in lines[k + 2 + q].split()]
        etype, ntags = t[1], t[2]
        nodes = t[3 + ntags:]
        if etype == 1:
            lines_1d.append(nodes)
        elif etype == 2:
            tris.append(nodes)
        elif etype == 4:
            tets.append(nodes)
    dim = 3 if tets else 2
    cells = np.array(tets if tets else tris, np.int64)
    remap = {int(v): q + 1 for q, v in enumerate(ids)}
    conn = np.vectorize(remap.__getitem__)(cells).astype(np.int32)
    # the file's surface elements (triangles of a tet mesh, lines of a
    # triangle mesh), in file order, 1-based compacted node numbers
    sf = tris if tets else lines_1d
    facets = [[remap[int(v)] for v in f] for f in sf]
    X = np.ascontiguousarray(xyz[:, :dim])
    loc = dim + 1
    faces = {}
    for e in range(conn.shape[0]):
        for skip in range(loc):
            fkey = tuple(sorted(int(conn[e, a]) for a in range(loc) if a != skip))
            faces[fkey] = faces.get(fkey, 0) + 1
    bnd = sorted({v for fk, cnt in faces.items() if cnt == 1 for v in fk})
    nd = conn.ravel()
    name = os.path.splitext(os.path.basename(path))[0]
    return Mesh(dim, conn.shape[0], nn, nd, X, 1, nn, nd, X, np.array(bnd, np.int32), name=name,
                meta={"surface_facets": facets})


def grf_lognormal(pos: np.ndarray, seed: int, corr_len: float = 0.1, contrast: float = 1e4,
                  nfeat: int = 64) -> np.ndarray:
    """exp of a Gaussian random field (random Fourier features, correlation
    length corr_len), scaled so that max/min over the nodes = contrast."""
    dim = pos.shape[1]
    fid = np.arange(nfeat * dim, dtype=np.uint64)
    omega = normal(seed, fid).reshape(nfeat, dim) / corr_len
    phi = 2.0 * np.pi * uniform01(seed ^ 0x5DEECE66D, np.arange(nfeat, dtype=np.uint64))
    g = np.zeros(pos.shape[0])
    for f in range(nfeat):
        g += np.cos(pos @ omega[f] + phi[f])
    g *= np.sqrt(2.0 / nfeat)
    g = (g - g.mean()) / max(g.std(), 1e-300)
    span = g.max() - g.min()
    sigma = np.log(contrast) / span if span > 0 else 0.0
    return np.exp(sigma * (g - g.min()))


def abc_flow(pos: np.ndarray, A: float = 1.0, B: float = 1.0, Cc: float = 1.0) -> np.ndarray:
    """Arnold-Beltrami-Childress velocity on the unit cube (2 pi periodic
    scaled to [0,1]); divergence free."""
    x, y, z = (2.0 * np.pi * pos[:, d] for d in range(3))
    return np.stack([A * np.sin(z) + Cc * np.cos(y),
                     B * np.sin(x) + A * np.cos(z),
                     Cc * np.sin(y) + B * np.cos(x)], axis=1)


def sin_source(pos: np.ndarray) -> np.ndarray:
    out = np.ones(pos.shape[0])
    for d in range(pos.shape[1]):
        out *= np.sin(2.0 * np.pi * pos[:, d])
    return out


@dataclass
class Config:
    """A benchmark configuration: mesh, fields and equation options."""
    name: str
    mesh: Mesh
    quad_degree: int
    T: np.ndarray
    U: np.ndarray | None
    kappa: np.ndarray | float | None     # nodal (N,) iso, or constant
    source: np.ndarray | None
    absorption: np.ndarray | None
    dt: float
    theta: float
    beta: float
    have_mass: bool
    have_advection: bool
    solver: str                           # "cg" or "gmres"
    rtol: float


def config(name: str, seed: int = 1, scale: int | None = None) -> Config:
    """Configurations of the benchmark (scale overrides the cells per side
    for quick tests)."""
    if name == "c1":
        n = scale or 64
        m = square_tris(n, jitter=0.2, seed=seed)
        p = m.node_pos
        T = np.exp(-((p[:, 0] - 0.3) ** 2 + (p[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))
        U = np.stack([-(p[:, 1] - 0.5), p[:, 0] - 0.5], axis=1)
        return Config("c1", m, 2, T, U, 1e-3, None, None, 0.01, 0.5, 0.0, True, True, "gmres", 1e-10)
    if name == "c1p2":
        # the 2D configuration on P2 triangles (straight-sided, 3-point rule)
        n = scale or 32
        m = square_tris(n, jitter=0.2, seed=seed, degree=2)
        p = m.node_pos
        T = np.exp(-((p[:, 0] - 0.3) ** 2 + (p[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))
        U = np.stack([-(p[:, 1] - 0.5), p[:, 0] - 0.5], axis=1)
        return Config("c1p2", m, 2, T, U, 1e-3, np.cos(3 * p[:, 0]), None, 0.01, 0.5, 0.0, True, True, "gmres",
                      1e-10)
    if name == "c2":
        n = scale or 64
        m = cube_tets(n, jitter=0.15, seed=seed)
        return Config("c2", m, 2, np.zeros(m.nnodes), None, 1.0, sin_source(m.node_pos), None,
                      1.0, 1.0, 0.0, False, False, "cg", 1e-8)
    if name == "c3":
        n = scale or 40
        m = cube_tets(n, jitter=0.0, seed=seed, degree=2)
        p = m.node_pos
        T = np.sin(np.pi * p[:, 0]) * np.sin(np.pi * p[:, 1]) * np.sin(np.pi * p[:, 2])
        return Config("c3", m, 4, T, abc_flow(p), 1e-3, None, None, 0.005, 0.5, 0.0, True, True,
                      "gmres", 1e-8)
    if name == "c4":
        n = scale or 40
        m = cube_tets(n, jitter=0.15, seed=seed)
        kap = grf_lognormal(m.node_pos, seed, 0.1, 1e4)
        return Config("c4", m, 2, np.zeros(m.nnodes), None, kap, sin_source(m.node_pos), None,
                      1.0, 1.0, 0.0, False, False, "cg", 1e-8)
    if name == "c5":
        n = scale or 110
        m = cube_tets(n, jitter=0.15, seed=seed, grading=100.0)
        kap = grf_lognormal(m.node_pos, seed, 0.1, 1e3)
        return Config("c5", m, 2, np.zeros(m.nnodes), None, kap, sin_source(m.node_pos), None,
                      1.0, 1.0, 0.0, False, False, "cg", 1e-8)
    if name in ("interval", "intervalp2"):
        # 1-D advection-diffusion with every term on a jittered interval
        # (P1, or P2 on a 3-point rule): a travelling bump, variable
        # velocity and diffusivity, a source
        deg = 2 if name == "intervalp2" else 1
        n = scale or (4096 if deg == 1 else 2048)
        m = interval_mesh(n, 0.0, 1.0, jitter=0.3, seed=seed, degree=deg)
        p = m.node_pos
        T = np.exp(-((p[:, 0] - 0.3) ** 2) / (2 * 0.05 ** 2))
        U = (0.5 + 0.3 * np.sin(2 * np.pi * p[:, 0]))[:, None]
        kap = 1e-3 * (1.0 + 0.5 * np.cos(3 * p[:, 0]))
        return Config(name, m, 2 if deg == 1 else 3, T, U, kap, np.sin(3 * p[:, 0]), None, 0.01, 0.5, 0.0, True,
                      True, "gmres", 1e-10)
    if name in ("dsd1d", "dsd1dp2"):
        # tests/diff_src_dirichlet_1d: -T'' = 2.13 on [0, 10] (dx = 1 in the
        # reference), T(10) = 2 (surface id 2), zero flux at x = 0, no mass
        # or advection terms, theta = 1, dt = 5; the P1 (and P2) solution is
        # nodally exact: T = 2.13/2 (100 - x^2) + 2
        deg = 2 if name == "dsd1dp2" else 1
        n = scale or 10
        m = interval_mesh(n, 0.0, 10.0, degree=deg)
        m.boundary = np.array([m.nnodes if deg == 1 else 2 * n + 1], np.int32)
        return Config(name, m, 2 if deg == 1 else 3, np.zeros(m.nnodes), None, 1.0, np.full(m.nnodes, 2.13), None,
                      5.0, 1.0, 0.0, False, False, "gmres", 1e-10)
    if name == "cube1":
        # Fluidity's small tetrahedral test mesh (tests/data/cube.1.msh: the
        # cube [-3, 3]^3 in 12 tetrahedra around one interior node), with
        # advection-diffusion on every term
        here = os.path.dirname(os.path.abspath(__file__))
        m = read_gmsh(os.path.join(os.path.dirname(here), "tests", "golden", "cube.1.msh"))
        p = m.node_pos
        T = np.cos(0.3 * p[:, 0]) + 0.2 * p[:, 1]
        U = np.tile(np.array([0.3, -0.2, 0.1]), (m.nnodes, 1))
        return Config("cube1", m, 2, T, U, 0.5, np.sin(0.5 * p[:, 2]), None, 0.1, 0.5, 0.0, True, True, "gmres",
                      1e-10)
    if name == "lapgrid":
        # an unstructured triangle mesh of the unit square from Fluidity's
        # tests (tests/data/laplacian_grid.1.msh), rotating bump
        here = os.path.dirname(os.path.abspath(__file__))
        m = read_gmsh(os.path.join(os.path.dirname(here), "tests", "golden", "laplacian_grid.1.msh"))
        p = m.node_pos
        T = np.exp(-((p[:, 0] - 0.3) ** 2 + (p[:, 1] - 0.5) ** 2) / (2 * 0.1 ** 2))
        U = np.stack([-(p[:, 1] - 0.5), p[:, 0] - 0.5], axis=1)
        return Config("lapgrid", m, 2, T, U, 1e-2, None, None, 0.01, 0.5, 0.0, True, True, "gmres", 1e-10)
    if name == "cavity":
        # the square-cavity mesh of Fluidity's colouring test
        # (tests/data/square-cavity-2d.msh), with the rotating-bump
        # advection-diffusion of the 2D configuration
        here = os.path.dirname(os.path.abspath(__file__))
        m = read_gmsh(os.path.join(os.path.dirname(here), "tests", "golden", "square-cavity-2d.msh"))
        p = m.node_pos
        T = np.exp(-((p[:, 0] - 0.3) ** 2 + (p[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))
        U = np.stack([-(p[:, 1] - 0.5), p[:, 0] - 0.5], axis=1)
        return Config("cavity", m, 2, T, U, 1e-3, None, None, 0.01, 0.5, 0.0, True, True, "gmres", 1e-10)
    raise ValueError(name)
