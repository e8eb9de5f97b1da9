"""Structured hexahedral box meshes and 1D tensor-product tabulation.

This is the PETSc-free stand-in for what HONEE gets from DMPlex and PetscFE:

* ``PetscFECreateLagrange(..., degree, q_order=degree+q_extra)`` (src/dm-utils.c:1071-1117) gives a tensor Lagrange
  basis on Gauss-Lobatto-Legendre nodes (PETSc's default Lagrange dual-space nodes) and a Gauss-Legendre rule with
  ``degree + 1 + q_extra`` points per direction on the reference interval [-1, 1].
* ``Create1DTabulation_Tensor`` / ``ComputeFieldTabulationP2C`` (src/dm-utils.c:594-729) hand libCEED the 1D
  interpolation ``B[q][p]`` and derivative ``D[q][p]`` with nodes in lexicographic (left-to-right) order.
* ``DMPlexCeedElemRestrictionCreate`` (src/dm-utils.c:83-121) gives int32 offsets ``[nelem][P^3]`` (x fastest) into an
  L-vector of ``[nnodes][5]`` interleaved components.
* The coordinate field uses a linear (P=2) basis over per-element ("localized", src/dm-utils.c:1138-1153) vertex
  coordinates, so periodic directions need no special casing.

Node numbering is lexicographic over the global node grid.  Everything here is setup, off the hot path.
"""
from __future__ import annotations

import dataclasses

import numpy as np


def gll_nodes(n: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre nodes on [-1, 1] (n >= 2), ascending."""
    if n == 2:
        return np.array([-1.0, 1.0])
    # interior nodes are the roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1.0
    interior = np.polynomial.legendre.Legendre(c).deriv().roots()
    x = np.concatenate(([-1.0], np.sort(interior.real), [1.0]))
    # polish with Newton on (1 - x^2) P'_{n-1}(x)
    leg = np.polynomial.legendre.Legendre(c)
    d1, d2 = leg.deriv(1), leg.deriv(2)
    for _ in range(3):
        xi = x[1:-1]
        x[1:-1] = xi - d1(xi) / d2(xi)
    return x


def gauss_rule(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre points and weights on [-1, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def lagrange_tabulate(nodes: np.ndarray, pts: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Interpolation B[q][p] = l_p(x_q) and derivative D[q][p] = l_p'(x_q) for the Lagrange basis on ``nodes``."""
    P, Q = len(nodes), len(pts)
    B = np.zeros((Q, P))
    D = np.zeros((Q, P))
    for p in range(P):
        others = [nodes[j] for j in range(P) if j != p]
        denom = np.prod([nodes[p] - o for o in others])
        for q in range(Q):
            x = pts[q]
            B[q, p] = np.prod([x - o for o in others]) / denom
            s = 0.0
            for k in range(len(others)):
                s += np.prod([x - others[j] for j in range(len(others)) if j != k])
            D[q, p] = s / denom
    return B, D


@dataclasses.dataclass
class Tabulation:
    P: int
    Q: int
    B: np.ndarray  # [Q][P]
    D: np.ndarray  # [Q][P]
    qref: np.ndarray  # [Q]
    qweight: np.ndarray  # [Q]
    Bx: np.ndarray  # [Q][2] linear coordinate basis
    Dx: np.ndarray


def tabulate(degree: int, q_extra: int = 0) -> Tabulation:
    P = degree + 1
    Q = degree + 1 + q_extra
    qref, qw = gauss_rule(Q)
    B, D = lagrange_tabulate(gll_nodes(P), qref)
    Bx, Dx = lagrange_tabulate(np.array([-1.0, 1.0]), qref)
    return Tabulation(P, Q, B, D, qref, qw, Bx, Dx)


@dataclasses.dataclass
class BoxMesh:
    """Hex box with ``shape=(nx, ny, nz)`` elements of polynomial degree ``degree``."""

    shape: tuple[int, int, int]
    degree: int
    periodic: tuple[bool, bool, bool]
    nnodes: int
    node_dims: tuple[int, int, int]
    offsets: np.ndarray  # int32 [nelem][P^3]
    xoffsets: np.ndarray  # int32 [nelem][8]
    xcoord: np.ndarray  # float64 [nelem*8][3]
    vertex_grid: np.ndarray  # float64 [nz+1][ny+1][nx+1][3]

    @property
    def nelem(self) -> int:
        return int(np.prod(self.shape))

    @property
    def P(self) -> int:
        return self.degree + 1

    def element_node_coords(self, tab: Tabulation | None = None, fast: bool = False) -> np.ndarray:
        """Physical coordinates of every element's nodes, [nelem][P^3][3] (trilinear map at GLL nodes).  fast: the
        contraction in pairwise steps (for large meshes; the sums run in another order, so values differ in the last
        bits)."""
        nodes = gll_nodes(self.P)
        Bn, _ = lagrange_tabulate(np.array([-1.0, 1.0]), nodes)  # [P][2]
        xe = self.xcoord.reshape(self.nelem, 2, 2, 2, 3)  # [e][z][y][x][c]
        out = np.einsum("ka,jb,ic,eabcd->ekjid", Bn, Bn, Bn, xe, optimize=fast)
        return out.reshape(self.nelem, self.P**3, 3)

    def multiplicity(self) -> np.ndarray:
        m = np.zeros(self.nnodes)
        np.add.at(m, self.offsets.ravel(), 1.0)
        return m


def box_mesh(
    shape,
    degree: int,
    lengths=(2 * np.pi, 2 * np.pi, 2 * np.pi),
    lower=(0.0, 0.0, 0.0),
    periodic=(True, True, True),
    stretch: tuple[float, float, float] = (1.0, 1.0, 1.0),
    symmetric_stretch: tuple[bool, bool, bool] = (False, False, False),
    perturb: float = 0.0,
    seed: int = 0,
) -> BoxMesh:
    """Build a box mesh.

    ``stretch[d]`` is a geometric growth ratio of element sizes along ``d``; with ``symmetric_stretch[d]`` the growth
    runs from both walls toward the centre (channel wall-normal clustering).  ``perturb`` moves interior vertices by
    up to that fraction of the local spacing (seeded), giving non-affine trilinear elements.
    """
    nx, ny, nz = (int(s) for s in shape)
    P = degree + 1
    ne = (nx, ny, nz)
    edges = []
    for d in range(3):
        n = ne[d]
        if stretch[d] == 1.0:
            h = np.ones(n)
        elif symmetric_stretch[d]:
            half = (n + 1) // 2
            g = stretch[d] ** np.arange(half)
            h = np.concatenate([g, g[::-1][n % 2:]]) if n % 2 == 0 else np.concatenate([g, g[::-1][1:]])
            h = h[:n]
        else:
            h = stretch[d] ** np.arange(n)
        x = np.concatenate([[0.0], np.cumsum(h)])
        x = lower[d] + lengths[d] * x / x[-1]
        edges.append(x)
    vz, vy, vx = np.meshgrid(edges[2], edges[1], edges[0], indexing="ij")
    vgrid = np.stack([vx, vy, vz], axis=-1)  # [nz+1][ny+1][nx+1][3]
    if perturb > 0:
        rng = np.random.default_rng(seed)
        hmin = min(np.min(np.diff(e)) for e in edges)
        delta = rng.uniform(-perturb * hmin, perturb * hmin, size=vgrid.shape)
        # keep boundary vertices on their planes; periodic images must move together
        for d, axis in zip(range(3), (2, 1, 0)):
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[axis] = 0
            sl_hi[axis] = -1
            if periodic[d]:
                delta[tuple(sl_hi)] = delta[tuple(sl_lo)]
            else:
                delta[tuple(sl_lo) + (d,)] = 0.0
                delta[tuple(sl_hi) + (d,)] = 0.0
        # a periodic face plane is shared: its normal coordinate must stay fixed too
        for d, axis in zip(range(3), (2, 1, 0)):
            if periodic[d]:
                for idx in (0, -1):
                    sl = [slice(None)] * 3
                    sl[axis] = idx
                    delta[tuple(sl) + (d,)] = 0.0
        vgrid = vgrid + delta

    ndims = tuple(ne[d] * degree + (0 if periodic[d] else 1) for d in range(3))
    Nx, Ny, Nz = ndims
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = (a.transpose(2, 1, 0).ravel() for a in (ex, ey, ez))  # element id = ex + nx*(ey + ny*ez)
    li = np.arange(P)
    k, j, i = np.meshgrid(li, li, li, indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()  # local node n = i + P*(j + P*k)
    gx = (ex[:, None] * degree + i[None, :]) % Nx
    gy = (ey[:, None] * degree + j[None, :]) % Ny
    gz = (ez[:, None] * degree + k[None, :]) % Nz
    offsets = (gx + Nx * (gy + Ny * gz)).astype(np.int32)

    # localized coordinates: 8 vertices per element, lexicographic (x fastest)
    xc = np.empty((nx * ny * nz, 2, 2, 2, 3))
    for c in range(2):
        for b in range(2):
            for a in range(2):
                xc[:, c, b, a, :] = vgrid[ez + c, ey + b, ex + a]
    xcoord = xc.reshape(-1, 3)
    xoffsets = np.arange(nx * ny * nz * 8, dtype=np.int32).reshape(-1, 8)
    return BoxMesh((nx, ny, nz), degree, tuple(bool(p) for p in periodic), Nx * Ny * Nz, ndims, offsets, xoffsets, xcoord, vgrid)


def boundary_faces(mesh: BoxMesh, face_id: int) -> np.ndarray:
    """Elements owning the boundary face ``face_id`` (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)."""
    nx, ny, nz = mesh.shape
    d, side = face_id // 2, face_id % 2
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    eid = ex + nx * (ey + ny * ez)
    idx = [ex, ey, ez][d]
    target = 0 if side == 0 else mesh.shape[d] - 1
    return np.sort(eid[idx == target].ravel()).astype(np.int32)


# Face-local node (i, j) -> element node (a, b, c) for face ids 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z, with the first face
# coordinate X1 along i and X2 along j chosen so that dx/dX1 x dx/dX2 points out of the element.
_FACE_ABC = {
    0: lambda i, j, L: (0, j, i),
    1: lambda i, j, L: (L, i, j),
    2: lambda i, j, L: (i, 0, j),
    3: lambda i, j, L: (j, L, i),
    4: lambda i, j, L: (j, i, 0),
    5: lambda i, j, L: (i, j, L),
}


def face_node_tables(mesh: BoxMesh, face_elem, face_id) -> tuple[np.ndarray, np.ndarray]:
    """Face node indices [nfaces][P^2] and coordinate node indices [nfaces][4] (face-local, X1 fastest)."""
    P = mesh.P
    face_elem = np.asarray(face_elem, dtype=np.int64)
    face_id = np.asarray(face_id, dtype=np.int64)
    nodes = np.zeros((len(face_elem), P * P), dtype=np.int32)
    xnodes = np.zeros((len(face_elem), 4), dtype=np.int32)
    for k, (e, f) in enumerate(zip(face_elem, face_id)):
        for j in range(P):
            for i in range(P):
                a, b, c = _FACE_ABC[int(f)](i, j, P - 1)
                nodes[k, i + P * j] = mesh.offsets[e, a + P * (b + P * c)]
        for j in range(2):
            for i in range(2):
                a, b, c = _FACE_ABC[int(f)](i, j, 1)
                xnodes[k, i + 2 * j] = mesh.xoffsets[e, a + 2 * (b + 2 * c)]
    return nodes, xnodes


def box_boundary(mesh: BoxMesh, sides) -> tuple[np.ndarray, np.ndarray]:
    """(face_elem, face_id) of all faces on the given box sides."""
    elems, ids = [], []
    for f in sides:
        e = boundary_faces(mesh, f)
        elems.append(e)
        ids.append(np.full(len(e), f, dtype=np.int32))
    return np.concatenate(elems).astype(np.int32), np.concatenate(ids).astype(np.int32)


def boundary_nodes(mesh: BoxMesh, sides) -> np.ndarray:
    """Sorted indices of the nodes on the given box sides (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)."""
    fe, fi = box_boundary(mesh, sides)
    nodes, _ = face_node_tables(mesh, fe, fi)
    return np.unique(nodes.ravel()).astype(np.int32)


def essential_dofs(mesh: BoxMesh, sides, comps) -> np.ndarray:
    """L-vector entries 5 node + c of the components comps on the nodes of the given sides (HONEE's essential_comps
    of a DM_BC_ESSENTIAL boundary, src/bc_definition.c:249-280)."""
    nodes = boundary_nodes(mesh, sides)
    return np.sort((5 * nodes[:, None] + np.asarray(comps, dtype=np.int64)[None, :]).ravel()).astype(np.int32)
