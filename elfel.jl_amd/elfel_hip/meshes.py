"""Meshes at the boundary of the assembly path (stand-in for MeshSteward/MeshCore).

A Mesh holds what the assembly path consumes from MeshCore: the base
incidence relation (connectivity, element-major, 1-based Int64) and the
"geom" attribute (node coordinates).  Generators:

* T3block / Q4block: nodes x-fastest; cells x-index outer, y-index inner;
  T3 diagonal (i,j)-(i+1,j+1).  This is the order of the reference's fixture
  test/qmesh-conn.dat and of the element lists in test/test_feiterators.jl:21-32.
* T6block: the T3 block's vertices first, then midside nodes numbered by
  first appearance (element order, local edges 1-2, 2-3, 3-1), midpoint
  coordinates (a + b) / 2.
* T6toT3: the corner nodes of a T6 mesh, renumbered compactly in increasing
  order (the pressure mesh of test/test_stokes.jl:32,37).
* load: the reference's .mesh manifest (xyz file, shape name, conn file;
  test/qmesh.mesh:1-3).
* L2block: nodes L i / n, segments (i, i+1) (the interval of
  examples/klein_gordon/klein_gordon.jl:29).
* T4block: nodes x-fastest, then y, then z; each cell (x-index fastest, then
  y, then z) split into six positively oriented tetrahedra around its
  diagonal (i, j, k)-(i+1, j+1, k+1) (the cube of
  examples/heat/poisson/t4.jl:23).  MeshSteward's own L2block / T4block are
  not in the reference tree: these orders are this package's, shared with
  the oracle.
"""
import os

import numpy as np


class Mesh:
    def __init__(self, conn, xy, shape):
        self.conn = np.ascontiguousarray(conn, dtype=np.int64)
        self.xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.shape = shape

    @property
    def nelements(self):
        return self.conn.shape[0]

    @property
    def nnodes(self):
        return self.xy.shape[0]

    def geometry(self):
        return self.xy

    def boundary_nodes(self):
        """connectedv(boundary(mesh)): nodes on edges that belong to one element."""
        return boundary_nodes(self)

    @property
    def sdim(self):
        return self.xy.shape[1]

    def vselect(self, box, inflate=0.0):
        """Nodes inside box = [xmin, xmax, ymin, ymax] enlarged by inflate (MeshSteward.vselect)."""
        x, y = self.xy[:, 0], self.xy[:, 1]
        m = ((x >= box[0] - inflate) & (x <= box[1] + inflate) & (y >= box[2] - inflate) & (y <= box[3] + inflate))
        return np.nonzero(m)[0] + 1

    def nearestto(self, p):
        d = np.hypot(self.xy[:, 0] - p[0], self.xy[:, 1] - p[1])
        return int(np.argmin(d)) + 1


def _lin(L, n):
    return L * np.arange(n + 1, dtype=np.float64) / float(n)


def _block_nodes(A, B, nA, nB):
    xs, ys = _lin(A, nA), _lin(B, nB)
    X, Y = np.meshgrid(xs, ys)  # rows = y, cols = x: x fastest when raveled
    return np.stack([X.ravel(), Y.ravel()], 1)


def _cell_origins(nA, nB):
    i, j = np.meshgrid(np.arange(nA), np.arange(nB), indexing="ij")  # i outer, j inner
    return (j.ravel() * (nA + 1) + i.ravel() + 1).astype(np.int64)


def T3block(A, B, nA, nB):
    f = _cell_origins(nA, nB)
    c1 = np.stack([f, f + 1, f + (nA + 1) + 1], 1)
    c2 = np.stack([f, f + (nA + 1) + 1, f + (nA + 1)], 1)
    conn = np.stack([c1, c2], 1).reshape(-1, 3)
    return Mesh(conn, _block_nodes(A, B, nA, nB), "T3")


def Q4block(A, B, nA, nB):
    f = _cell_origins(nA, nB)
    conn = np.stack([f, f + 1, f + (nA + 1) + 1, f + (nA + 1)], 1)
    return Mesh(conn, _block_nodes(A, B, nA, nB), "Q4")


def T3toT6(mesh):
    c3 = mesh.conn
    E = c3.shape[0]
    Nv = mesh.nnodes
    a = c3[:, [0, 1, 2]].ravel()
    b = c3[:, [1, 2, 0]].ravel()
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (Nv + 1) + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # number edges by first appearance in the (element, local edge) order
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    mid = (Nv + 1 + rank[inv]).reshape(E, 3)
    ea = (uniq[order] // (Nv + 1)) - 1
    eb = (uniq[order] % (Nv + 1)) - 1
    # coordinates: (locs[a] + locs[b]) / 2 with a, b in first-appearance orientation
    fa, fb = a[first[order]] - 1, b[first[order]] - 1
    del ea, eb
    xy = np.concatenate([mesh.xy, (mesh.xy[fa] + mesh.xy[fb]) / 2], 0)
    conn = np.concatenate([c3, mid], 1)
    return Mesh(conn, xy, "T6")


def T6block(A, B, nA, nB):
    return T3toT6(T3block(A, B, nA, nB))


def T6toT3(mesh):
    corners = np.unique(mesh.conn[:, :3].ravel())
    newnum = np.zeros(mesh.nnodes + 1, np.int64)
    newnum[corners] = np.arange(1, corners.size + 1)
    return Mesh(newnum[mesh.conn[:, :3]], mesh.xy[corners - 1], "T3")


def L2block(L, n):
    conn = np.stack([np.arange(1, n + 1), np.arange(2, n + 2)], 1)
    return Mesh(conn, _lin(L, n).reshape(-1, 1), "L2")


# corners of the unit cube (bit 0 = x, bit 1 = y, bit 2 = z) of the six
# tetrahedra (0, a, a|b, 7) along the monotone paths from corner 0 to 7; the
# last two are swapped where that makes the orientation positive
_KUHN = [(0, 1, 3, 7), (0, 1, 7, 5), (0, 2, 7, 3), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 7, 6)]


def T4block(A, B, C, nA, nB, nC):
    xs, ys, zs = _lin(A, nA), _lin(B, nB), _lin(C, nC)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")  # x fastest when raveled
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    W, S = nA + 1, (nA + 1) * (nB + 1)
    k, j, i = np.meshgrid(np.arange(nC), np.arange(nB), np.arange(nA), indexing="ij")  # x-index fastest
    base = (k.ravel() * S + j.ravel() * W + i.ravel() + 1).astype(np.int64)
    corner = [((b >> 2) & 1) * S + ((b >> 1) & 1) * W + (b & 1) for b in range(8)]
    conn = np.stack([np.stack([base + corner[t[0]], base + corner[t[1]], base + corner[t[2]], base + corner[t[3]]], 1)
                     for t in _KUHN], 1).reshape(-1, 4)
    return Mesh(conn, xyz, "T4")


def transform(mesh, f):
    mesh.xy = np.ascontiguousarray(f(mesh.xy), dtype=np.float64)
    return mesh


def boundary_nodes(mesh):
    c = mesh.conn
    if mesh.shape in ("L2", "T4"):
        # facets (end points, triangles) that belong to one element
        faces = [(0,), (1,)] if mesh.shape == "L2" else [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
        f = np.sort(np.concatenate([c[:, list(t)] for t in faces]), axis=1)
        _, inv, cnt = np.unique(f, axis=0, return_inverse=True, return_counts=True)
        return np.unique(f[cnt[inv.ravel()] == 1].ravel())
    if mesh.shape == "T3" or mesh.shape == "T6":
        ends = [(0, 1, 3), (1, 2, 4), (2, 0, 5)]
    elif mesh.shape == "Q4":
        ends = [(0, 1, None), (1, 2, None), (2, 3, None), (3, 0, None)]
    else:
        raise ValueError(mesh.shape)
    a = np.concatenate([c[:, e[0]] for e in ends])
    b = np.concatenate([c[:, e[1]] for e in ends])
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (mesh.nnodes + 1) + hi
    _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    onb = cnt[inv] == 1
    nodes = [a[onb], b[onb]]
    if mesh.shape == "T6":
        mids = np.concatenate([c[:, e[2]] for e in ends])
        nodes.append(mids[onb])
    return np.unique(np.concatenate(nodes))


def load(path):
    """Reads a .mesh manifest (xyz file, shape, conn file; 1-based connectivity)."""
    d = os.path.dirname(os.path.abspath(path))
    with open(path) as fh:
        lines = [l.strip() for l in fh if l.strip()]
    xy = np.loadtxt(os.path.join(d, lines[0]), ndmin=2)
    conn = np.loadtxt(os.path.join(d, lines[2]), dtype=np.int64, ndmin=2)
    return Mesh(conn, xy[:, :2], lines[1])
