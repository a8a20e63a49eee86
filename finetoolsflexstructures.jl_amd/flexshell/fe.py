"""Host-side finite-element utilities the shell FEMMs are called with.

These mirror the FinEtools (8.2.x) calls that the reference's drivers and
tests use around the hot path (mesh generators, equation numbering, node and
boundary selection, distributed loads).  They produce the plain arrays that
cross the C ABI; none of the hot-path arithmetic is here.

Index conventions: node and element numbers are 0-based; dof component
numbers (1..6 in FinEtools) are given 1-based to setebc to keep driver code
readable next to the reference.
"""
import numpy as np

# GaussRule(1, 2) / GaussRule(2, 2) abscissa as FinEtools tabulates it
GAUSS2 = 0.577350269189626


def linearspace(a, b, n):
    return np.linspace(a, b, n)


def T3blockx(xs, ys, orientation="a"):
    """Structured triangle mesh (FinEtools MeshTriangleModule.T3blockx).

    Nodes run along x fastest; each rectangle (i, j) is split into two
    triangles, orientation :a along the diagonal from (i+1, j) to (i, j+1).
    Returns (xy: N x 2, conn: E x 3, 0-based)."""
    xs = np.asarray(xs, dtype=float)
    ys = np.asarray(ys, dtype=float)
    nL, nW = len(xs) - 1, len(ys) - 1
    X, Y = np.meshgrid(xs, ys)  # row j, column i -> node j*(nL+1)+i
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    # element order: i (along x) outer, j inner; two triangles per rectangle
    I, J = np.meshgrid(np.arange(nL), np.arange(nW), indexing="ij")
    f = (J * (nL + 1) + I).ravel().astype(np.int64)
    conn = np.zeros((2 * nL * nW, 3), dtype=np.int64)
    if orientation == "a":
        conn[0::2] = np.stack([f, f + 1, f + (nL + 1)], axis=1)
        conn[1::2] = np.stack([f + 1, f + (nL + 1) + 1, f + (nL + 1)], axis=1)
    elif orientation == "b":
        conn[0::2] = np.stack([f, f + 1, f + (nL + 1) + 1], axis=1)
        conn[1::2] = np.stack([f, f + (nL + 1) + 1, f + (nL + 1)], axis=1)
    else:
        raise ValueError(f"Cannot handle orientation : {orientation}")
    return xy, conn


def T3block(Length, Width, nL, nW, orientation="a"):
    return T3blockx(linearspace(0.0, Length, nL + 1), linearspace(0.0, Width, nW + 1), orientation)


def Q4blockx(xs, ys):
    """Structured quadrilateral mesh (FinEtools MeshQuadrilateralModule.Q4blockx)."""
    xs = np.asarray(xs, dtype=float)
    ys = np.asarray(ys, dtype=float)
    nL, nW = len(xs) - 1, len(ys) - 1
    X, Y = np.meshgrid(xs, ys)
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    I, J = np.meshgrid(np.arange(nL), np.arange(nW), indexing="ij")
    f = (J * (nL + 1) + I).ravel().astype(np.int64)
    conn = np.stack([f, f + 1, f + (nL + 1) + 1, f + (nL + 1)], axis=1)
    return xy, conn


def Q4block(Length, Width, nL, nW):
    return Q4blockx(linearspace(0.0, Length, nL + 1), linearspace(0.0, Width, nW + 1))


def xyz3(xy):
    """Pad 2-D coordinates with z = 0 (FinEtools xyz3)."""
    xy = np.asarray(xy, dtype=float)
    if xy.shape[1] == 3:
        return xy.copy()
    return np.hstack([xy, np.zeros((xy.shape[0], 3 - xy.shape[1]))])


def selectnode(xyz, box, inflate=0.0):
    """Nodes inside an axis-aligned box [xmin xmax ymin ymax zmin zmax] (inflated)."""
    xyz = np.asarray(xyz)
    box = np.asarray(box, dtype=float).reshape(-1)
    keep = np.ones(xyz.shape[0], dtype=bool)
    for d in range(xyz.shape[1]):
        lo, hi = box[2 * d] - inflate, box[2 * d + 1] + inflate
        keep &= (xyz[:, d] >= lo) & (xyz[:, d] <= hi)
    return np.nonzero(keep)[0]


def nearest_node(xyz, point):
    d = np.linalg.norm(np.asarray(xyz) - np.asarray(point, dtype=float)[None, :], axis=1)
    return int(np.argmin(d))


def boundary_edges(conn):
    """Edges used by exactly one element (meshboundary of a T3/Q4 mesh)."""
    conn = np.asarray(conn)
    nen = conn.shape[1]
    e = np.concatenate([conn[:, [a, (a + 1) % nen]] for a in range(nen)])
    key = np.sort(e, axis=1)
    _, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    return e[np.sort(idx[cnt == 1])]


def boundary_nodes(conn):
    """connectednodes(meshboundary(fes))."""
    return np.unique(boundary_edges(conn))


class DofField:
    """Equation numbering of a 6-dof-per-node field (FinEtools NodalField
    setebc!/applyebc!/numberdofs!): free dofs first in node-major order (or in
    the order of a node permutation), then the fixed ones."""

    def __init__(self, nnodes, ndofn=6):
        self.N = nnodes
        self.ndofn = ndofn
        self.is_fixed = np.zeros((nnodes, ndofn), dtype=bool)
        self.fixed_values = np.zeros((nnodes, ndofn))
        self.dofnums = None
        self.nfreedofs = 0
        self.nalldofs = nnodes * ndofn

    def setebc(self, nodes, comps, value=0.0):
        for c in np.atleast_1d(comps):
            self.is_fixed[np.asarray(nodes, dtype=np.int64), int(c) - 1] = True
            self.fixed_values[np.asarray(nodes, dtype=np.int64), int(c) - 1] = value
        return self

    def numberdofs(self, perm=None):
        order = np.arange(self.N) if perm is None else np.asarray(perm)
        fixed = self.is_fixed[order]  # nodes in numbering order, node-major
        dn_ord = np.empty(fixed.shape, dtype=np.int64)
        free = ~fixed
        self.nfreedofs = int(free.sum())
        dn_ord[free] = np.arange(self.nfreedofs)
        dn_ord[fixed] = self.nfreedofs + np.arange(int(fixed.sum()))
        dn = np.empty_like(dn_ord)
        dn[order] = dn_ord
        self.dofnums = dn
        return self

    def gathersysvec(self, values):
        """Free-dof vector from an N x 6 nodal array."""
        out = np.zeros(self.nfreedofs)
        m = ~self.is_fixed
        out[self.dofnums[m]] = np.asarray(values)[m]
        return out

    def scattersysvec(self, vec):
        """N x 6 nodal array from a free-dof vector (fixed dofs get their prescribed values)."""
        out = self.fixed_values.copy()
        m = ~self.is_fixed
        out[m] = np.asarray(vec)[self.dofnums[m]]
        return out


def edge_load(xyz, edges, force, nalldofs, dofnums):
    """distribloads on L2 boundary elements with GaussRule(1, 2) (uniform force 6-vector)."""
    F = np.zeros(nalldofs)
    f = np.asarray(force, dtype=float)
    for (a, b) in np.asarray(edges):
        L = np.linalg.norm(np.asarray(xyz)[b] - np.asarray(xyz)[a])
        for xi in (-GAUSS2, GAUSS2):
            N = np.array([(1 - xi) / 2, (1 + xi) / 2])
            F[dofnums[a]] += N[0] * f * (L / 2)
            F[dofnums[b]] += N[1] * f * (L / 2)
    return F
