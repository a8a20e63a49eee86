"""Fused assembly (element tiles writing the CSR values of the node pairs they
complete, boundary pairs finalised from tile partial sums) against the
two-kernel form (element blocks through HBM + pair scatter, sums in element
order) and against the oracle: identical pattern, values within 1e-13 of
the scatter form and 1e-12 of the oracle, bitwise reproducible, for every
element family and shear formulation, both blocks, element numberings
shuffled so that tiles and element order disagree, a fan of 48
triangles around one node (a pair with more contributions in one tile than
one tile entry holds), a one-element model and two disconnected patches."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import _lib, fe, femm as FM, workloads as W
from flexshell.material import MatDeforElastIso
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dofs(xyz):
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [-np.inf, 1e-9, -np.inf, np.inf, -np.inf, np.inf], 1e-6), [1, 2, 3, 4, 5, 6])
    d.setebc(fe.selectnode(xyz, [0.99, np.inf, -np.inf, np.inf, -np.inf, np.inf], 1e-6), [1, 2, 3])
    d.numberdofs()
    return d


def _both_forms(ctx, f, xyz, d):
    out = {}
    for mode in (0, 1):
        ctx.set_assembly(mode)
        K = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), xyz, None, None, d).tocsr()
        Kf = f.stiffness(FM.SysmatAssemblerFFBlock(d.nfreedofs), xyz, None, None, d).tocsr()
        out[mode] = (K, Kf)
    ctx.set_assembly(1)
    return out


@pytest.mark.parametrize("nen,layered,form,shuffle", [(3, False, 0, False), (3, False, 1, False), (3, True, 0, True),
                                                      (3, True, 1, False), (4, False, 0, False), (4, True, 0, False),
                                                      (4, False, 0, True), (4, True, 0, True), (3, False, 0, True)])
def test_fused_matches_scatter_and_oracle(nen, layered, form, shuffle):
    print(f"[gpu] fused assembly nen={nen} layered={layered} form={form} shuffle={shuffle}", flush=True)
    ctx = _lib.Context(0)
    xyz, conn = SC.distorted_mesh(nen, 37, 29, seed=9, warp=0.4)
    if shuffle:
        conn = conn[np.random.default_rng(2).permutation(conn.shape[0])]
    d = _dofs(xyz)
    if layered:
        layup = W.c5_layup(4, [0.0, 45.0, -45.0, 90.0])
        f = FM.make(conn, layup, ctx=ctx)
    else:
        f = FM.make(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
    f.transv_shear_formulation = form
    f.associategeometry(xyz)
    forms = _both_forms(ctx, f, xyz, d)
    Ke = f.element_matrices(xyz)
    for i, nrows in enumerate((d.nalldofs, d.nfreedofs)):
        Ks, Kt = forms[0][i], forms[1][i]
        assert np.array_equal(Ks.indptr, Kt.indptr) and np.array_equal(Ks.indices, Kt.indices)
        assert np.linalg.norm(Kt.data - Ks.data) <= 1e-13 * np.linalg.norm(Ks.data)
        KT = Kt.T.tocsr()
        KT.sort_indices()
        assert np.array_equal(KT.data, Kt.data)  # bitwise symmetric
        Ko = O.assemble_csr(nen, conn, d.dofnums, nrows, Ke)
        assert np.array_equal(Ko.indices, Kt.indices)
        assert np.linalg.norm(Kt.data - Ko.data) <= 1e-12 * np.linalg.norm(Ko.data)
    # bitwise reproducible
    K2 = f.stiffness(FM.SysmatAssemblerFFBlock(d.nfreedofs), xyz, None, None, d).tocsr()
    assert np.array_equal(K2.data, forms[1][1].data)
    f._model.close()
    ctx.close()


def test_fused_fan_hub():
    """48 triangles around one node: the hub's diagonal block collects 48
    contributions, more than one tile entry carries, in one or two tiles."""
    print("[gpu] fused assembly: 48-triangle fan", flush=True)
    ctx = _lib.Context(0)
    n = 48
    ang = np.arange(n) * 2 * np.pi / n
    ring = np.stack([np.cos(ang), np.sin(ang), 0.05 * np.sin(3 * ang)], axis=1)
    ring2 = 2.0 * ring
    xyz = np.vstack([[0.0, 0.0, 0.0], ring, ring2])
    conn = [[0, 1 + k, 1 + (k + 1) % n] for k in range(n)]
    conn += [[1 + k, 1 + n + k, 1 + (k + 1) % n] for k in range(n)]
    conn += [[1 + (k + 1) % n, 1 + n + k, 1 + n + (k + 1) % n] for k in range(n)]
    conn = np.array(conn)
    d = fe.DofField(xyz.shape[0])
    d.setebc(np.arange(1 + n, 1 + 2 * n), [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    f = FM.FEMMShellT3FF(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.02, ctx=ctx)
    f.associategeometry(xyz)
    forms = _both_forms(ctx, f, xyz, d)
    Ke = f.element_matrices(xyz)
    for i, nrows in enumerate((d.nalldofs, d.nfreedofs)):
        Ks, Kt = forms[0][i], forms[1][i]
        assert np.array_equal(Ks.indices, Kt.indices)
        assert np.linalg.norm(Kt.data - Ks.data) <= 1e-13 * np.linalg.norm(Ks.data)
        Ko = O.assemble_csr(3, conn, d.dofnums, nrows, Ke)
        assert np.linalg.norm(Kt.data - Ko.data) <= 1e-12 * np.linalg.norm(Ko.data)
    f._model.close()
    ctx.close()


@pytest.mark.parametrize("nen", [3, 4])
def test_fused_single_element_and_split_patches(nen):
    """Tiles that are mostly empty and tiles cut by components: a model of
    one element (one tile whose unused slots repeat that element), and a
    model of two disconnected patches of 3 and 14 (T3: 6 and 28) elements,
    the second curved and rotated, each patch filling a partial tile of its
    own.  Fused form against the two-kernel form and the oracle."""
    print(f"[gpu] fused assembly: one element and two patches, nen={nen}", flush=True)
    ctx = _lib.Context(0)
    xa, ca = SC.distorted_mesh(nen, 1, 1, seed=3, warp=0.3)
    one = (xa, ca[:1])
    xb, cb = SC.distorted_mesh(nen, 3, 1, seed=4, warp=0.2)
    xc, cc = SC.distorted_mesh(nen, 7, 2, seed=6, warp=0.6)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    xc = xc @ R.T + np.array([3.0, -1.0, 0.5])
    two = (np.vstack([xb, xc]), np.vstack([cb, cc + xb.shape[0]]))
    for name, (xyz, conn) in (("one element", one), ("two patches", two)):
        d = fe.DofField(xyz.shape[0])
        d.setebc([0], [1, 2, 3, 4, 5, 6])
        d.setebc([xyz.shape[0] - 1], [3])
        d.numberdofs()
        f = FM.make(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
        f.associategeometry(xyz)
        forms = _both_forms(ctx, f, xyz, d)
        Dps, Dt = O.shell_D(O.iso_D6(70e9, 0.33))
        nrm, val = O.normals(nen, xyz, conn)
        Ke = O.element_matrices(nen, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=0.01)
        for i, nrows in enumerate((d.nalldofs, d.nfreedofs)):
            Ks, Kt = forms[0][i], forms[1][i]
            assert np.array_equal(Ks.indptr, Kt.indptr) and np.array_equal(Ks.indices, Kt.indices), name
            assert np.linalg.norm(Kt.data - Ks.data) <= 1e-13 * np.linalg.norm(Ks.data), name
            KT = Kt.T.tocsr()
            KT.sort_indices()
            assert np.array_equal(KT.data, Kt.data), name
            Ko = O.assemble_csr(nen, conn, d.dofnums, nrows, Ke)
            assert np.array_equal(Ko.indices, Kt.indices), name
            assert np.linalg.norm(Kt.data - Ko.data) <= 1e-12 * np.linalg.norm(Ko.data), name
        f._model.close()
    ctx.close()
