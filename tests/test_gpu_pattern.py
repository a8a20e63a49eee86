"""CSR patterns from the host planner (build_block, fs_model_pattern_get)
against an independent construction: the structural union of the element
dof pairs built with scipy from the connectivity and the dof numbers.
Free-free and all-dof blocks, with runs of three and six free dofs, a node
permutation, and a hub node of 48 neighbours; plus the planner's refusal of
a numbering whose free dofs of a node are not consecutive."""
import numpy as np
import pytest
import scipy.sparse as sp

import test_gpu_cd_forms as CF
from flexshell import _lib, femm as FM
from flexshell.material import MatDeforElastIso

pytestmark = pytest.mark.gpu


def _reference(conn, dofnums, nrows):
    N = dofnums.shape[0]
    E, nen = conn.shape
    inc = sp.csr_matrix((np.ones(E * nen, dtype=np.int8), (np.repeat(np.arange(E), nen), conn.ravel())),
                        shape=(E, N))
    adj = (inc.T @ inc).tocoo()
    I, J = adj.row.astype(np.int64), adj.col.astype(np.int64)
    r = np.repeat(dofnums[I], 6, axis=1)
    c = np.tile(dofnums[J], (1, 6))
    keep = (r < nrows) & (c < nrows)
    pat = sp.csr_matrix((np.ones(int(keep.sum()), dtype=np.int8), (r[keep], c[keep])), shape=(nrows, nrows))
    pat.sort_indices()
    return pat


@pytest.mark.parametrize("mesh", [CF._panel_q4, CF._plate_t3_perm, CF._fan_t3], ids=["q4_ss", "t3_perm", "t3_fan"])
def test_pattern_matches_independent_construction(ctx, mesh):
    print(f"[gpu] pattern {mesh.__name__}", flush=True)
    nen, xyz, conn, d = mesh()
    cls = FM.FEMMShellT3FF if nen == 3 else FM.FEMMShellQ4RS
    f = cls(conn, MatDeforElastIso(210e9, 0.3, 7850.0), thickness=0.01, ctx=ctx)
    f.associategeometry(xyz)
    nf, na = d.nfreedofs, d.nalldofs
    Kff = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    Kall = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), xyz, None, None, d)
    for K, n in ((Kff, nf), (Kall, na)):
        P = _reference(conn, d.dofnums, n)
        assert K.shape == (n, n)
        assert np.array_equal(K.indptr, P.indptr) and np.array_equal(K.indices, P.indices)
        KT = K.T.tocsr()
        KT.sort_indices()
        assert np.array_equal(KT.data, K.data)  # bitwise symmetric
    # one-based expansion of the same pattern
    m = f._model
    nr, rp, cv = m.pattern(_lib.FS_BLOCK_FF)
    rp1 = np.zeros_like(rp)
    cv1 = np.zeros_like(cv)
    ctx.check(ctx.lib.fs_model_pattern_get(m.h, _lib.FS_BLOCK_FF, _lib.ptr(rp1, _lib.C.c_int64),
                                           _lib.ptr(cv1, _lib.C.c_int64), 1), "fs_model_pattern_get")
    assert np.array_equal(rp1, rp + 1) and np.array_equal(cv1, cv + 1)


def test_pattern_refuses_split_free_dofs(ctx):
    """Free dofs of a node that are not consecutive numbers: the planner
    names the lowest such node."""
    nen, xyz, conn, d = CF._panel_q4()
    dn = d.dofnums.copy()
    nf = d.nfreedofs
    # swap one free dof of node 5 with one of node 9 (both interior, all six free)
    assert (dn[5] < nf).all() and (dn[9] < nf).all()
    dn[5, 2], dn[9, 2] = dn[9, 2], dn[5, 2]
    m = _lib.Model(ctx, nen, xyz, conn)
    try:
        m.set_dofs(dn, d.nalldofs, nf)
        with pytest.raises(_lib.FlexShellError, match="node 5: free dofs are not consecutive"):
            m.pattern(_lib.FS_BLOCK_FF)
    finally:
        m.close()
