"""The generic assembler (assemble! / makematrix! with the element matrices
computed on the host, reduced on the device): the pattern and values of
sparse(I, J, V) as the oracle's COO assembly forms them, bit for bit (same
entries, duplicates summed in the same order), for two-node elements of 12
dofs (a beam-like chain with branches) in the all-dof and the free-free
block, for shell element matrices of the library's own element kernel, and
for rectangular element matrices against scipy's COO -> CSR."""
import numpy as np
import pytest
import scipy.sparse as sp

from flexshell import femm as FM, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _run(asm, mats, drs, dcs, nrows, ncols):
    asm.startassembly(mats.shape[1], mats.shape[2], mats.shape[0], nrows, ncols)
    for m, r, c in zip(mats, drs, dcs):
        asm.assemble(m, r, c)
    return asm.makematrix()


def test_generic_two_node_elements(ctx):
    print("[gpu] generic assembler: 2-node 12-dof elements", flush=True)
    rng = np.random.default_rng(3)
    N = 400
    conn = np.array([[i, i + 1] for i in range(N - 1)] + [[i, (7 * i + 11) % N] for i in range(0, N, 5)
                                                          if (7 * i + 11) % N != i])
    E = conn.shape[0]
    fixed = np.zeros((N, 6), dtype=bool)
    fixed[0] = True
    fixed[rng.choice(N, 30, replace=False), 2] = True
    dn, nf = O.number_dofs(fixed)
    A = rng.normal(size=(E, 12, 12))
    mats = A @ np.swapaxes(A, 1, 2)  # symmetric element matrices
    asm = FM.SysmatAssemblerHIPGeneric(ctx)
    for nrows in (6 * N, nf):
        drs = np.stack([np.concatenate([dn[a], dn[b]]) for a, b in conn])
        K = _run(asm, mats, drs, drs, nrows, nrows)
        Ko = O.assemble_csr(2, conn, dn, nrows, mats)
        assert np.array_equal(K.indptr, Ko.indptr) and np.array_equal(K.indices, Ko.indices)
        assert np.array_equal(K.data, Ko.data)
        # the plan is reused for new values with the same dof numbers
        K2 = _run(asm, 2.0 * mats, drs, drs, nrows, nrows)
        assert np.array_equal(K2.data, 2.0 * K.data)
    asm.close()


def test_generic_shell_element_matrices(ctx):
    print("[gpu] generic assembler: shell element matrices of the Q4 kernel", flush=True)
    c = W.cylinder_panel_c2(n=24)
    d, xyz, conn = c["dchi"], c["xyz"], c["conn"]
    f = FM.FEMMShellQ4RS(conn, c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(xyz)
    nf = d.nfreedofs
    Ke = f.element_matrices(xyz)
    K_dev = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    drs = d.dofnums[conn].reshape(conn.shape[0], 24)
    K = _run(FM.SysmatAssemblerHIPGeneric(ctx), Ke, drs, drs, nf, nf)
    Ko = O.assemble_csr(4, conn, d.dofnums, nf, Ke)
    assert np.array_equal(K.indices, Ko.indices) and np.array_equal(K.data, Ko.data)
    # same pattern as the shell path, values within rounding
    assert np.array_equal(K.indices, K_dev.indices) and np.array_equal(K.indptr, K_dev.indptr)
    assert np.linalg.norm(K.data - K_dev.data) <= 1e-13 * np.linalg.norm(K.data)


def test_generic_rectangular(ctx):
    print("[gpu] generic assembler: rectangular element matrices", flush=True)
    rng = np.random.default_rng(5)
    nm, kr, kc, nrows, ncols = 300, 5, 3, 120, 80
    drs = rng.integers(0, nrows + 10, (nm, kr))
    dcs = rng.integers(0, ncols + 10, (nm, kc))
    mats = rng.normal(size=(nm, kr, kc))
    K = _run(FM.SysmatAssemblerHIPGeneric(ctx), mats, drs, dcs, nrows, ncols)
    I = np.repeat(drs[:, :, None], kc, axis=2).ravel()
    J = np.repeat(dcs[:, None, :], kr, axis=1).ravel()
    keep = (I < nrows) & (J < ncols)
    S = sp.coo_matrix((mats.ravel()[keep], (I[keep], J[keep])), shape=(nrows, ncols)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(K.indptr, S.indptr) and np.array_equal(K.indices, S.indices)
    np.testing.assert_allclose(K.data, S.data, rtol=1e-14, atol=1e-14)
