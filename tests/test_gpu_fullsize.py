"""The two single-GPU benchmark workloads at their full sizes, checked through
properties that do not depend on size:

- the 512 x 512 Q4RS cylinder panel (262,144 elements, 1,575,936 free dofs):
  the free-free pattern equals the structural union of the element dof
  pairs built independently here (scipy, from the connectivity and the dof
  numbers), bit for bit; K is bitwise symmetric; all 85 M values match the
  oracle's COO assembly of every element matrix within 1e-12 (relative
  Frobenius, and entrywise against the largest entry); a seeded sample of
  element matrices matches the oracle's within 1e-12; 20 central-difference steps
  match the oracle's loop on the same K and M within 1e-10; the loop is
  linear, so doubling the initial velocity doubles the state exactly (a
  power-of-two scaling commutes with every rounding); both mixed precisions
  stay within 1e-6 of the oracle over those steps and of the FP64 run over
  the benchmark's 1000 Gaussian-velocity steps;
- the 1024 x 512 T3 plate (1,048,576 elements, 3,136,518 free dofs): the
  same checks."""
import numpy as np
import pytest
import scipy.sparse as sp

from flexshell import _lib, explicit, femm as FM, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pattern_reference(conn, dofnums, nf):
    """Free-free CSR pattern of sum_e (dof pairs of e), built from scratch."""
    N = dofnums.shape[0]
    E, nen = conn.shape
    inc = sp.csr_matrix((np.ones(E * nen, dtype=np.int8), (np.repeat(np.arange(E), nen), conn.ravel())),
                        shape=(E, N))
    adj = (inc.T @ inc).tocoo()  # node pairs sharing an element (including I = J)
    I, J = adj.row.astype(np.int64), adj.col.astype(np.int64)
    r = np.repeat(dofnums[I], 6, axis=1)  # (pairs, 36): dof of I for each (a, b)
    c = np.tile(dofnums[J], (1, 6))
    keep = (r < nf) & (c < nf)
    pat = sp.csr_matrix((np.ones(int(keep.sum()), dtype=np.int8), (r[keep], c[keep])), shape=(nf, nf))
    pat.sort_indices()
    return pat


def _step_checks(f, xyz, d, seed, nsteps=20):
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d).diagonal()
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, None, None, d)
    f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, d)
    rng = np.random.default_rng(seed)
    V0 = W.gaussian_velocity(d, 1e-3, seed=seed)
    dt = 0.9 * explicit.stable_dt(explicit.omega_max(f, rng.random(nf), maxit=30))
    z = np.zeros(nf)
    g = np.zeros(nsteps + 1)
    r = explicit.run(f._model, 0.0, z, V0, dt, nsteps, z, g)
    Uo, Vo, _ = O.cd_loop(K, M, 0.0, dt, nsteps, z, g, z, V0)
    assert np.linalg.norm(r["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo)
    assert np.linalg.norm(r["V"] - Vo) <= 1e-10 * np.linalg.norm(Vo)
    r2 = explicit.run(f._model, 0.0, z, 2.0 * V0, dt, nsteps, z, g)
    assert np.array_equal(r2["U"], 2.0 * r["U"]) and np.array_equal(r2["V"], 2.0 * r["V"])
    # mixed precisions: the same steps against the oracle, and the benchmark's
    # 1000 Gaussian-velocity steps against the FP64 run (max over the state)
    gl = np.zeros(1001)
    rl = explicit.run(f._model, 0.0, z, V0, dt, 1000, z, gl)
    for prec in (_lib.FS_PREC_MIXED, _lib.FS_PREC_MIXED_F32):
        rm = explicit.run(f._model, 0.0, z, V0, dt, nsteps, z, g, precision=prec)
        assert np.abs(rm["U"] - Uo).max() <= 1e-6 * np.abs(Uo).max(), prec
        rm = explicit.run(f._model, 0.0, z, V0, dt, 1000, z, gl, precision=prec)
        err = np.abs(rm["U"] - rl["U"]).max() / np.abs(rl["U"]).max()
        print(f"[gpu] full size: precision {prec}, 1000 steps, max error {err:.2e}", flush=True)
        assert err <= 1e-6, (prec, err)
    return K


def _matrix_checks(f, nen, xyz, conn, d, K, E, nu, t):
    """Bit-exact pattern, bitwise symmetry, every value against the oracle's
    assembly of all element matrices, and a seeded element sample."""
    P = _pattern_reference(conn, d.dofnums, d.nfreedofs)
    assert np.array_equal(K.indptr, P.indptr) and np.array_equal(K.indices, P.indices)
    del P
    KT = K.T.tocsr()
    KT.sort_indices()
    assert np.array_equal(KT.indptr, K.indptr) and np.array_equal(KT.indices, K.indices)
    assert np.array_equal(KT.data, K.data)
    del KT
    nrm, val = O.normals(nen, xyz, conn)
    Dps, Dt = O.shell_D(O.iso_D6(E, nu))
    # all values: the oracle's COO -> CSR assembly of every element matrix
    # (dense transforms, duplicates summed in element order)
    Ko = O.assemble_csr(nen, conn, d.dofnums, d.nfreedofs,
                        O.element_matrices(nen, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=t))
    assert np.array_equal(K.indptr, Ko.indptr) and np.array_equal(K.indices, Ko.indices)
    rel = np.linalg.norm(K.data - Ko.data) / np.linalg.norm(Ko.data)
    print(f"[gpu] full-size K vs oracle: {K.nnz} values, relative Frobenius error {rel:.2e}", flush=True)
    assert rel <= 1e-12, rel
    assert np.abs(K.data - Ko.data).max() <= 1e-12 * np.abs(Ko.data).max()
    del Ko
    sample = np.sort(np.random.default_rng(7).choice(conn.shape[0], 256, replace=False))
    Ko = O.element_matrices(nen, xyz, conn[sample], nrm, val, Dps=Dps, Dt=Dt, t=t)
    Kg = f.element_matrices(xyz)[sample]
    err = np.linalg.norm(Kg - Ko, axis=(1, 2)) / np.linalg.norm(Ko, axis=(1, 2))
    assert err.max() < 1e-12, err.max()


@pytest.mark.timeout(600)
def test_cylinder_q4_512_full_size(ctx):
    print("[gpu] full size: Q4RS cylinder 512x512", flush=True)
    c = W.cylinder_panel_c2(n=512)
    d, xyz, conn = c["dchi"], c["xyz"], c["conn"]
    f = FM.FEMMShellQ4RS(conn, c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(xyz)
    K = _step_checks(f, xyz, d, seed=31)
    _matrix_checks(f, 4, xyz, conn, d, K, 70e9, 0.33, c["t"])


@pytest.mark.timeout(600)
def test_plate_t3_1024x512_full_size(ctx):
    print("[gpu] full size: T3 plate 1024x512", flush=True)
    c = W.plate_c4()
    d, xyz, conn = c["dchi"], c["xyz"], c["conn"]
    f = FM.FEMMShellT3FF(conn, c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(xyz)
    K = _step_checks(f, xyz, d, seed=32)
    _matrix_checks(f, 3, xyz, conn, d, K, W.STEEL["E"], W.STEEL["nu"], c["t"])
    # the mixed-precision benchmark inputs: Gaussian pressure pulse (sigma
    # 0.1 m, 1e4 Pa, 1e-4 s), 5000 steps at 0.9 of the stable dt, 64 probe
    # dofs every 10 steps, both mixed precisions against the FP64 run
    nf = d.nfreedofs
    x0 = c["x0"]
    F = W.pressure_load_t3(xyz, conn, d, lambda X: 1e4 * np.exp(-np.sum((X - x0) ** 2, axis=1) / (2 * 0.1 ** 2)))
    dt = explicit.stable_dt(explicit.omega_max(f, np.random.default_rng(7).random(nf), maxit=30))
    nsteps = 5000
    g = (explicit.sample_times(nsteps, dt) <= 1e-4).astype(float)
    wd = d.dofnums[:, 2][d.dofnums[:, 2] < nf]
    probes = wd[np.linspace(0, len(wd) - 1, 64).astype(int)]
    z = np.zeros(nf)
    ref = explicit.run(f._model, 0.0, z, z, dt, nsteps, F, g, probes, 10)["probes"]
    for prec in (_lib.FS_PREC_MIXED, _lib.FS_PREC_MIXED_F32):
        pm = explicit.run(f._model, 0.0, z, z, dt, nsteps, F, g, probes, 10, precision=prec)["probes"]
        err = np.abs(pm - ref).max() / np.abs(ref).max()
        print(f"[gpu] full size plate pulse: precision {prec}, {nsteps} steps, max error {err:.2e}", flush=True)
        assert err <= 1e-6, (prec, err)
