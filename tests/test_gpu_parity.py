"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerances (the accuracy contract of the drop-in): stiffness within 1e-12 relative Frobenius
error with a bit-exact CSR pattern; lowest frequencies from the GPU-assembled
K and M within 1e-9 of the oracle's; explicit trajectories within 1e-10
relative to the FP64 oracle trajectory; integer structure bit-exact."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import shell_cases as SC
from flexshell import _lib, explicit, fe, femm as FM
from flexshell.material import CompositeLayup, MatDeforElastIso, Ply, lamina_material
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KTOL = 1e-12


def _rel(a, b):
    a = a.toarray() if sp.issparse(a) else np.asarray(a)
    b = b.toarray() if sp.issparse(b) else np.asarray(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _homog_femm(ctx, nen, conn, E, nu, rho, t, stab_c=None):
    cls = FM.FEMMShellT3FF if nen == 3 else FM.FEMMShellQ4RS
    return cls(conn, MatDeforElastIso(E, nu, rho), thickness=t, stab_c=stab_c, ctx=ctx)


def _oracle_homog(nen, xyz, conn, E, nu, t, stab_c=None, form=0, drill=1.0):
    Dps, Dt = O.shell_D(O.iso_D6(E, nu))
    nrm, val = O.normals(nen, xyz, conn)
    return O.element_matrices(nen, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=t, stab_c=stab_c, formulation=form,
                              drill=drill)


@pytest.mark.parametrize("nen,warp,form,drill", [(3, 0.0, 0, 1.0), (3, 0.5, 0, 1.0), (3, 0.5, 1, 0.3),
                                                 (4, 0.0, 0, 1.0), (4, 0.5, 0, 1.0), (4, 0.8, 0, 0.1)])
def test_element_matrices_homogeneous(ctx, nen, warp, form, drill):
    print(f"[gpu] element matrices nen={nen} warp={warp} form={form}", flush=True)
    xyz, conn = SC.distorted_mesh(nen, 7, 5, seed=11, warp=warp)
    E, nu, t = 70e9, 0.33, 0.01
    f = _homog_femm(ctx, nen, conn, E, nu, 0.0, t)
    f.transv_shear_formulation = form
    f.drilling_stiffness_scale = drill
    f.associategeometry(xyz)
    Kg = f.element_matrices(xyz)
    Ko = _oracle_homog(nen, xyz, conn, E, nu, t, form=form, drill=drill)
    per = np.linalg.norm(Kg - Ko, axis=(1, 2)) / np.linalg.norm(Ko, axis=(1, 2))
    assert per.max() < KTOL, per.max()


@pytest.mark.parametrize("nen", [3, 4])
def test_normals_match_oracle(ctx, nen):
    print(f"[gpu] normals nen={nen}", flush=True)
    xyz, conn = SC.distorted_mesh(nen, 8, 6, seed=5, warp=0.7)
    f = _homog_femm(ctx, nen, conn, 70e9, 0.3, 0.0, 0.01)
    f.associategeometry(xyz)
    n_o, v_o = O.normals(nen, xyz, conn)
    np.testing.assert_allclose(f._normals, n_o, rtol=0, atol=1e-14)
    assert (f._normal_valid == v_o).all()
    # a second call recomputes from the coordinates: bitwise the same
    n1, v1 = f._normals.copy(), f._normal_valid.copy()
    f.associategeometry(xyz)
    assert np.array_equal(f._normals, n1) and np.array_equal(f._normal_valid, v1)
    # folded plate: the crease invalidates the fold nodes
    xy, c3 = fe.T3block(1.0, 1.0, 4, 4) if nen == 3 else fe.Q4block(1.0, 1.0, 4, 4)
    fold = fe.xyz3(xy)
    m = fold[:, 0] > 0.5
    fold[m, 2] = fold[m, 0] - 0.5
    fold[m, 0] = 0.5
    f2 = _homog_femm(ctx, nen, c3, 70e9, 0.3, 0.0, 0.01)
    f2.associategeometry(fold)
    n_o, v_o = O.normals(nen, fold, c3)
    assert (f2._normal_valid == v_o).all() and (~v_o).sum() == 5
    np.testing.assert_allclose(f2._normals, n_o, rtol=0, atol=1e-14)


@pytest.mark.parametrize("nen,block", [(3, 0), (3, 1), (4, 0), (4, 1)])
def test_csr_pattern_and_values(ctx, nen, block):
    """Scordelis-Lo (curved, fixed dofs interleaved): CSR pattern bit-exact
    against the oracle's COO assembly; values within 1e-12."""
    print(f"[gpu] csr nen={nen} block={block}", flush=True)
    c = SC.scordelis(nen, 8)
    d = c["dchi"]
    f = _homog_femm(ctx, nen, c["conn"], c["E"], c["nu"], 0.0, c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    asm = FM.SysmatAssemblerSparseCSRSymm() if block == 0 else FM.SysmatAssemblerFFBlock(d.nfreedofs)
    Kg = f.stiffness(asm, c["xyz"], None, None, d)
    nrows = d.nalldofs if block == 0 else d.nfreedofs
    Ko = O.assemble_csr(nen, c["conn"], d.dofnums, nrows,
                        _oracle_homog(nen, c["xyz"], c["conn"], c["E"], c["nu"], c["t"], stab_c=c["stab_c"]))
    assert np.array_equal(Kg.indptr, Ko.indptr) and np.array_equal(Kg.indices, Ko.indices)
    assert _rel(Kg, Ko) < KTOL
    # bitwise symmetric by construction, and the solve reproduces the reference value
    assert (Kg != Kg.T).nnz == 0
    if block == 1:
        U = sla.spsolve(Kg.tocsc(), c["F"][:d.nfreedofs])
        w = d.scattersysvec(U)[c["probe"], 2]
        ref = (SC.SCORDELIS_T3 if nen == 3 else SC.SCORDELIS_Q4)[8]
        assert abs(w / -0.3024 * 100 - ref) / ref < 1e-8


def _flexshell_layup(case, rho=0.0):
    E1, E2, nu12, G12, G13, G23 = case["lamina"]
    mat = lamina_material(rho, E1, E2, nu12, G12, G13, G23)
    plies = [Ply(f"p{i}", mat, t, ang) for i, (_, t, ang, _) in enumerate(case["plies"])]
    return CompositeLayup("layup", plies, case["csys"])


@pytest.mark.parametrize("nen", [3, 4])
def test_layered_barbero(ctx, nen):
    print(f"[gpu] layered barbero nen={nen}", flush=True)
    c = SC.barbero(nen)
    d = c["dchi"]
    lay = _flexshell_layup(c)
    cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
    f = cls(c["conn"], lay, ctx=ctx)
    f.associategeometry(c["xyz"])
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    A, B, D, H, _ = O.laminate(c["plies"])
    nrm, val = O.csys_normals(nen, c["xyz"], c["conn"], c["csys"])
    Ko = O.assemble_csr(nen, c["conn"], d.dofnums, d.nalldofs,
                        O.element_matrices(nen, c["xyz"], c["conn"], nrm, val,
                                           groups=[(A, B, D, H, c["csys"], c["th"])]))
    assert np.array_equal(Kg.indices, Ko.indices)
    assert _rel(Kg, Ko) < KTOL
    nf = d.nfreedofs
    U = sla.spsolve(Kg[:nf, :nf].tocsc(), c["F"][:nf])
    wmax = d.scattersysvec(U)[:, 2].max() / SC.mm
    ref = SC.BARBERO_T3 if nen == 3 else SC.BARBERO_Q4
    assert abs(wmax - ref) / ref < 1e-10


@pytest.mark.parametrize("nen", [3, 4])
def test_layered_two_layup_groups(ctx, nen):
    """Two layup groups on alternating elements of the Barbero plate (the
    reference's layup_groups): a [0/90] and a [30/-60/30] laminate of other
    thicknesses whose frame is turned about x. Every element matrix must use
    its own group's laminate and frame: K against the oracle within 1e-12."""
    print(f"[gpu] layered two groups nen={nen}", flush=True)
    c = SC.barbero(nen)
    d = c["dchi"]
    D6 = c["plies"][0][0]
    plies_b = [(D6, 3 * SC.mm, 30.0, 0.0), (D6, 4 * SC.mm, -60.0, 0.0), (D6, 3 * SC.mm, 30.0, 0.0)]
    ca, sa = np.cos(0.3), np.sin(0.3)
    csys_b = np.array([[1.0, 0.0, 0.0], [0.0, ca, -sa], [0.0, sa, ca]])
    lay_a = _flexshell_layup(c)
    lay_b = _flexshell_layup(dict(c, plies=plies_b, csys=csys_b))
    E = c["conn"].shape[0]
    eg = (np.arange(E) % 2).astype(np.int32)
    cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
    f = cls(c["conn"], None, layup_groups=[(lay_a, np.nonzero(eg == 0)[0]), (lay_b, np.nonzero(eg == 1)[0])],
            ctx=ctx)
    f.associategeometry(c["xyz"])
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    A, B, D, H, _ = O.laminate(c["plies"])
    Ab, Bb, Db, Hb, _ = O.laminate(plies_b)
    groups = [(A, B, D, H, c["csys"], c["th"]), (Ab, Bb, Db, Hb, csys_b, 10 * SC.mm)]
    Ko = O.assemble_csr(nen, c["conn"], d.dofnums, d.nalldofs,
                        O.element_matrices(nen, c["xyz"], c["conn"], f._normals, f._normal_valid, groups=groups,
                                           elem_group=eg))
    assert np.array_equal(Kg.indices, Ko.indices)
    assert _rel(Kg, Ko) < KTOL
    # and each group alone: a model of only its elements gives the same element matrices
    Ke = f.element_matrices(c["xyz"])
    for gi, lay in enumerate((lay_a, lay_b)):
        sel = np.nonzero(eg == gi)[0]
        fg = cls(c["conn"][sel], lay, ctx=ctx)
        fg.associategeometry(c["xyz"])
        fg._normals, fg._normal_valid = f._normals.copy(), f._normal_valid.copy()
        fg._model.normals_upload(fg._normals, fg._normal_valid)
        Kge = fg.element_matrices(c["xyz"])
        assert np.abs(Kge - Ke[sel]).max() <= 1e-12 * np.abs(Ke[sel]).max(), gi
        fg._model.close()
    f._model.close()


@pytest.mark.parametrize("nen", [3, 4])
@pytest.mark.parametrize("case", ["c1", "c2", "c6", "c7", "c8+", "c9-", "c10w", "c10u"])
def test_composite_reference_values_gpu(ctx, case, nen):
    """The T3FFComp / Q4RSComp plates of test_composite_shell_statics.jl and
    test_composite_shell_statics_q4rs.jl with K and the pressure load from the
    device: the values of their @test lines."""
    print(f"[gpu] composite reference {case} nen={nen}", flush=True)
    c = SC.comp_t3(case, nen)
    d = c["dchi"]
    lay = _flexshell_layup(c)
    cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
    f = cls(c["conn"], lay, ctx=ctx)
    f.associategeometry(c["xyz"])
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    F = c["F"].copy()
    if c["pressure"] is not None:
        F += f.distribloads(c["xyz"], d, c["pressure"], c["load_npts"], _lib.FS_BLOCK_ALL)
    comp, sign, scale = c["quantity"]
    v = SC._csr_solve_max(Kg, F, d, comp, sign) * scale
    assert abs(v - c["ref"]) / abs(c["ref"]) < c["rtol"], (v, c["ref"])


@pytest.mark.parametrize("nen", [3, 4])
@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_twisted_beam_gpu(ctx, nen, which):
    """Twisted cantilever (test_shell_statics.jl:350-386,
    test_q4rs_shell_statics.jl:346-382) on the 16 x 8 mesh with K from the
    device: CSR within 1e-12 of the oracle's and the reference percentage
    (bars as in the oracle pin: 1e-8 thick, 1e-4 thin)."""
    print(f"[gpu] twisted beam nen={nen} case={which}", flush=True)
    c = SC.twisted_beam(nen, which, 8)
    d = c["dchi"]
    f = _homog_femm(ctx, nen, c["conn"], c["E"], c["nu"], 0.0, c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    Ko = O.assemble_csr(nen, c["conn"], d.dofnums, d.nalldofs,
                        _oracle_homog(nen, c["xyz"], c["conn"], c["E"], c["nu"], c["t"], stab_c=c["stab_c"]))
    assert np.array_equal(Kg.indices, Ko.indices) and _rel(Kg, Ko) < KTOL
    nf = d.nfreedofs
    U = sla.spsolve(Kg[:nf, :nf].tocsc(), c["F"][:nf])
    v = d.scattersysvec(U)[c["probe"], c["dir"] - 1] / c["uex"] * 100
    assert abs(v - c["ref"]) / c["ref"] < (1e-8 if which < 2 else 1e-4), (v, c["ref"])


@pytest.mark.parametrize("composite", [False, True])
def test_petyt_plate_frequencies_gpu(ctx, composite):
    """mcompshelldyn0 of test_composite_shell_dynamics.jl with the FF-block K
    and lumped M from the device: the 8 reference frequencies within 1e-9."""
    print(f"[gpu] petyt plate composite={composite}", flush=True)
    c = SC.petyt(composite)
    d = c["dchi"]
    nf = d.nfreedofs
    if composite:
        mat = lamina_material(c["rho"], c["E"], c["nu"])  # isotropic lamina, as the reference's
        lay = CompositeLayup("petyt", [Ply(f"p{i}", mat, t, ang) for i, (_, t, ang, _) in enumerate(c["plies"])],
                             c["csys"])
        f = FM.FEMMShellT3FFComp(c["conn"], lay, ctx=ctx)
    else:
        f = _homog_femm(ctx, 3, c["conn"], c["E"], c["nu"], c["rho"], c["t"])
    f.associategeometry(c["xyz"])
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), c["xyz"], None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), c["xyz"], d).diagonal()
    w = sl.eigh(K.toarray(), np.diag(M), eigvals_only=True, subset_by_index=[0, 7])
    fs = np.sqrt(w) / (2 * np.pi)
    assert np.linalg.norm(fs - SC.PETYT_FS) < 1e-9 * np.linalg.norm(fs), fs


def test_le5_z_cantilever_gpu(ctx):
    """LE5 Z-section cantilever (test_shell_statics.jl:438-531) with K from
    the device, normals from the device (folds: invalid normals): the
    reference's minimum and maximum deflection."""
    print("[gpu] LE5 Z cantilever", flush=True)
    c = SC.le5()
    d = c["dchi"]
    f = _homog_femm(ctx, 3, c["conn"], c["E"], c["nu"], 0.0, c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    assert (~f._normal_valid).any()
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    nf = d.nfreedofs
    w = d.scattersysvec(sla.spsolve(Kg[:nf, :nf].tocsc(), c["F"][:nf]))[:, 2]
    assert abs(w.min() - SC.LE5_W[0]) / abs(SC.LE5_W[0]) < 1e-9, w.min()
    assert abs(w.max() - SC.LE5_W[1]) / abs(SC.LE5_W[1]) < 1e-9, w.max()


@pytest.mark.parametrize("mesh", ["1x9", "10x72"])
def test_raasch_hook_gpu(ctx, mesh):
    """GPU stiffness on the reference's Abaqus meshes (test_shell_statics.jl:
    148-234): CSR equal to the oracle's and the tip deflection percentage of
    the reference."""
    print(f"[gpu] raasch hook {mesh}", flush=True)
    c = SC.raasch(mesh)
    d = c["dchi"]
    f = _homog_femm(ctx, 3, c["conn"], c["E"], c["nu"], 0.0, c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    Kg = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    Ko = O.assemble_csr(3, c["conn"], d.dofnums, d.nalldofs,
                        _oracle_homog(3, c["xyz"], c["conn"], c["E"], c["nu"], c["t"], stab_c=c["stab_c"]))
    assert np.array_equal(Kg.indices, Ko.indices) and np.array_equal(Kg.indptr, Ko.indptr)
    assert _rel(Kg, Ko) < KTOL
    nf = d.nfreedofs
    U = sla.spsolve(Kg[:nf, :nf].tocsc(), c["F"][:nf])
    pct = d.scattersysvec(U)[c["probe"], 2] / 5.02 * 100
    assert abs(pct - SC.RAASCH_T3[mesh]) / SC.RAASCH_T3[mesh] < 1e-9


@pytest.mark.parametrize("nen", [3, 4])
def test_layered_normals_device(ctx, nen):
    """Layered normals on the device (fs_model_normals_layered): one layup
    against the oracle's constant-csys normals, and two layups with a tilted
    csys (so that the sum and the validity both vary) against the per-element
    loop of associategeometry! on the host (generic csys closure path)."""
    print(f"[gpu] layered normals nen={nen}", flush=True)
    from flexshell import workloads as W
    xyz, conn = SC.distorted_mesh(nen, 6, 5, seed=11, warp=0.6)
    cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
    lay_a = W.c5_layup(3, [0.0, 45.0, 90.0])
    f = cls(conn, lay_a, ctx=ctx)
    f.associategeometry(xyz)
    nrm, val = O.csys_normals(nen, xyz, conn, lay_a.csys)
    assert np.allclose(f._normals, nrm, rtol=0, atol=1e-14)
    assert np.array_equal(f._normal_valid, val)
    lay_b = W.c5_layup(4, [0.0, -45.0, 45.0, 90.0])
    c, s = np.cos(0.9), np.sin(0.9)
    lay_b.csys = np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=float)
    E = conn.shape[0]
    f2 = cls(conn, None, layup_groups=[(lay_a, np.arange(0, E, 2)), (lay_b, np.arange(1, E, 2))], ctx=ctx)
    f2.associategeometry(xyz)
    n_loop, v_loop = FM._ShellFEMM._host_csys_normals(f2, xyz, f2._csys_normal_function())
    assert np.allclose(f2._normals, n_loop, rtol=0, atol=1e-14)
    assert np.array_equal(f2._normal_valid, v_loop)
    assert (~v_loop).any() and v_loop.any()
    f._model.close()
    f2._model.close()


def test_fv12_frequencies_gpu(ctx):
    print("[gpu] fv12", flush=True)
    c = SC.fv12()
    d = c["dchi"]
    f = _homog_femm(ctx, 3, c["conn"], c["E"], c["nu"], c["rho"], c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    K = f.stiffness(c["xyz"], None, None, d)
    M = f.mass(c["xyz"], d)
    sh = (0.5 * 2 * np.pi) ** 2
    w = sla.eigsh(K + sh * M, k=14, M=M, sigma=0, which="LM", return_eigenvectors=False)
    fs = np.sqrt(np.abs(np.sort(w) - sh)) / (2 * np.pi)
    np.testing.assert_allclose(fs[6:], SC.FV12_REF, rtol=1e-9)
    assert np.all(fs[:6] < 1e-4)


@pytest.mark.parametrize("nplies,axes", [(10, (1, 2, 3)), (3, (2, -1, 3))])
def test_nayak_gpu(ctx, nplies, axes):
    print(f"[gpu] nayak {nplies} {axes}", flush=True)
    c = SC.nayak(nplies, axes)
    d = c["dchi"]
    lay = _flexshell_layup(c, rho=c["rho"])
    f = FM.FEMMShellT3FFComp(c["conn"], lay, ctx=ctx)
    f.associategeometry(c["xyz"])
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), c["xyz"], None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), c["xyz"], d)
    w = sl.eigh(K.toarray(), M.toarray(), eigvals_only=True, subset_by_index=[0, 8])
    nd = c["th"] * np.sqrt(c["rho"] / c["C11"]) * np.sqrt(w)
    assert np.linalg.norm(SC.NAYAK_REF - nd) < 1e-12


@pytest.mark.parametrize("nen,composite", [(3, False), (4, False), (3, True), (4, True)])
def test_lumped_mass(ctx, nen, composite):
    print(f"[gpu] mass nen={nen} composite={composite}", flush=True)
    xyz, conn = SC.distorted_mesh(nen, 6, 6, seed=9, warp=0.4)
    d = fe.DofField(xyz.shape[0])
    d.setebc([0, 3, 17], [1, 4])
    d.numberdofs()
    if composite:
        c = SC.nayak(4, (1, 2, 3))
        lay = _flexshell_layup(c, rho=c["rho"])
        cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
        f = cls(conn, lay, ctx=ctx)
        md, mi = lay.laminate_inertia()
        tm, rm = O.element_masses(nen, xyz, conn, m0=md, m1=mi)
    else:
        f = _homog_femm(ctx, nen, conn, 70e9, 0.3, 2700.0, 0.02)
        tm, rm = O.element_masses(nen, xyz, conn, rho=2700.0, t=0.02)
    f.associategeometry(xyz)
    for block, nrows in ((0, d.nalldofs), (1, d.nfreedofs)):
        asm = FM.SysmatAssemblerSparseDiag() if block == 0 else FM.SysmatAssemblerFFBlock(
            FM.SysmatAssemblerSparseDiag(), nrows, nrows)
        Mg = f.mass(asm, xyz, d).diagonal()
        Mo = O.lumped_diag(nen, conn, d.dofnums, nrows, tm, rm)
        np.testing.assert_allclose(Mg, Mo, rtol=1e-14, atol=0)


def _c1_femm(ctx):
    from flexshell import workloads as W
    c = W.plate_c1(n=16)
    d = c["dchi"]
    m = c["material"]
    f = FM.FEMMShellT3FF(c["conn"], m, thickness=c["t"], ctx=ctx)
    f.associategeometry(c["xyz"])
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), c["xyz"], None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), c["xyz"], d)
    return c, f, K, M


def test_explicit_matches_oracle(ctx):
    """Central-difference loop with damping and a time-varying load against
    the oracle's restatement of _cd_loop! on the same K and M."""
    print("[gpu] explicit vs oracle", flush=True)
    c, f, K, M = _c1_femm(ctx)
    d = c["dchi"]
    nf = d.nfreedofs
    rng = np.random.default_rng(3)
    U0 = rng.normal(0, 1e-6, nf)
    V0 = rng.normal(0, 1e-3, nf)
    Fmag = rng.normal(0, 1.0, nf)
    om = explicit.omega_max(f, rng.random(nf), maxit=50)
    dt = explicit.stable_dt(om)
    nsteps = 300
    tt = explicit.sample_times(nsteps, dt)
    g = np.sin(2 * np.pi * 3000 * tt)
    cdamp = 2 * 0.02 * 2 * np.pi * 1000
    probes = np.array([0, 7, nf // 2, nf - 1])
    r = explicit.run(f._model, cdamp, U0, V0, dt, nsteps, Fmag, g, probes, 1)
    Uo, Vo, po = O.cd_loop(K, M.diagonal(), cdamp, dt, nsteps, Fmag, g, U0, V0, probes)
    assert np.linalg.norm(r["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo)
    assert np.linalg.norm(r["V"] - Vo) <= 1e-10 * np.linalg.norm(Vo)
    assert np.abs(r["probes"] - po).max() <= 1e-10 * np.abs(po).max()
    # bitwise reproducible
    r2 = explicit.run(f._model, cdamp, U0, V0, dt, nsteps, Fmag, g, probes, 1)
    assert np.array_equal(r["U"], r2["U"]) and np.array_equal(r["V"], r2["V"])


def test_omega_max_matches_oracle(ctx):
    print("[gpu] omega_max", flush=True)
    c, f, K, M = _c1_femm(ctx)
    v0 = np.random.default_rng(1).random(c["dchi"].nfreedofs)
    og = explicit.omega_max(f, v0, maxit=30)
    oo = O.pwr(K, M.diagonal(), v0, maxit=30)
    assert abs(og - oo) < 1e-10 * oo


def test_stiffness_bitwise_reproducible(ctx):
    print("[gpu] reproducible", flush=True)
    c = SC.scordelis(4, 12)
    f = _homog_femm(ctx, 4, c["conn"], c["E"], c["nu"], 0.0, c["t"])
    f.associategeometry(c["xyz"])
    K1 = f.stiffness(c["xyz"], None, None, c["dchi"])
    K2 = f.stiffness(c["xyz"], None, None, c["dchi"])
    assert np.array_equal(K1.data, K2.data)


def test_isolated_node_and_errors(ctx):
    """A node used by no element keeps empty rows (as the COO assembly gives);
    a degenerate quadrilateral raises the reference's singular-metric error;
    a repeated node is rejected."""
    print("[gpu] edge cases", flush=True)
    xy, conn = fe.T3block(1.0, 1.0, 2, 2)
    xyz = np.vstack([fe.xyz3(xy), [[5.0, 5.0, 0.0]]])
    d = fe.DofField(xyz.shape[0]).numberdofs()
    f = _homog_femm(ctx, 3, conn, 70e9, 0.3, 0.0, 0.01)
    f.associategeometry(xyz)
    K = f.stiffness(xyz, None, None, d)
    Ko = O.assemble_csr(3, conn, d.dofnums, d.nalldofs, _oracle_homog(3, xyz, conn, 70e9, 0.3, 0.01))
    assert np.array_equal(K.indptr, Ko.indptr) and np.array_equal(K.indices, Ko.indices)
    assert K[-6:].nnz == 0
    # collapsed quadrilateral (all four nodes at one point): the metric is
    # exactly singular, which the reference reports (detG ≈ 0.0 with atol 0,
    # FEMMShellQ4RSModule.jl:307-310)
    xyzq = np.zeros((4, 3))
    fq = _homog_femm(ctx, 4, np.array([[0, 1, 2, 3]]), 70e9, 0.3, 0.0, 0.01)
    fq.associategeometry(xyzq)
    dq = fe.DofField(4).numberdofs()
    with pytest.raises(Exception):
        fq.stiffness(xyzq, None, None, dq)
    with pytest.raises(Exception):
        FM.FEMMShellT3FF(np.array([[0, 1, 1]]), MatDeforElastIso(1.0, 0.3), thickness=0.1,
                         ctx=ctx).associategeometry(xyzq[:3])
