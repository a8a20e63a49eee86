"""Pin the CPU oracle against the reference's own test values, known answers
and closed forms (CPU only).  A mistake in the oracle -- a dropped term, a
wrong sign or index, a transposed operand -- moves at least one of these."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import shell_cases as SC
from flexshell import fe
from oracle import oracle as O


def _homog_K(case, nrows=None):
    Dps, Dt = O.shell_D(O.iso_D6(case["E"], case["nu"]))
    nrm, val = O.normals(case["nen"], case["xyz"], case["conn"])
    Ke = O.element_matrices(case["nen"], case["xyz"], case["conn"], nrm, val, Dps=Dps, Dt=Dt, t=case["t"],
                            stab_c=case["stab_c"])
    d = case["dchi"]
    return O.assemble_csr(case["nen"], case["conn"], d.dofnums, nrows or d.nalldofs, Ke)


def _layered_K(case, nrows=None):
    A, B, D, H, inert = O.laminate(case["plies"])
    nrm, val = O.csys_normals(case["nen"], case["xyz"], case["conn"], case["csys"])
    Ke = O.element_matrices(case["nen"], case["xyz"], case["conn"], nrm, val,
                            groups=[(A, B, D, H, case["csys"], case["th"])])
    d = case["dchi"]
    return O.assemble_csr(case["nen"], case["conn"], d.dofnums, nrows or d.nalldofs, Ke), inert


def test_fv12_frequencies():
    """T3FF stiffness + lumped mass: the 8 elastic FV12 frequencies of
    test/test_shell_dynamics.jl:113-132 (reference tolerance rtol 1e-6)."""
    c = SC.fv12()
    K = _homog_K(c)
    tm, rm = O.element_masses(3, c["xyz"], c["conn"], rho=c["rho"], t=c["t"])
    d = c["dchi"]
    M = O.lumped_diag(3, c["conn"], d.dofnums, d.nalldofs, tm, rm)
    sh = (0.5 * 2 * np.pi) ** 2
    w = sla.eigsh(K + sh * sp.diags(M), k=14, M=sp.diags(M), sigma=0, which="LM", return_eigenvectors=False)
    fs = np.sqrt(np.abs(np.sort(w) - sh)) / (2 * np.pi)
    assert np.all(fs[:6] < 1e-4)  # six rigid-body modes
    np.testing.assert_allclose(fs[6:], SC.FV12_REF, rtol=1e-9)


@pytest.mark.parametrize("nen,n", [(3, 4), (3, 8), (3, 12), (4, 4), (4, 8), (4, 12)])
def test_scordelis_lo(nen, n):
    """T3FF / Q4RS on a curved roof with isoparametric nodal normals
    (test/test_shell_statics.jl:97-108, test/test_q4rs_shell_statics.jl:92-110)."""
    c = SC.scordelis(nen, n)
    K = _homog_K(c)
    d = c["dchi"]
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), c["F"][:nf])
    w = d.scattersysvec(U)[c["probe"], 2]
    ref = (SC.SCORDELIS_T3 if nen == 3 else SC.SCORDELIS_Q4)[n]
    assert abs(w / -0.3024 * 100 - ref) / ref < 1e-8


@pytest.mark.parametrize("nen", [3, 4])
def test_barbero_layered_statics(nen):
    """T3FFComp / Q4RSComp with membrane-bending coupling (B != 0) under edge
    loads (test/test_composite_shell_statics.jl:97, ..._q4rs.jl:97, `≈`)."""
    c = SC.barbero(nen)
    K, _ = _layered_K(c)
    d = c["dchi"]
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), c["F"][:nf])
    wmax = d.scattersysvec(U)[:, 2].max() / SC.mm
    ref = SC.BARBERO_T3 if nen == 3 else SC.BARBERO_Q4
    assert abs(wmax - ref) / ref < 1e-10


@pytest.mark.parametrize("nplies,axes", [(10, (1, 2, 3)), (3, (2, -1, 3)), (5, (-1, -2, 3)), (4, (-2, 1, 3))])
def test_nayak_layered_frequencies(nplies, axes):
    """T3FFComp + laminate inertia with rotated layup csys: the reference's own
    abs 1e-13 check (test/test_composite_shell_dynamics.jl:432-443)."""
    c = SC.nayak(nplies, axes)
    K, (md, mi) = _layered_K(c)
    tm, rm = O.element_masses(3, c["xyz"], c["conn"], m0=md, m1=mi)
    d = c["dchi"]
    M = O.lumped_diag(3, c["conn"], d.dofnums, d.nalldofs, tm, rm)
    nf = d.nfreedofs
    w = sl.eigh(K[:nf, :nf].toarray(), np.diag(M[:nf]), eigvals_only=True, subset_by_index=[0, 8])
    nd = c["th"] * np.sqrt(c["rho"] / c["C11"]) * np.sqrt(w)
    assert np.linalg.norm(SC.NAYAK_REF - nd) < 1.0e-13


def test_layered_equals_homogeneous_for_isotropic_ply():
    """T3FFComp with one isotropic ply equals T3FF (A = t Dps, D = t^3/12 Dps,
    H = 5/6 t G for the Vinson-Sierakowski model)."""
    xyz, conn = SC.distorted_mesh(3, 4, 3, seed=1, warp=0.3)
    E, nu, t = 70e9, 0.3, 0.01
    D6 = O.iso_D6(E, nu)
    Dps, Dt = O.shell_D(D6)
    nrm, val = O.normals(3, xyz, conn)
    K1 = O.element_matrices(3, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=t)
    A, B, D, H, _ = O.laminate([(D6, t, 0.0, 0.0)])
    K2 = O.element_matrices(3, xyz, conn, nrm, val, groups=[(A, B, D, H, np.eye(3), t)])
    assert np.linalg.norm(K1 - K2) < 1e-13 * np.linalg.norm(K1)


def test_layup_known_answers():
    """test/test_composite_layup.jl:321-328 (isotropic ply, any angle),
    :431-451 (Barbero Ex 5.7), :594 (cross-ply B = 0)."""
    E, nu, t = 200e9, 0.3, 0.005
    Atrue = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]]) * t
    for ang in (0.0, 47.0, 90.0, 129.0, 180.0):
        A, B, D, H, _ = O.laminate([(O.iso_D6(E, nu), t, ang, 0.0)])
        assert np.linalg.norm(A - Atrue) < 1e-15 * np.linalg.norm(Atrue)
        np.testing.assert_allclose(H, 5 / 6 * t * E / 2 / (1 + nu) * np.eye(2), rtol=1e-14, atol=1e-6)
    MPa = 1e6
    E1, E2, G12, nu12 = 67192 * MPa, 12139 * MPa, 7638 * MPa, 0.365
    D6 = O.ortho_D6(E1, E2, E2, nu12, nu12, nu12, G12, G12, G12)
    Dps, _ = O.ply_matrices(D6)
    ref = np.array([[68849.10243290635, 4540.006665496835, 0.0], [4540.006665496834, 12438.374426018723, 0.0],
                    [0.0, 0.0, 7637.999999999999]])
    assert np.linalg.norm(Dps / MPa - ref) < 1e-15 * E2
    A, _, _, _, _ = O.laminate([(D6, 1.0, 45.0, 0.0)])
    refA = np.array([[30229.872547479692, 14953.872547479687, 14102.68200172191],
                     [14953.872547479687, 30229.872547479685, 14102.682001721907],
                     [14102.682001721909, 14102.682001721909, 18051.865881982852]])
    assert np.linalg.norm(A / MPa - refA) < 1e-15 * E2
    D6l = O.lamina_D6(E1, E2, nu12, G12, G12, G12)
    plies = [(D6l, 0.001, a, 0.0) for a in (0, 90, 0, 90, 0)]
    _, B, _, _, _ = O.laminate(plies)
    assert np.linalg.norm(B) < 1e-15 * E2


def test_assembly_matches_dense_sum():
    """COO -> CSR with duplicate summation against a brute-force dense sum,
    with fixed dofs interleaved (FF and ALL blocks)."""
    rng = np.random.default_rng(7)
    xy, conn = fe.T3block(1.0, 1.0, 3, 2)
    N = xy.shape[0]
    d = fe.DofField(N)
    d.setebc([0, 5], [1, 3, 6])
    d.numberdofs(perm=rng.permutation(N))
    Ke = rng.standard_normal((conn.shape[0], 18, 18))
    dense = np.zeros((d.nalldofs, d.nalldofs))
    for e in range(conn.shape[0]):
        dofs = d.dofnums[conn[e]].ravel()
        dense[np.ix_(dofs, dofs)] += Ke[e]
    for nrows in (d.nalldofs, d.nfreedofs):
        K = O.assemble_csr(3, conn, d.dofnums, nrows, Ke)
        np.testing.assert_allclose(K.toarray(), dense[:nrows, :nrows], rtol=0, atol=1e-12)
        # structural union of element dof pairs, zeros retained
        pattern = np.zeros((nrows, nrows), dtype=bool)
        for e in range(conn.shape[0]):
            dofs = d.dofnums[conn[e]].ravel()
            dofs = dofs[dofs < nrows]
            pattern[np.ix_(dofs, dofs)] = True
        assert K.nnz == pattern.sum()


@pytest.mark.parametrize("nen", [3, 4])
def test_element_rigid_body_null_space(nen):
    """A free flat element has the rigid translations and the rigid rotations
    about in-plane axes in its null space.  The in-plane rotation (about z)
    is resisted by the artificial drilling stiffness only (kavg on the drilling
    dof, FEMMShellT3FFModule.jl:714-727; FEMMShellQ4RSModule.jl:840-859), so its
    energy is bounded by that penalty, which is of bending scale."""
    xyz, conn = SC.distorted_mesh(nen, 1, 1, seed=3, warp=0.0)
    conn = conn[:1]
    used = np.unique(conn)
    nrm, val = O.normals(nen, xyz, conn)
    Dps, Dt = O.shell_D(O.iso_D6(70e9, 0.33))
    Ke = O.element_matrices(nen, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=0.01)[0]
    X = xyz[conn[0]]
    modes = []
    for k in range(3):
        u = np.zeros((nen, 6))
        u[:, k] = 1.0
        modes.append(u.ravel())
        th = np.zeros(3)
        th[k] = 1.0
        r = np.zeros((nen, 6))
        r[:, :3] = np.cross(th, X)
        r[:, 3:] = th
        modes.append(r.ravel())
    for i, m in enumerate(modes):
        r = np.linalg.norm(Ke @ m) / (np.linalg.norm(Ke) * np.linalg.norm(m))
        if i == 5:  # rotation about z
            assert r < 1e-4
        else:
            assert r < 1e-9
    assert used.size == nen


def test_normals_flat_and_crease():
    """Flat plate: normals +z, all valid.  Folded plate (90 degree crease):
    nodes on the fold are marked invalid (threshold 30 degrees)."""
    xy, conn = fe.T3block(1.0, 1.0, 4, 4)
    xyz = fe.xyz3(xy)
    n, v = O.normals(3, xyz, conn)
    np.testing.assert_allclose(n, np.tile([0, 0, 1.0], (xyz.shape[0], 1)), atol=1e-15)
    assert v.all()
    fold = xyz.copy()
    m = fold[:, 0] > 0.5
    fold[m, 2] = fold[m, 0] - 0.5
    fold[m, 0] = 0.5
    for nen in (3, 4):
        c = conn if nen == 3 else fe.Q4block(1.0, 1.0, 4, 4)[1]
        n, v = O.normals(nen, fold, c)
        on_fold = np.isclose(fold[:, 0], 0.5) & np.isclose(fold[:, 2], 0.0)
        assert (~v[on_fold]).all() and v[~on_fold].all()


@pytest.mark.parametrize("nen", [3, 4])
def test_lumped_mass_total(nen):
    """Translational lumped mass sums to rho t area; rotary to rho t^3/12 area."""
    xyz, conn = SC.distorted_mesh(nen, 5, 4, seed=2, warp=0.0)
    rho, t = 2700.0, 0.02
    tm, rm = O.element_masses(nen, xyz, conn, rho=rho, t=t)
    # area of the jittered (flat) mesh, independently: triangles, or quads split in two
    tri = conn if nen == 3 else np.concatenate([conn[:, [0, 1, 2]], conn[:, [0, 2, 3]]])
    X = xyz[tri]
    area = 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1).sum()
    assert abs(tm.sum() * nen - rho * t * area) < 1e-12 * rho * t * area
    assert abs(rm.sum() * nen - rho * t ** 3 / 12 * area) < 1e-12 * rho * t ** 3 / 12 * area


def test_cd_loop_single_dof_closed_form():
    """Undamped CD recurrence of one dof from an initial displacement.  The
    reference starts from A0 = invMC * F(0) (plate_expl_examples.jl:80-81),
    i.e. without the elastic force, so u1 = u0 and thereafter
    u_{n+1} = 2 u_n - u_{n-1} - (omega dt)^2 u_n, whose solution is
    u_n = u0 cos(n theta - theta/2) / cos(theta/2), cos(theta) = 1 - (omega dt)^2/2."""
    k, m, u0 = 4.0e6, 2.0, 1e-3
    omega = np.sqrt(k / m)
    dt = 0.5 * 2 / omega
    nsteps = 1000
    U, V, pr = O.cd_loop(sp.csr_matrix([[k]]), np.array([m]), 0.0, dt, nsteps, np.zeros(1), np.zeros(nsteps + 1),
                         np.array([u0]), np.zeros(1), probes=[0])
    th = np.arccos(1 - (omega * dt) ** 2 / 2)
    n = np.arange(nsteps + 1)
    np.testing.assert_allclose(pr[:, 0], u0 * np.cos(n * th - th / 2) / np.cos(th / 2), rtol=0, atol=1e-12)


def _cd_closed_form(m, c, k, dt, f, u_a, u_b, n_a, n, dt_mass=None):
    """Single-dof solution of the central-difference recurrence.  With the
    velocity update V_{n+1} = V_n + dt/2 (A_n + A_{n+1}) the scheme is
    Newmark (beta 0, gamma 1/2), so u_{n+1} - 2 u_n + u_{n-1} = dt^2 a_n and
    v_n = (u_{n+1} - u_{n-1}) / (2 dt); the step's force balance
    A_{n+1} = (f - k U_{n+1} - c (V_n + dt/2 A_n)) / (m + dt_m/2 c) then reads
    m' a_n + c v_n + k u_n = f with m' = m + (dt_m - dt) c / 2.  That
    three-term recurrence, fitted to u at steps n_a - 1 = u_a and n_a = u_b,
    has the solution f/k + r^j (P cos j theta + Q sin j theta), j = n - n_a,
    from the roots r e^(+-i theta) of alpha z^2 + beta z + gamma."""
    dt_m = dt if dt_mass is None else dt_mass
    mp = m + (dt_m - dt) * c / 2
    al, be, ga = mp / dt ** 2 + c / (2 * dt), k - 2 * mp / dt ** 2, mp / dt ** 2 - c / (2 * dt)
    r = np.sqrt(ga / al)
    th = np.arccos(-be / (2 * np.sqrt(al * ga)))
    up = f / k
    P = u_b - up
    # u_a = up + r^-1 (P cos th - Q sin th)
    Q = (P * np.cos(th) - r * (u_a - up)) / np.sin(th)
    j = np.asarray(n, dtype=float) - n_a
    return up + r ** j * (P * np.cos(j * th) + Q * np.sin(j * th))


def test_cd_loop_damped_free_decay():
    """Damped free vibration of one dof (C = ksi 2 omega M, plate_expl_examples.jl:
    69-71) from an initial displacement and velocity.  The damping force enters
    as C (V + dt/2 A) and invMC = 1/(M + dt/2 C) (:86-89), which makes the
    step an exact Newmark force balance; its closed form fixes the decay per
    step r^2 = (m/dt^2 - c/(2 dt)) / (m/dt^2 + c/(2 dt)) and the phase, so
    damping taken at V_n, or left out of invMC, fails."""
    k, m, u0, v0 = 4.0e6, 2.0, 1e-3, 0.3
    omega = np.sqrt(k / m)
    ksi = 0.05
    cdamp = ksi * 2 * omega
    dt = 0.5 * 2 / omega * (np.sqrt(1 + ksi ** 2) - ksi)
    nsteps = 1000
    U, V, pr = O.cd_loop(sp.csr_matrix([[k]]), np.array([m]), cdamp, dt, nsteps, np.zeros(1), np.zeros(nsteps + 1),
                         np.array([u0]), np.array([v0]), probes=[0])
    u1 = u0 + dt * v0  # the reference's A0 = invMC F(0) = 0 (force only, :80-81)
    n = np.arange(nsteps + 2)
    u = _cd_closed_form(m, cdamp * m, k, dt, 0.0, u0, u1, 1, n)  # fitted at steps 0 and 1
    scale = np.abs(pr[:, 0]).max()
    np.testing.assert_allclose(pr[:, 0], u[:nsteps + 1], rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(V[0], (u[nsteps + 1] - u[nsteps - 1]) / (2 * dt), rtol=1e-8)
    assert abs(pr[-1, 0]) < 0.05 * u0  # it decays


@pytest.mark.parametrize("ksi", [0.0, 0.02])
def test_cd_loop_step_load(ksi):
    """A load switched on at step n0 (g[n] = 1 for n >= n0; F(t_n) = g[n] Fmag,
    force! called after the displacement update, :84-85) on a dof at rest:
    u stays 0 up to and including step n0 and then follows the closed form of
    the forced recurrence about the static value F/k.  Reading the load of
    the previous step (g[n-1]) delays the response by one step and fails."""
    k, m, F0 = 9.0e6, 3.0, 250.0
    omega = np.sqrt(k / m)
    cdamp = ksi * 2 * omega
    dt = 0.8 * 2 / omega * (np.sqrt(1 + ksi ** 2) - ksi)
    nsteps, n0 = 600, 7
    g = (np.arange(nsteps + 1) >= n0).astype(float)
    U, V, pr = O.cd_loop(sp.csr_matrix([[k]]), np.array([m]), cdamp, dt, nsteps, np.array([F0]), g, np.zeros(1),
                         np.zeros(1), probes=[0])
    n = np.arange(nsteps + 1)
    u = np.where(n <= n0, 0.0, _cd_closed_form(m, cdamp * m, k, dt, F0, 0.0, 0.0, n0, n))
    np.testing.assert_allclose(pr[:, 0], u, rtol=0, atol=1e-10 * F0 / k)
    assert pr[n0 + 1, 0] > 0.0


def test_cd_loop_adjusted_dt_keeps_invmc_dt():
    """_cd_loop! forms invMC from the dt it is called with (:70-71) and only
    then re-derives dt = tend / (nsteps + 1) when round(tend/dt) dt < tend
    (:72-75).  With damping the loop therefore runs at the new dt with the
    old dt in invMC: the force balance carries m' = m + (dt_in - dt) c / 2,
    which the closed form resolves from the single-dt one."""
    k, m, u0 = 2.0e6, 1.0, 1e-3
    omega = np.sqrt(k / m)
    ksi = 0.05
    dt_in, tend = 1e-4, 0.01234
    U, V, pr, dt, nsteps = O.cd_loop_tend(sp.csr_matrix([[k]]), np.array([m]), ksi, omega, np.array([u0]),
                                          np.zeros(1), tend, dt_in, np.zeros(1), lambda t: 0.0, probes=[0])
    assert nsteps == 123 and dt == tend / 124
    c = ksi * 2 * omega * m
    n = np.arange(nsteps + 1)
    u = _cd_closed_form(m, c, k, dt, 0.0, u0, u0, 1, n, dt_mass=dt_in)
    np.testing.assert_allclose(pr[:, 0], u, rtol=0, atol=1e-11 * u0)
    u_one_dt = _cd_closed_form(m, c, k, dt, 0.0, u0, u0, 1, n)
    assert np.abs(u_one_dt - u).max() > 1e-8 * u0  # the two forms are told apart


def test_pwr_largest_eigenvalue():
    n = 200
    kd = np.ones(n)
    kd[17] = 10.0
    md = np.full(n, 2.0)
    v0 = np.random.default_rng(0).random(n)
    om = O.pwr(sp.diags(kd).tocsr(), md, v0, maxit=60)
    assert abs(om - np.sqrt(10.0 / 2.0)) < 1e-12


@pytest.mark.parametrize("mesh", ["1x9", "3x18", "5x36", "10x72"])
def test_raasch_hook_from_abaqus_meshes(mesh):
    """T3FF on the reference's Abaqus meshes (raasch_s4_*.inp read by
    flexshell.meshio, compacted, split by Q4toT3): the tip deflection in
    percent of 5.02 of test/test_shell_statics.jl:226-234 (rtol 1e-4 there)."""
    c = SC.raasch(mesh)
    d = c["dchi"]
    K = _homog_K(c)
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), c["F"][:nf])
    pct = d.scattersysvec(U)[c["probe"], 2] / 5.02 * 100
    assert abs(pct - SC.RAASCH_T3[mesh]) / SC.RAASCH_T3[mesh] < 1e-9


@pytest.mark.parametrize("nen", [3, 4])
@pytest.mark.parametrize("case", ["c1", "c2", "c6", "c7", "c8+", "c8-", "c9+", "c9-", "c10w", "c10u"])
def test_composite_reference_values(case, nen):
    """T3FFComp / Q4RSComp plates of test/test_composite_shell_statics.jl and
    test_composite_shell_statics_q4rs.jl (layups of 2 to 20 plies at 0, 90,
    +-45 degrees; edge and pressure loads; a plate and layup csys rotated in
    3-D): the values of their @test lines."""
    c = SC.comp_t3(case, nen)
    d = c["dchi"]
    K, _ = _layered_K(c)
    F = c["F"].copy()
    if c["pressure"] is not None:
        F += O.surface_load(nen, c["xyz"], c["conn"], c["pressure"], d.nalldofs, d.dofnums, c["load_npts"])
    comp, sign, scale = c["quantity"]
    v = SC._csr_solve_max(K, F, d, comp, sign) * scale
    assert abs(v - c["ref"]) / abs(c["ref"]) < c["rtol"], (v, c["ref"])


@pytest.mark.parametrize("nen", [3, 4])
@pytest.mark.parametrize("which", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_twisted_beam(nen, which, n):
    """Twisted cantilever of test_shell_statics.jl:350-386 (T3FF) and
    test_q4rs_shell_statics.jl:346-382 (Q4RS), thick and thin, both load
    directions, four meshes: the reference percentages (rtol 1e-3 there).
    Thick section (t = 0.32): within 1e-8.  Thin section (t = 0.0032 on a
    12-long beam): the stiffness condition number grows like (L/t)^2 ~ 1e7, so
    rounding-level differences of K and of the sparse solver (not the
    reference's) show at up to 2e-5 on the finest mesh; held to 1e-4."""
    c = SC.twisted_beam(nen, which, n)
    d = c["dchi"]
    Dps, Dt = O.shell_D(O.iso_D6(c["E"], c["nu"]))
    nrm, val = O.normals(nen, c["xyz"], c["conn"])
    Ke = O.element_matrices(nen, c["xyz"], c["conn"], nrm, val, Dps=Dps, Dt=Dt, t=c["t"], stab_c=c["stab_c"])
    K = O.assemble_csr(nen, c["conn"], d.dofnums, d.nalldofs, Ke)
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), c["F"][:nf])
    v = d.scattersysvec(U)[c["probe"], c["dir"] - 1] / c["uex"] * 100
    assert abs(v - c["ref"]) / c["ref"] < (1e-8 if which < 2 else 1e-4), (v, c["ref"])


def test_le5_z_cantilever():
    """LE5 Z-section cantilever of test_shell_statics.jl:438-531 on the
    reference's folded Abaqus mesh (nodal normals invalid along the folds):
    the minimum and maximum deflection of its @test lines."""
    c = SC.le5()
    d = c["dchi"]
    K = _homog_K(c)
    nf = d.nfreedofs
    w = d.scattersysvec(sla.spsolve(K[:nf, :nf].tocsc(), c["F"][:nf]))[:, 2]
    assert abs(w.min() - SC.LE5_W[0]) / abs(SC.LE5_W[0]) < 1e-9, w.min()
    assert abs(w.max() - SC.LE5_W[1]) / abs(SC.LE5_W[1]) < 1e-9, w.max()


@pytest.mark.parametrize("composite", [False, True])
def test_petyt_plate_frequencies(composite):
    """mcompshelldyn0 of test_composite_shell_dynamics.jl: the 8 lowest
    frequencies of the homogeneous T3FF plate and of the [45/-45/45/-45]
    T3FFComp plate of isotropic plies (reference bar: 1e-3 of the norm)."""
    c = SC.petyt(composite)
    d = c["dchi"]
    nf = d.nfreedofs
    if composite:
        K, (m0, m1) = _layered_K(c)
        tm, rm = O.element_masses(3, c["xyz"], c["conn"], m0=m0, m1=m1)
    else:
        K = _homog_K(c)
        tm, rm = O.element_masses(3, c["xyz"], c["conn"], rho=c["rho"], t=c["t"])
    M = O.lumped_diag(3, c["conn"], d.dofnums, d.nalldofs, tm, rm)
    w = sl.eigh(K[:nf, :nf].toarray(), np.diag(M[:nf]), eigvals_only=True, subset_by_index=[0, 7])
    fs = np.sqrt(w) / (2 * np.pi)
    assert np.linalg.norm(fs - SC.PETYT_FS) < 1e-9 * np.linalg.norm(fs), fs
