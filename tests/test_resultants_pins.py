"""Pin the oracle's stress-resultant recovery (inspectintegpoints,
FEMMShellT3FFModule.jl:850-962, FEMMShellQ4RSModule.jl:1061-1170 and the
layered forms) on the CPU: closed-form membrane, bending and shear patches,
rigid translations, and the energy identity u'K u = sum w (N'A^-1 N + ...)
that ties the recovery to the element stiffness pinned in
test_oracle_pins.py.  The reference's own tests have no shell resultant
values, so these closed forms and identities are what fixes the oracle."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import fe
from oracle import oracle as O

G2 = 0.577350269189626
QP = [(-G2, -G2), (-G2, G2), (G2, -G2), (G2, G2)]
E_, NU, T_ = 70e9, 0.3, 0.01


def _mat():
    return O.shell_D(O.iso_D6(E_, NU))


def _flat(nen, nx=4, ny=3, seed=1, rot=None, jitter=0.2):
    xyz, conn = SC.distorted_mesh(nen, nx, ny, seed=seed, warp=0.0, jitter=jitter)
    if rot is not None:
        xyz = xyz @ rot.T
    return xyz, conn


def _rot3(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def _q4_jac(X, xi, eta):
    gN = 0.25 * np.array([[-(1 - eta), -(1 - xi)], [(1 - eta), -(1 + xi)], [(1 + eta), (1 + xi)],
                          [-(1 + eta), (1 - xi)]])
    return X.T @ gN


def _weights(nen, xyz, conn):
    """Integration weight x surface Jacobian per point (nelem x npts)."""
    if nen == 3:
        X = xyz[conn]
        return 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)[:, None]
    w = np.zeros((conn.shape[0], 4))
    for e, c in enumerate(conn):
        for q, (xi, eta) in enumerate(QP):
            J = _q4_jac(xyz[c], xi, eta)
            w[e, q] = np.linalg.norm(np.cross(J[:, 0], J[:, 1]))
    return w


def _stab(nen, xyz, conn, t, c):
    X = xyz[conn]
    if nen == 3:
        h2 = np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)  # 2 Ae
    else:
        h2 = np.max(np.sum((X[:, 1:] - X[:, :1]) ** 2, axis=2), axis=1)
    return t * t / (t * t + c * h2)


def _stab_c(nen):
    return 5 / 12 / 1.5 if nen == 3 else 0.1


def _run(nen, xyz, conn, u, **kw):
    nrm, val = O.normals(nen, xyz, conn)
    Dps, Dt = _mat()
    return O.resultants(nen, xyz, conn, nrm, val, u, Dps=Dps, Dt=Dt, t=T_, **kw)


@pytest.mark.parametrize("nen", [3, 4])
@pytest.mark.parametrize("tilted", [False, True])
def test_membrane_patch(nen, tilted):
    """u = G x (in-plane, constant gradient): N = t Dps [exx, eyy, gxy] in the
    plane's own axes, M = Q = 0.  Tilted: the same patch on a plane rotated
    in 3-D, reported in the rotated axes."""
    R = _rot3(0.4, -0.7, 0.3) if tilted else np.eye(3)
    xyz0, conn = _flat(nen)
    xyz = xyz0 @ R.T
    Gm = np.array([[1.3e-4, -0.4e-4], [0.9e-4, -0.6e-4]])
    u = np.zeros((xyz.shape[0], 6))
    u2 = xyz0[:, :2] @ Gm.T
    u[:, :3] = np.hstack([u2, np.zeros((xyz.shape[0], 1))]) @ R.T
    out = _run(nen, xyz, conn, u, ocs=R)
    Dps, _ = _mat()
    eps = np.array([Gm[0, 0], Gm[1, 1], Gm[0, 1] + Gm[1, 0]])
    Nref = T_ * Dps @ eps
    scale = np.abs(Nref).max()
    assert np.abs(out[..., :3] - Nref).max() < 1e-11 * scale
    assert np.abs(out[..., 3:]).max() < 1e-11 * scale * T_


@pytest.mark.parametrize("nen", [3, 4])
def test_bending_patch(nen):
    """w = (A x^2 + B y^2)/2 + C x y with theta_x = w,y and theta_y = -w,x
    (no shear): M = t^3/12 Dps [-A, -B, -2C], N = 0, Q = 0 in global axes."""
    xyz, conn = _flat(nen)
    A, B, Cc = 2.0e-3, -1.1e-3, 0.7e-3
    x, y = xyz[:, 0], xyz[:, 1]
    u = np.zeros((xyz.shape[0], 6))
    u[:, 2] = 0.5 * A * x * x + 0.5 * B * y * y + Cc * x * y
    u[:, 3] = B * y + Cc * x
    u[:, 4] = -(A * x + Cc * y)
    out = _run(nen, xyz, conn, u, ocs=np.eye(3))
    Dps, _ = _mat()
    Mref = T_ ** 3 / 12 * Dps @ np.array([-A, -B, -2 * Cc])
    scale = np.abs(Mref).max()
    assert np.abs(out[..., 3:6] - Mref).max() < 1e-9 * scale
    assert np.abs(out[..., 6:]).max() < 1e-9 * scale / T_
    assert np.abs(out[..., :3]).max() < 1e-9 * scale / T_


@pytest.mark.parametrize("nen", [3, 4])
def test_shear_patch(nen):
    """w = g1 x + g2 y, no rotations: Q = t stab(t, h) 5/6 G [g1, g2] with the
    element's own h, N = M = 0.  The Q4 MITC strains of the reference
    (Bathe-Dvorkin edge factors, FEMMShellQ4RSModule.jl:753-804) reproduce a
    constant shear exactly on rectangles, so the Q4 mesh is regular; the T3
    mesh is jittered."""
    xyz, conn = _flat(nen, jitter=0.2 if nen == 3 else 0.0)
    g = np.array([1.5e-4, -2.5e-4])
    u = np.zeros((xyz.shape[0], 6))
    u[:, 2] = g[0] * xyz[:, 0] + g[1] * xyz[:, 1]
    out = _run(nen, xyz, conn, u, ocs=np.eye(3))
    Gs = E_ / 2 / (1 + NU)
    st = _stab(nen, xyz, conn, T_, _stab_c(nen))
    Qref = (T_ * st * 5 / 6 * Gs)[:, None, None] * g[None, None, :]
    scale = np.abs(Qref).max()
    assert np.abs(out[..., 6:] - Qref).max() < 1e-11 * scale
    assert np.abs(out[..., :6]).max() < 1e-11 * scale


@pytest.mark.parametrize("nen", [3, 4])
def test_rigid_motions(nen):
    """Rigid translations and rotations give zero resultants on a flat mesh
    tilted in 3-D.  (On a curved mesh the nodal-to-element transformation
    mixes a node's translations through its own triad,
    FEMMShellT3FFModule.jl:421-463, so rigid motions leave small residual
    resultants there, in the reference as here.)"""
    R = _rot3(-0.5, 0.6, 0.2)
    xyz, conn = _flat(nen, rot=R)
    u = np.zeros((xyz.shape[0], 6))
    u[:, :3] = [1e-3, -2e-3, 0.5e-3]
    assert np.abs(_run(nen, xyz, conn, u)).max() < 1e-6
    w = np.array([0.3e-3, -0.2e-3, 0.5e-3])
    u = np.zeros((xyz.shape[0], 6))
    u[:, :3] = np.cross(w, xyz)
    u[:, 3:] = w
    assert np.abs(_run(nen, xyz, conn, u)).max() < 1e-6


@pytest.mark.parametrize("nen", [3, 4])
def test_energy_identity_homogeneous(nen):
    """u_e' K_e u_e (drilling scale 0, AVERAGE_B) equals
    sum_q w_q (N'(t Dps)^-1 N + M'(t^3/12 Dps)^-1 M + Q'(t stab 5/6 Dt)^-1 Q)
    per element on a warped, jittered mesh with random displacements."""
    xyz, conn = SC.distorted_mesh(nen, 5, 4, seed=4, warp=0.5)
    nrm, val = O.normals(nen, xyz, conn)
    Dps, Dt = _mat()
    Ke = O.element_matrices(nen, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=T_, drill=0.0)
    u = np.random.default_rng(3).normal(0, 1e-3, (xyz.shape[0], 6))
    out = O.resultants(nen, xyz, conn, nrm, val, u, Dps=Dps, Dt=Dt, t=T_)
    ue = u[conn].reshape(conn.shape[0], -1)
    ref = np.einsum("ei,eij,ej->e", ue, Ke, ue)
    w = _weights(nen, xyz, conn)
    st = _stab(nen, xyz, conn, T_, _stab_c(nen))
    Cm = np.linalg.inv(T_ * Dps)
    Cb = np.linalg.inv(T_ ** 3 / 12 * Dps)
    en = np.einsum("epi,ij,epj->ep", out[..., :3], Cm, out[..., :3])
    en += np.einsum("epi,ij,epj->ep", out[..., 3:6], Cb, out[..., 3:6])
    Cs = np.linalg.inv(5 / 6 * Dt)
    en += np.einsum("epi,ij,epj->ep", out[..., 6:], Cs, out[..., 6:]) / (T_ * st)[:, None]
    got = np.sum(w * en, axis=1)
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("nen", [3, 4])
def test_energy_identity_layered(nen):
    """Layered (T applied once): with the output csys equal to the layup csys
    the resultants are the unrotated laminate stiffnesses times the strains,
    so u_e'K_e u_e = sum_q w_q ([N;M]' ABD^-1 [N;M] + Q'(stab H)^-1 Q)."""
    xyz, conn = SC.distorted_mesh(nen, 5, 4, seed=5, warp=0.5)
    D6 = O.lamina_D6(133.86e9, 7.706e9, 0.301, 4.306e9, 4.306e9, 2.76e9)
    plies = [(D6, 0.002, 0.0, 0.0), (D6, 0.003, 35.0, 0.0), (D6, 0.002, -60.0, 0.0)]
    A, B, D, H, _ = O.laminate(plies)
    cs = _rot3(0.35, 0.0, 0.0)  # layup axes turned 20 degrees about z
    th = 0.007
    nrm, val = O.csys_normals(nen, xyz, conn, cs)
    groups = [(A, B, D, H, cs, th)]
    Ke = O.element_matrices(nen, xyz, conn, nrm, val, groups=groups, drill=0.0)
    u = np.random.default_rng(6).normal(0, 1e-3, (xyz.shape[0], 6))
    out = O.resultants(nen, xyz, conn, nrm, val, u, groups=groups, ocs=cs)
    ue = u[conn].reshape(conn.shape[0], -1)
    ref = np.einsum("ei,eij,ej->e", ue, Ke, ue)
    ABD = np.block([[A, B], [B, D]])
    Ci = np.linalg.inv(ABD)
    w = _weights(nen, xyz, conn)
    st = _stab(nen, xyz, conn, th, _stab_c(nen))
    en = np.einsum("epi,ij,epj->ep", out[..., :6], Ci, out[..., :6])
    en += np.einsum("epi,ij,epj->ep", out[..., 6:], np.linalg.inv(H), out[..., 6:]) / st[:, None]
    got = np.sum(w * en, axis=1)
    assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max()


def test_q4_layered_double_transform():
    """The layered Q4 form applies T twice.  On a flat grid rotated by theta
    about z every element has E_G = Rz(theta) and identity nodal triads, so
    T = blockdiag(Rz(-theta)) and the twice-transformed result is the
    once-transformed result of the displacements rotated by Rz(-theta)."""
    th_ = 0.37
    Rz = np.array([[np.cos(th_), -np.sin(th_), 0], [np.sin(th_), np.cos(th_), 0], [0, 0, 1]])
    xy, conn = fe.Q4block(1.0, 0.7, 4, 3)
    xyz = fe.xyz3(xy) @ Rz.T
    D6 = O.lamina_D6(133.86e9, 7.706e9, 0.301, 4.306e9, 4.306e9, 2.76e9)
    A, B, D, H, _ = O.laminate([(D6, 0.002, 0.0, 0.0), (D6, 0.003, 50.0, 0.0)])
    groups = [(A, B, D, H, np.eye(3), 0.005)]
    nrm, val = O.csys_normals(4, xyz, conn, np.eye(3))
    u = np.random.default_rng(7).normal(0, 1e-3, (xyz.shape[0], 6))
    twice = O.resultants(4, xyz, conn, nrm, val, u, groups=groups, twice=True)
    ur = u.copy()
    ur[:, :3] = u[:, :3] @ Rz
    ur[:, 3:] = u[:, 3:] @ Rz
    once = O.resultants(4, xyz, conn, nrm, val, ur, groups=groups, twice=False)
    assert np.abs(twice - once).max() < 1e-10 * np.abs(once).max()
    # with theta = 0 the two forms coincide (T = I)
    xyz0 = fe.xyz3(xy)
    a = O.resultants(4, xyz0, conn, nrm, val, u, groups=groups, twice=True)
    b = O.resultants(4, xyz0, conn, nrm, val, u, groups=groups, twice=False)
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()


def test_layered_equals_homogeneous_for_isotropic_ply():
    """One isotropic ply (A = t Dps, B = 0, D = t^3/12 Dps, H = 5/6 t G) gives the
    homogeneous resultants (T3, and Q4 with T applied once)."""
    D6 = O.iso_D6(E_, NU)
    A, B, D, H, _ = O.laminate([(D6, T_, 0.0, 0.0)])
    for nen in (3, 4):
        xyz, conn = SC.distorted_mesh(nen, 4, 3, seed=8, warp=0.3)
        nrm, val = O.normals(nen, xyz, conn)
        u = np.random.default_rng(9).normal(0, 1e-3, (xyz.shape[0], 6))
        h = _run(nen, xyz, conn, u, ocs=np.eye(3))
        c = O.resultants(nen, xyz, conn, nrm, val, u, groups=[(A, B, D, H, np.eye(3), T_)], ocs=np.eye(3))
        for s in (slice(0, 3), slice(3, 6), slice(6, 8)):
            assert np.abs(h[..., s] - c[..., s]).max() < 1e-11 * np.abs(h[..., s]).max()


def test_barrel_vault_resultant_fields():
    """The barrel vault of test_shell_statics.jl:578-728: minimum and maximum
    of the nodal moment, membrane and shear fields in the cylindrical csys,
    within the reference's rtol 0.01 (the only shell resultant values the
    reference's tests hold; its nodal interpolation is FinEtools'
    inverse-distance fieldfromintegpoints, restated by the test)."""
    import scipy.sparse.linalg as sla
    c = SC.barrel_vault()
    xyz, conn, d = c["xyz"], c["conn"], c["dchi"]
    Dps, Dt = O.shell_D(O.iso_D6(c["E"], c["nu"]))
    nrm, val = O.normals(3, xyz, conn)
    Ke = O.element_matrices(3, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=c["t"], stab_c=c["stab_c"], drill=c["drill"])
    K = O.assemble_csr(3, conn, d.dofnums, d.nalldofs, Ke)
    F = O.surface_load(3, xyz, conn, c["force"], d.nalldofs, d.dofnums, 3)
    nf = d.nfreedofs
    u = d.scattersysvec(sla.spsolve(K[:nf, :nf].tocsc(), F[:nf]))
    loc = xyz[conn].mean(axis=1)[:, None, :]
    ocs = np.array([c["csys"](p) for p in loc[:, 0]]).reshape(-1, 1, 3, 3)
    res = O.resultants(3, xyz, conn, nrm, val, u, Dps=Dps, Dt=Dt, t=c["t"], stab_c=c["stab_c"], ocs=ocs)
    off = {"membrane": 0, "moment": 3, "shear": 6}
    for q, refs in SC.BARREL_RESULTANTS.items():
        for nc, (lo, hi) in enumerate(refs):
            fl = SC.inverse_distance_field(xyz, conn, loc, res[:, :, off[q] + nc])
            assert abs(fl.min() - lo) <= 0.01 * abs(lo), (q, nc, fl.min(), lo)
            assert abs(fl.max() - hi) <= 0.01 * abs(hi), (q, nc, fl.max(), hi)
