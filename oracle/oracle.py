"""CPU oracle for the shell hot path -- TEST INFRASTRUCTURE ONLY.

This module and oracle/shell_oracle.c restate the reference algorithms of
FinEtoolsFlexStructures.jl v3.6.4 (plus the unvendored FinEtools 8.2.5 /
FinEtoolsDeforLinear 3.0.6 primitives they call) on the CPU.  They are the
checker the HIP product is compared against.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product (finetoolsflexstructures.jl_amd/) never does.

The reference is Julia and cannot be built or run here (no julia, no depot,
FinEtools not vendored), so there is no oracle/_ref.

What pins each function (tests/test_oracle_pins.py unless noted):

  or_t3ff_elmat (T3FF)        FV12 free-plate frequencies, 8 elastic modes,
                              rtol 1e-6 (test/test_shell_dynamics.jl:113-132);
                              Scordelis-Lo T3FF percentages, rtol 1e-4
                              (test/test_shell_statics.jl:97-108); rigid-body
                              null space of a free curved element.
  or_t3ff_elmat (T3FFComp)    Barbero 3.1 max deflection, rtol 1e-8
                              (test/test_composite_shell_statics.jl:97);
                              Nayak 4.3 nondimensional frequencies, abs 1e-10
                              (test/test_composite_shell_dynamics.jl:432-443);
                              equality with T3FF for an isotropic one-ply layup.
  or_q4rs_elmat (Q4RS)        Scordelis-Lo Q4RS percentages, rtol 1e-4
                              (test/test_q4rs_shell_statics.jl:92-110);
                              rigid-body null space.
  or_q4rs_elmat (Q4RSComp)    Barbero 3.1 max deflection, rtol 1e-8
                              (test/test_composite_shell_statics_q4rs.jl:97).
  or_t3_normals/or_q4_normals flat plate -> +z, valid; folded plate crease ->
                              invalid along the fold (threshold logic).
  or_element_masses           closed form total mass rho*t*area; the FV12 and
                              Nayak frequencies above (they need M).
  or_assemble_csr             brute-force dense sum of element matrices
                              (test/test_utilities.jl:14-46 semantics).
  or_laminate / ply matrices  Barbero Ex 5.7 Dps and A, single isotropic ply A,
                              cross-ply B = 0 (test/test_composite_layup.jl:
                              321-328, 431-451, 594); H of one ply = 5/6 t G.
  or_cd_loop                  single-dof closed forms of the CD recurrence:
                              undamped free vibration from the force-only
                              A0; damped free decay (amplitude ratio and
                              phase per step); a step load switched on at a
                              given step (the g index); the invMC of a damped
                              loop whose dt was adjusted (two step sizes).
  or_pwr                      largest eigenvalue by scipy on small models.
  surface_load, number_dofs   Scordelis-Lo deflections (the load of
                              test_shell_statics.jl:84-86); exactness of each
                              rule on polynomial loads of its degree; the
                              numbering by brute force on small flag sets
                              (tests/test_frontend_pins.py).
  or_resultants_batch         membrane, bending and shear patch closed forms;
  (inspectintegpoints)        rigid motions on a tilted flat mesh; energy
                              identity against or_*_elmat (homogeneous and
                              layered); isotropic ply = homogeneous; the
                              layered Q4 double transform on a rotated grid
                              (tests/test_resultants_pins.py).
"""
import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libshell_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing at {LIB_PATH}; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.or_t3ff_elmat.argtypes = [_dp, _dp, _u8p, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                    C.c_double, _dp]
        L.or_t3ff_elmat.restype = None
        L.or_q4rs_elmat.argtypes = [_dp, _dp, _u8p, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.or_q4rs_elmat.restype = C.c_int
        L.or_stiffness_batch.argtypes = [C.c_int, C.c_int64, _dp, _i64p, _dp, _u8p, C.c_int, _dp, _dp, _dp,
                                         _i32p, C.c_double, C.c_double, C.c_int, C.c_double, _dp]
        L.or_stiffness_batch.restype = C.c_int
        for f in (L.or_t3_normals, L.or_q4_normals):
            f.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_double, _dp, _u8p]
            f.restype = None
        L.or_element_masses.argtypes = [C.c_int, C.c_int64, _dp, _i64p, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, _dp, _dp]
        L.or_element_masses.restype = None
        L.or_assemble_csr.argtypes = [C.c_int, C.c_int64, _i64p, _i64p, C.c_int64, _dp, _i64p, _i64p, _dp]
        L.or_assemble_csr.restype = C.c_int64
        L.or_assemble_count.argtypes = [C.c_int, C.c_int64, _i64p, _i64p, C.c_int64]
        L.or_assemble_count.restype = C.c_int64
        L.or_cd_loop.argtypes = [C.c_int64, _i64p, _i64p, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int64,
                                 _dp, _dp, _dp, _dp, C.c_int, _i64p, _dp]
        L.or_cd_loop.restype = None
        L.or_pwr.argtypes = [C.c_int64, _i64p, _i64p, _dp, _dp, _dp, C.c_int]
        L.or_pwr.restype = C.c_double
        L.or_resultants_batch.argtypes = [C.c_int, _dp, _i64p, _dp, _u8p, C.c_int, _dp, _dp, _dp, _i32p,
                                          C.c_double, C.c_double, _dp, C.c_int64, _i64p, C.c_int, _dp, C.c_int, _dp]
        L.or_resultants_batch.restype = C.c_int
        L.or_ply_matrices.argtypes = [_dp, _dp, _dp]
        L.or_ply_matrices.restype = None
        L.or_laminate.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                  _dp]
        L.or_laminate.restype = None
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------- materials
def iso_D6(E, nu):
    """MatDeforElastIso 3-D stiffness (FinEtoolsDeforLinear), ordering 11,22,33,12,13,23."""
    lam = E * nu / (1 + nu) / (1 - 2 * nu)
    mu = E / 2 / (1 + nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    return D


def ortho_D6(E1, E2, E3, nu12, nu13, nu23, G12, G13, G23):
    """MatDeforElastOrtho 3-D stiffness: inverse of the compliance."""
    Cm = np.array([[1 / E1, -nu12 / E1, -nu13 / E1, 0, 0, 0], [-nu12 / E1, 1 / E2, -nu23 / E2, 0, 0, 0],
                   [-nu13 / E1, -nu23 / E2, 1 / E3, 0, 0, 0], [0, 0, 0, 1 / G12, 0, 0], [0, 0, 0, 0, 1 / G13, 0],
                   [0, 0, 0, 0, 0, 1 / G23]])
    return np.linalg.inv(Cm)


def lamina_D6(E1, E2, nu12, G12, G13, G23):
    """lamina_material (CompositeLayupModule.jl:105-123): E3 = E2, nu13 = nu12, nu23 = 0."""
    return ortho_D6(E1, E2, E2, nu12, nu12, 0.0, G12, G13, G23)


def ply_matrices(D6):
    """(Dps, Dts) of Ply (CompositeLayupModule.jl:71-97), via the C restatement."""
    D6 = _c(D6, np.float64)
    Dps = np.zeros(9)
    Dts = np.zeros(4)
    lib().or_ply_matrices(_p(D6), _p(Dps), _p(Dts))
    return Dps.reshape(3, 3), Dts.reshape(2, 2)


def laminate(plies, offset=0.0, tsc=0.0):
    """plies: list of (D6, thickness, angle_deg, rho). Returns A, B, D, H, (md, mi)."""
    n = len(plies)
    Dps = np.zeros((n, 9))
    Dts = np.zeros((n, 4))
    th = np.zeros(n)
    an = np.zeros(n)
    rho = np.zeros(n)
    for i, (D6, t, a, r) in enumerate(plies):
        P, T = ply_matrices(D6)
        Dps[i] = P.ravel()
        Dts[i] = T.ravel()
        th[i], an[i], rho[i] = t, a, r
    A = np.zeros(9)
    B = np.zeros(9)
    D = np.zeros(9)
    H = np.zeros(4)
    inert = np.zeros(2)
    lib().or_laminate(n, _p(Dps), _p(Dts), _p(th), _p(an), _p(rho), float(offset), float(tsc), _p(A), _p(B),
                      _p(D), _p(H), _p(inert))
    return A.reshape(3, 3), B.reshape(3, 3), D.reshape(3, 3), H.reshape(2, 2), tuple(inert)


def shell_D(D6):
    """(Dps, Dt) of _shell_material_stiffness (same condensation as Ply)."""
    return ply_matrices(D6)


# ---------------------------------------------------------------- geometry
def normals(nen, xyz, conn, threshold=30.0):
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int64)
    N = xyz.shape[0]
    nrm = np.zeros((N, 3))
    valid = np.zeros(N, dtype=np.uint8)
    f = lib().or_t3_normals if nen == 3 else lib().or_q4_normals
    f(N, _p(xyz), conn.shape[0], _p(conn, C.c_int64), float(threshold), _p(nrm), _p(valid, C.c_uint8))
    return nrm, valid.astype(bool)


def csys_normals(nen, xyz, conn, csys, threshold=30.0):
    """Normals of the layered FEMMs from a constant layup csys (associategeometry!,
    FEMMShellT3FFCompModule.jl:489-542; FEMMShellQ4RSCompModule.jl:447-509)."""
    xyz = np.asarray(xyz, dtype=float)
    conn = np.asarray(conn)
    N = xyz.shape[0]
    n3 = np.asarray(csys, dtype=float)[:, 2]
    nrm = np.zeros((N, 3))
    used = np.zeros(N, dtype=bool)
    used[conn.ravel()] = True
    nrm[used] = n3 / np.linalg.norm(n3)
    valid = np.ones(N, dtype=bool)
    if nen == 3:
        s = np.sin(threshold / 180 * np.pi)
        ntol = 1 - np.sqrt(1 - s * s)
        for (i, j, k) in conn:
            e1 = (xyz[j] - xyz[i]) / np.linalg.norm(xyz[j] - xyz[i])
            e3 = np.cross(e1, xyz[k] - xyz[i])
            e3 /= np.linalg.norm(e3)
            for n in (i, j, k):
                if nrm[n] @ e3 < 1 - ntol:
                    valid[n] = False
    return nrm, valid


# ---------------------------------------------------------------- element matrices
def element_matrices(nen, xyz, conn, normals_, valid, *, Dps=None, Dt=None, t=None, groups=None, elem_group=None,
                     stab_c=None, formulation=0, drill=1.0):
    """All element stiffness matrices (E x k x k).  Homogeneous: Dps, Dt
    (without 5/6), t.  Layered: groups = [(A, B, D, H, csys, thickness), ...]."""
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int64)
    nrm = _c(normals_, np.float64)
    val = _c(valid, np.uint8)
    E = conn.shape[0]
    k = 6 * nen
    Ke = np.zeros((E, k, k))
    if stab_c is None:
        stab_c = 5 / 12 / 1.5 if nen == 3 else 0.1
    if groups is None:
        mat = _c(np.concatenate([np.asarray(Dps, float).ravel(), np.asarray(Dt, float).ravel()]), np.float64)
        cs = np.eye(3).ravel().copy()
        st = lib().or_stiffness_batch(nen, E, _p(xyz), _p(conn, C.c_int64), _p(nrm), _p(val, C.c_uint8), 0, _p(mat),
                                      _p(cs), None, None, float(t), float(stab_c), int(formulation), float(drill),
                                      _p(Ke))
    else:
        mat = _c(np.concatenate([np.concatenate([np.asarray(A).ravel(), np.asarray(B).ravel(), np.asarray(D).ravel(),
                                                 np.asarray(H).ravel()]) for (A, B, D, H, _, _) in groups]),
                 np.float64)
        cs = _c(np.concatenate([np.asarray(c_, float).ravel() for (_, _, _, _, c_, _) in groups]), np.float64)
        tg = _c([g[5] for g in groups], np.float64)
        eg = None if elem_group is None else _c(elem_group, np.int32)
        st = lib().or_stiffness_batch(nen, E, _p(xyz), _p(conn, C.c_int64), _p(nrm), _p(val, C.c_uint8), 1, _p(mat),
                                      _p(cs), _p(tg), _p(eg, C.c_int32), 0.0, float(stab_c), int(formulation),
                                      float(drill), _p(Ke))
    if st != 0:
        raise ArithmeticError("Singular metric matrix in _gradN_e!")
    return Ke


def resultants(nen, xyz, conn, normals_, valid, u, *, Dps=None, Dt=None, t=None, groups=None, elem_group=None,
               stab_c=None, felist=None, ocs=None, twice=False):
    """Stress resultants N11 N22 N12 M11 M22 M12 Q1 Q2 at the integration
    points of the listed elements (nlist x npts x 8; npts 1 for T3, 4 for Q4):
    inspectintegpoints for all three quantities.  u: N x 6 nodal values.
    ocs: None (isoparametric csys), one 3 x 3 matrix, or nlist x npts x 3 x 3.
    twice: the layered Q4 form that applies T to the displacements twice."""
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int64)
    nrm = _c(normals_, np.float64)
    val = _c(valid, np.uint8)
    u = _c(u, np.float64)
    E = conn.shape[0]
    fl = _c(np.arange(E) if felist is None else felist, np.int64)
    npts = 1 if nen == 3 else 4
    out = np.zeros((fl.shape[0], npts, 8))
    if stab_c is None:
        stab_c = 5 / 12 / 1.5 if nen == 3 else 0.1
    if groups is None:
        mat = _c(np.concatenate([np.asarray(Dps, float).ravel(), np.asarray(Dt, float).ravel()]), np.float64)
        cs, tg, eg, comp = None, None, None, 0
    else:
        mat = _c(np.concatenate([np.concatenate([np.asarray(A).ravel(), np.asarray(B).ravel(), np.asarray(D).ravel(),
                                                 np.asarray(H).ravel()]) for (A, B, D, H, _, _) in groups]),
                 np.float64)
        cs = _c(np.concatenate([np.asarray(c_, float).ravel() for (_, _, _, _, c_, _) in groups]), np.float64)
        tg = _c([g[5] for g in groups], np.float64)
        eg = None if elem_group is None else _c(elem_group, np.int32)
        comp = 1
    if ocs is None:
        kind, oc = 0, None
    else:
        oc = _c(ocs, np.float64)
        kind = 1 if oc.size == 9 else 2
        if kind == 2 and oc.size != fl.shape[0] * npts * 9:
            raise ValueError("ocs must be one 3 x 3 matrix or nlist x npts x 3 x 3")
    st = lib().or_resultants_batch(nen, _p(xyz), _p(conn, C.c_int64), _p(nrm), _p(val, C.c_uint8), comp, _p(mat),
                                   _p(cs), _p(tg), _p(eg, C.c_int32), float(t or 0.0), float(stab_c), _p(u),
                                   fl.shape[0], _p(fl, C.c_int64), kind, _p(oc), int(bool(twice)), _p(out))
    if st != 0:
        raise ArithmeticError("Singular metric matrix in _gradN_e!")
    return out


# ---------------------------------------------------------------- loads and numbering
# FinEtools integration rules (IntegRuleModule): TriRule(1), TriRule(3),
# TriRule(6) and GaussRule(2, 2) as (points, weights).
G2 = 0.577350269189626
RULES = {
    (3, 1): (np.array([[1 / 3, 1 / 3]]), np.array([1 / 2])),
    (3, 3): (np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.array([1 / 6, 1 / 6, 1 / 6])),
    (3, 6): (np.array([[0.816847572980459, 0.091576213509771], [0.091576213509771, 0.816847572980459],
                       [0.091576213509771, 0.091576213509771], [0.108103018168070, 0.445948490915965],
                       [0.445948490915965, 0.108103018168070], [0.445948490915965, 0.445948490915965]]),
             np.array([0.109951743655322] * 3 + [0.223381589678011] * 3) / 2),
    (4, 4): (np.array([[-G2, -G2], [-G2, G2], [G2, -G2], [G2, G2]]), np.ones(4)),
}


def _tri_bfun(xi, eta):
    return np.array([1.0 - xi - eta, xi, eta]), np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])


def _quad_bfun(xi, eta):
    N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
    gN = 0.25 * np.array([[-(1 - eta), -(1 - xi)], [(1 - eta), -(1 + xi)], [(1 + eta), (1 + xi)],
                          [-(1 + eta), (1 - xi)]])
    return N, gN


def load_points(nen, xyz, conn, npts):
    """Locations of the integration points of the rule (E x npts x 3)."""
    pts, _ = RULES[(nen, npts)]
    bf = _tri_bfun if nen == 3 else _quad_bfun
    X = np.asarray(xyz)[np.asarray(conn)]
    return np.stack([np.einsum("a,eak->ek", bf(*p)[0], X) for p in pts], axis=1)


def surface_load(nen, xyz, conn, force, nalldofs, dofnums, npts=None):
    """distribloads(FEMMBase(IntegDomain(fes, rule)), geom, dchi,
    ForceIntensity(f), 2) (FinEtools; test_shell_statics.jl:84-86) for T3/Q4
    elements: per element Fe = sum_q outer(N, f) Jac w, assembled in element
    order.  force: a 6-vector, or E x npts x 6 values at the rule's points.
    Returns the vector over all dofs (length nalldofs)."""
    xyz = np.asarray(xyz)
    conn = np.asarray(conn)
    npts = (3 if nen == 3 else 4) if npts is None else npts
    pts, wts = RULES[(nen, npts)]
    bf = _tri_bfun if nen == 3 else _quad_bfun
    F = np.zeros(nalldofs)
    f_all = np.asarray(force, dtype=float)
    uniform = f_all.size == 6
    for e in range(conn.shape[0]):
        X = xyz[conn[e]]
        fe = np.zeros((nen, 6))
        for q, ((xi, eta), w) in enumerate(zip(pts, wts)):
            N, gN = bf(xi, eta)
            J = X.T @ gN
            Jac = np.linalg.norm(np.cross(J[:, 0], J[:, 1]))
            f = f_all.reshape(6) if uniform else f_all[e, q]
            fe += np.outer(N, f) * (Jac * w)
        for a in range(nen):
            F[dofnums[conn[e, a]]] += fe[a]
    return F


def number_dofs(is_fixed, perm=None):
    """numberdofs!(dchi[, perm]) (FinEtools NodalField): free dofs first,
    node by node in node order (or in the order of perm), then the fixed ones
    in the same order.  Returns (dofnums N x 6, nfree)."""
    fx = np.asarray(is_fixed, dtype=bool)
    N = fx.shape[0]
    order = np.arange(N) if perm is None else np.asarray(perm)
    dn = np.zeros((N, 6), dtype=np.int64)
    nfree = int((~fx).sum())
    nf = nx = 0
    for I in order:
        for c in range(6):
            if fx[I, c]:
                dn[I, c] = nfree + nx
                nx += 1
            else:
                dn[I, c] = nf
                nf += 1
    return dn, nfree


def element_masses(nen, xyz, conn, *, rho=None, t=None, m0=None, m1=None):
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int64)
    E = conn.shape[0]
    tm = np.zeros(E)
    rm = np.zeros(E)
    hom = 1 if rho is not None else 0
    lib().or_element_masses(nen, E, _p(xyz), _p(conn, C.c_int64), float(m0 or 0.0), float(m1 or 0.0), hom,
                            float(rho or 0.0), float(t or 0.0), _p(tm), _p(rm))
    return tm, rm


# ---------------------------------------------------------------- assembly
def assemble_csr(nen, conn, dofnums, nrows, Ke):
    """COO append in element order + duplicate summation in input order."""
    conn = _c(conn, np.int64)
    dn = _c(dofnums, np.int64)
    Ke = _c(Ke, np.float64)
    E = conn.shape[0]
    m = lib().or_assemble_count(nen, E, _p(conn, C.c_int64), _p(dn, C.c_int64), nrows)
    if m < 0:
        raise MemoryError("oracle assembly")
    rp = np.zeros(nrows + 1, dtype=np.int64)
    cv = np.zeros(max(m, 1), dtype=np.int64)
    vals = np.zeros(max(m, 1))
    nnz = lib().or_assemble_csr(nen, E, _p(conn, C.c_int64), _p(dn, C.c_int64), nrows, _p(Ke), _p(rp, C.c_int64),
                                _p(cv, C.c_int64), _p(vals))
    if nnz < 0:
        raise MemoryError("oracle assembly")
    return sp.csr_matrix((vals[:nnz].copy(), cv[:nnz].copy(), rp), shape=(nrows, nrows))


def lumped_diag(nen, conn, dofnums, nrows, tm, rm):
    """Diagonal of the lumped mass, accumulated in element order."""
    d = np.zeros(nrows)
    conn = np.asarray(conn)
    dn = np.asarray(dofnums)
    for e in range(conn.shape[0]):
        for a in range(nen):
            for c in range(6):
                dof = dn[conn[e, a], c]
                if dof < nrows:
                    d[dof] += tm[e] if c < 3 else rm[e]
    return d


# ---------------------------------------------------------------- dynamics
def cd_loop(K, Mdiag, cdamp, dt, nsteps, Fmag, g, U0, V0, probes=(), dt_mass=None):
    """nsteps of _cd_loop! at the (already adjusted) dt; invMC formed with
    dt_mass (the dt before the adjustment; None = dt)."""
    K = sp.csr_matrix(K)
    n = K.shape[0]
    rp = _c(K.indptr, np.int64)
    ci = _c(K.indices, np.int64)
    kv = _c(K.data, np.float64)
    U = _c(U0, np.float64).copy()
    V = _c(V0, np.float64).copy()
    pr = _c(probes, np.int64)
    pout = np.zeros((nsteps + 1, len(pr)))
    lib().or_cd_loop(n, _p(rp, C.c_int64), _p(ci, C.c_int64), _p(kv), _p(_c(Mdiag, np.float64)), float(cdamp),
                     float(dt), float(dt if dt_mass is None else dt_mass), int(nsteps), _p(_c(Fmag, np.float64)),
                     _p(_c(g, np.float64)), _p(U), _p(V), len(pr), _p(pr, C.c_int64), _p(pout))
    return U, V, pout


def cd_loop_tend(K, Mdiag, ksi, omegad, U0, V0, tend, dt, Fmag, gfun, probes=()):
    """_cd_loop!(M, K, ksi, U0, V0, tend, dt, force!, peek) as the reference
    calls it (plate_expl_examples.jl:61-94): C = ksi 2 omegad diag(M) and
    invMC from the given dt (:68-71), then nsteps = round(tend/dt) and dt
    re-derived when nsteps dt < tend (:72-75), force!(F, t) = gfun(t) Fmag
    at t accumulated as t = t + dt (:79-84).  Returns (U, V, probes, dt, nsteps)."""
    nsteps = int(round(tend / dt))
    dt_in = dt
    if nsteps * dt < tend:
        dt = tend / (nsteps + 1)
    t = 0.0
    g = np.zeros(nsteps + 1)
    g[0] = gfun(t)
    for n in range(1, nsteps + 1):
        t = t + dt
        g[n] = gfun(t)
    U, V, pout = cd_loop(K, Mdiag, ksi * 2 * omegad, dt, nsteps, Fmag, g, U0, V0, probes, dt_mass=dt_in)
    return U, V, pout, dt, nsteps


def pwr(K, Mdiag, v0, maxit=30):
    K = sp.csr_matrix(K)
    v = _c(v0, np.float64).copy()
    return lib().or_pwr(K.shape[0], _p(_c(K.indptr, np.int64), C.c_int64), _p(_c(K.indices, np.int64), C.c_int64),
                        _p(_c(K.data, np.float64)), _p(_c(Mdiag, np.float64)), _p(v), int(maxit))
