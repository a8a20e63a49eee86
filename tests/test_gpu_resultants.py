"""GPU parity of the stress-resultant recovery (inspectintegpoints) through
the C ABI: the kernel against the oracle on warped, jittered meshes for the
four FEMMs and every output-csys kind; the closed-form membrane patch at full
size; the Python inspectintegpoints mirror; argument errors.

Tolerances: 1e-11 of the largest magnitude of each quantity for an output
csys at a generic angle to the element frame.  With the isoparametric csys the
reference's own angle formula n = sign * sqrt(1 - m^2)
(TransformerModule.jl:100-103) turns a last-bit difference of m = 1 into an
angle of about 1.5e-8, so that case is held to 1e-7."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import _lib, fe, femm as FM
from flexshell.material import CompositeLayup, MatDeforElastIso, Ply, lamina_material
from oracle import oracle as O

pytestmark = pytest.mark.gpu

E_, NU, T_ = 70e9, 0.3, 0.01


def _layup(csys):
    mat = lamina_material(0.0, 133.86e9, 7.706e9, 0.301, 4.306e9, 4.306e9, 2.76e9)
    plies = [Ply("a", mat, 0.002, 0.0), Ply("b", mat, 0.003, 35.0), Ply("c", mat, 0.002, -60.0)]
    return CompositeLayup("lay", plies, csys)


def _rotz(a):
    return np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])


def _case(ctx, kind, nx=9, ny=7, warp=0.5, seed=21):
    nen = 3 if kind in ("t3", "t3comp") else 4
    xyz, conn = SC.distorted_mesh(nen, nx, ny, seed=seed, warp=warp)
    if kind in ("t3", "q4"):
        f = FM.make(conn, MatDeforElastIso(E_, NU, 0.0), thickness=T_, ctx=ctx)
    else:
        f = FM.make(conn, _layup(_rotz(0.35)), ctx=ctx)
    f.associategeometry(xyz)
    return nen, xyz, conn, f


def _oracle(f, nen, xyz, conn, u, felist=None, ocs=None, twice=False):
    if f.composite:
        lay = f.layup_groups[0][0]
        A, B, D = lay.laminate_stiffnesses()
        H = lay.laminate_transverse_stiffness()
        groups = [(A, B, D, H, lay.csys, lay.thickness())]
        return O.resultants(nen, xyz, conn, f._normals, f._normal_valid, u, groups=groups, felist=felist, ocs=ocs,
                            twice=twice, stab_c=f.stab_c)
    Dps, Dt = O.shell_D(O.iso_D6(E_, NU))
    return O.resultants(nen, xyz, conn, f._normals, f._normal_valid, u, Dps=Dps, Dt=Dt, t=T_, felist=felist,
                        ocs=ocs, stab_c=f.stab_c)


def _check(got, ref, tol):
    for s in (slice(0, 3), slice(3, 6), slice(6, 8)):
        scale = np.abs(ref[..., s]).max()
        err = np.abs(got[..., s] - ref[..., s]).max()
        assert err <= tol * scale, (s, err / scale)


@pytest.mark.parametrize("kind", ["t3", "q4", "t3comp", "q4comp"])
@pytest.mark.parametrize("csys", ["default", "matrix", "per_point"])
def test_resultants_match_oracle(ctx, kind, csys):
    print(f"[gpu] resultants {kind} csys={csys}", flush=True)
    nen, xyz, conn, f = _case(ctx, kind)
    rng = np.random.default_rng(4)
    u = rng.normal(0, 1e-3, (xyz.shape[0], 6))
    felist = rng.permutation(conn.shape[0])[: conn.shape[0] // 2]
    npts = 1 if nen == 3 else 4
    if csys == "default":
        ocs, tol = None, 1e-7
        if f.composite:
            ocs = f.layup_groups[0][0].csys
            tol = 1e-11
        got, _ = f.resultants(xyz, u, felist)
    elif csys == "matrix":
        ocs, tol = _rotz(-0.6), 1e-11
        got, _ = f.resultants(xyz, u, felist, outputcsys=ocs)
    else:
        ocs = np.array([_rotz(0.1 * (l % 7) - 0.3) for l in range(felist.shape[0] * npts)])
        tol = 1e-11
        got, _ = f.resultants(xyz, u, felist, outputcsys=lambda X, J, e, q, _o=ocs, _fl=list(felist):
                              _o[_fl.index(e) * npts + q])
        ocs = ocs.reshape(felist.shape[0], npts, 3, 3)
    ref = _oracle(f, nen, xyz, conn, u, felist=felist, ocs=ocs, twice=(kind == "q4comp"))
    _check(got, ref, tol)


def test_q4_layered_transform_once(ctx):
    """The FS_RES_LAYERED_Q4_SINGLE_T flag matches the oracle's single
    transformation, and differs from the reference form on a warped mesh."""
    print("[gpu] resultants q4comp transform once", flush=True)
    nen, xyz, conn, f = _case(ctx, "q4comp")
    u = np.random.default_rng(5).normal(0, 1e-3, (xyz.shape[0], 6))
    cs = f.layup_groups[0][0].csys
    got, _ = f.resultants(xyz, u, transform_once=True)
    ref = _oracle(f, nen, xyz, conn, u, ocs=cs, twice=False)
    _check(got, ref, 1e-11)
    twice, _ = f.resultants(xyz, u)
    assert np.abs(twice - got).max() > 1e-3 * np.abs(got).max()


@pytest.mark.parametrize("nen", [3, 4])
def test_membrane_patch_full_size(ctx, nen):
    """Uniform membrane strain on the full-size meshes (T3 plate 1024 x 512,
    Q4 512 x 512 flat grid): N = t Dps eps at every point, M = Q = 0."""
    print(f"[gpu] membrane patch full size nen={nen}", flush=True)
    if nen == 3:
        xy, conn = fe.T3block(2.0, 1.0, 1024, 512)
    else:
        xy, conn = fe.Q4block(2.0, 1.0, 512, 512)
    xyz = fe.xyz3(xy)
    f = FM.make(conn, MatDeforElastIso(E_, NU, 0.0), thickness=T_, ctx=ctx)
    f.associategeometry(xyz)
    G = np.array([[1.3e-4, -0.4e-4], [0.9e-4, -0.6e-4]])
    u = np.zeros((xyz.shape[0], 6))
    u[:, :2] = xyz[:, :2] @ G.T
    out, loc = f.resultants(xyz, u, outputcsys=np.eye(3))
    Dps, _ = O.shell_D(O.iso_D6(E_, NU))
    Nref = T_ * Dps @ np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
    scale = np.abs(Nref).max()
    assert out.shape == (conn.shape[0], 1 if nen == 3 else 4, 8)
    assert np.abs(out[..., :3] - Nref).max() < 1e-9 * scale
    assert np.abs(out[..., 3:]).max() < 1e-9 * scale * T_
    print(f"[gpu] resultants kernel {f._model.timing(5):.3f} ms for {conn.shape[0]} elements", flush=True)


def test_inspectintegpoints_mirror(ctx):
    """The Python inspectintegpoints calls the inspector per element and
    point with the selected quantity, in both call forms (with and without
    dT), and the default quantity is the bending moment."""
    print("[gpu] inspectintegpoints mirror", flush=True)
    nen, xyz, conn, f = _case(ctx, "q4", nx=4, ny=3)
    u = np.random.default_rng(6).normal(0, 1e-3, (xyz.shape[0], 6))
    res, loc = f.resultants(xyz, u)
    felist = [3, 0, 7]

    def insp(idat, e, c, xe, out, lc):
        idat.append((e, out.copy(), lc.copy()))
        return idat

    got = f.inspectintegpoints(xyz, u, felist, insp, [], "membrane")
    assert [g[0] for g in got] == [3] * 4 + [0] * 4 + [7] * 4
    for k, (e, out, lc) in enumerate(got):
        j = k % 4
        assert np.array_equal(out, res[e, j, 0:3])
        assert np.allclose(lc, loc[e, j])
    got = f.inspectintegpoints(xyz, u, None, felist, insp, [], ":shear")
    for k, (e, out, _) in enumerate(got):
        assert np.array_equal(out[:2], res[e, k % 4, 6:8]) and out[2] == 0.0
    got = f.inspectintegpoints(xyz, u, felist, insp, [])
    for k, (e, out, _) in enumerate(got):
        assert np.array_equal(out, res[e, k % 4, 3:6])


def test_resultants_errors(ctx):
    """Out-of-range element numbers and bad csys kinds are rejected on the
    host before any launch; an empty list returns nothing."""
    print("[gpu] resultants errors", flush=True)
    nen, xyz, conn, f = _case(ctx, "t3", nx=3, ny=2)
    u = np.zeros((xyz.shape[0], 6))
    out, _ = f.resultants(xyz, u, np.zeros(0, dtype=np.int64))
    assert out.shape == (0, 1, 8)
    with pytest.raises(IndexError):
        f.resultants(xyz, u, [0, conn.shape[0]])
    p = f._shell_params()
    o = np.zeros(16)
    # through the C ABI: element out of range, unknown csys kind, unknown flag
    for bad in ([conn.shape[0]], [-1]):
        fl = np.array(bad, dtype=np.int32)
        st = ctx.lib.fs_model_resultants(f._model.h, _lib.C.byref(p), _lib.ptr(u), 1, _lib.ptr(fl, _lib.C.c_int32),
                                         0, None, 0, _lib.ptr(o), None)
        assert st == -1
    st = ctx.lib.fs_model_resultants(f._model.h, _lib.C.byref(p), _lib.ptr(u), 1, None, 7, None, 0, _lib.ptr(o),
                                     None)
    assert st == -1
    st = ctx.lib.fs_model_resultants(f._model.h, _lib.C.byref(p), _lib.ptr(u), 1, None, 0, None, 4, _lib.ptr(o),
                                     None)
    assert st == -1
    st = ctx.lib.fs_model_resultants(f._model.h, _lib.C.byref(p), _lib.ptr(u), 2, None, 0, None, 0, _lib.ptr(o),
                                     None)
    assert st == 0


def test_barrel_vault_fields_gpu(ctx):
    """The barrel vault of test_shell_statics.jl:578-728 through the device
    (stiffness, TriRule(3) load, resultants) and the mirror's
    fieldfromintegpoints: the reference's minimum and maximum of every nodal
    resultant field within its rtol 0.01, and the same fields as the
    test-side interpolation of the oracle's resultants to 1e-9."""
    import scipy.sparse.linalg as sla
    print("[gpu] barrel vault resultant fields", flush=True)
    c = SC.barrel_vault()
    xyz, conn, d = c["xyz"], c["conn"], c["dchi"]
    f = FM.make(conn, MatDeforElastIso(c["E"], c["nu"], 0.0), thickness=c["t"], stab_c=c["stab_c"], ctx=ctx)
    f.drilling_stiffness_scale = c["drill"]
    f.associategeometry(xyz)
    K = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), xyz, None, None, d)
    F = f.distribloads(xyz, d, c["force"], 3, _lib.FS_BLOCK_ALL)
    nf = d.nfreedofs
    u = d.scattersysvec(sla.spsolve(K[:nf, :nf].tocsc(), F[:nf]))
    for q, refs in SC.BARREL_RESULTANTS.items():
        for nc, (lo, hi) in enumerate(refs):
            fl = f.fieldfromintegpoints(xyz, u, q, nc, outputcsys=c["csys"])
            assert abs(fl.min() - lo) <= 0.01 * abs(lo), (q, nc, fl.min(), lo)
            assert abs(fl.max() - hi) <= 0.01 * abs(hi), (q, nc, fl.max(), hi)
    res, loc = f.resultants(xyz, u, outputcsys=c["csys"])
    fl = f.fieldfromintegpoints(xyz, u, "moment", 0, outputcsys=c["csys"])
    ref = SC.inverse_distance_field(xyz, conn, loc, res[:, :, 3])
    assert np.abs(fl - ref).max() <= 1e-9 * np.abs(ref).max()
    ef = f.elemfieldfromintegpoints(xyz, u, "membrane", 1, outputcsys=c["csys"])
    assert np.array_equal(ef, res[:, 0, 1])
