"""GPU parity of the dof / load front end through the C ABI: numberdofs!
bit-exact against the oracle (with and without a node permutation, several
scan tiles), distributed surface loads against the oracle's distribloads for
every rule and both blocks, the Scordelis-Lo deflection with K and F both from
the device, and the pressure pulse of the 1024 x 512 plate at full size."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import shell_cases as SC
from flexshell import _lib, fe, femm as FM, workloads as W
from flexshell.material import MatDeforElastIso
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _femm(ctx, nen, conn, E=70e9, nu=0.3, t=0.01, stab_c=None):
    return FM.make(conn, MatDeforElastIso(E, nu, 0.0), thickness=t, stab_c=stab_c, ctx=ctx)


@pytest.mark.parametrize("nx,ny,perm", [(3, 2, False), (60, 40, False), (60, 40, True), (171, 97, True)])
def test_number_dofs_bitexact(ctx, nx, ny, perm):
    print(f"[gpu] number dofs {nx}x{ny} perm={perm}", flush=True)
    xyz, conn = SC.distorted_mesh(3, nx, ny, seed=2)
    N = xyz.shape[0]
    rng = np.random.default_rng(nx)
    d = fe.DofField(N)
    d.is_fixed[:] = rng.random((N, 6)) < 0.25
    p = rng.permutation(N) if perm else None
    f = _femm(ctx, 3, conn)
    f.associategeometry(xyz)
    f.number_dofs(xyz, d, p)
    dn, nf = O.number_dofs(d.is_fixed, p)
    assert d.nfreedofs == nf and np.array_equal(d.dofnums, dn)
    # the installed numbering drives the pattern like set_dofs would
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    assert K.shape == (nf, nf)


@pytest.mark.parametrize("nen,npts", [(3, 1), (3, 3), (3, 6), (4, 4)])
@pytest.mark.parametrize("block", [_lib.FS_BLOCK_FF, _lib.FS_BLOCK_ALL])
def test_surface_load_matches_oracle(ctx, nen, npts, block):
    print(f"[gpu] surface load nen={nen} npts={npts} block={block}", flush=True)
    xyz, conn = SC.distorted_mesh(nen, 9, 7, seed=4, warp=0.5)
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [-np.inf, 0.05, -np.inf, np.inf, -np.inf, np.inf], 0.02), [1, 2, 3])
    d.numberdofs()
    f = _femm(ctx, nen, conn)
    f.associategeometry(xyz)
    nrows = d.nfreedofs if block == _lib.FS_BLOCK_FF else d.nalldofs
    # uniform intensity
    fu = [0.3, -1.2, 4.0, 0.01, 0.0, -0.02]
    got = f.distribloads(xyz, d, fu, npts, block)
    ref = O.surface_load(nen, xyz, conn, fu, d.nalldofs, d.dofnums, npts)[:nrows]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    # position-dependent intensity (ForceIntensity function) at the rule's points
    pts = O.load_points(nen, xyz, conn, npts)

    def trac(X):
        out = np.zeros((X.shape[0], 6))
        out[:, 2] = np.exp(-((X[:, 0] - 0.4) ** 2 + (X[:, 1] - 0.3) ** 2) / 0.05)
        out[:, 0] = X[:, 1]
        return out

    got = f.distribloads(xyz, d, trac, npts, block)
    fp = trac(pts.reshape(-1, 3)).reshape(pts.shape[0], npts, 6)
    ref = O.surface_load(nen, xyz, conn, fp, d.nalldofs, d.dofnums, npts)[:nrows]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.allclose(f._model.load_points(npts), pts, rtol=0, atol=1e-15)


@pytest.mark.parametrize("nen", [3, 4])
def test_scordelis_load_and_stiffness_on_device(ctx, nen):
    """Scordelis-Lo n = 8 with the gravity load from the device distribloads
    and K from the device: the reference percentage of the analytical
    -0.3024 (test_shell_statics.jl:97-108, test_q4rs_shell_statics.jl:92-110)."""
    print(f"[gpu] scordelis device load nen={nen}", flush=True)
    c = SC.scordelis(nen, 8)
    d = c["dchi"]
    f = _femm(ctx, nen, c["conn"], E=c["E"], nu=c["nu"], t=c["t"], stab_c=c["stab_c"])
    f.associategeometry(c["xyz"])
    F = f.distribloads(c["xyz"], d, [0, 0, -90, 0, 0, 0], None, _lib.FS_BLOCK_ALL)
    assert np.abs(F - c["F"]).max() <= 1e-14 * np.abs(c["F"]).max()
    K = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), c["xyz"], None, None, d)
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), F[:nf])
    w = d.scattersysvec(U)[c["probe"], 2]
    ref = (SC.SCORDELIS_T3 if nen == 3 else SC.SCORDELIS_Q4)[8]
    assert abs(w / -0.3024 * 100 - ref) / ref < 1e-8


def test_pressure_pulse_full_size(ctx):
    """The 1024 x 512 T3 plate's Gaussian pressure (TriRule(3)) from the
    device against the vectorized host construction of the same loads."""
    print("[gpu] pressure pulse 1024x512", flush=True)
    c = W.plate_c4()
    d = c["dchi"]
    f = _femm(ctx, 3, c["conn"], t=c["t"])
    f.associategeometry(c["xyz"])
    x0 = c["x0"]

    def p_at(X):
        return 1e4 * np.exp(-np.sum((X - x0) ** 2, axis=1) / (2 * 0.1 ** 2))

    def trac(X):
        out = np.zeros((X.shape[0], 6))
        out[:, 2] = p_at(X)
        return out

    got = f.distribloads(c["xyz"], d, trac, 3)
    ref = W.pressure_load_t3(c["xyz"], c["conn"], d, p_at)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
