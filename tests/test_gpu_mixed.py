"""Mixed-precision explicit paths: membrane + bending stiffness stored as FP32
values with int16 residuals (a relative 2^-39) with FP64 transverse-shear
factor rows (FS_PREC_MIXED), or the shear in the same slices
(FS_PREC_MIXED_F32); FP64 drilling, mass, state and products.  The operator
agrees with the FP64 stiffness to 1e-10, trajectories stay within 1e-6 of
the FP64 oracle, runs are bitwise reproducible, and a batch gives each model
bitwise the same trajectory as running it alone."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import _lib, batch as B, explicit, fe, femm as FM, workloads as W
from flexshell.material import MatDeforElastIso
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MIXED = _lib.FS_PREC_MIXED
MIXED32 = _lib.FS_PREC_MIXED_F32


def _homog(ctx, nen, xyz, conn, d, E=210e9, nu=0.3, rho=7850.0, t=0.01, form=FM.AVERAGE_B):
    mat = MatDeforElastIso(E, nu, rho)
    f = FM.make(conn, mat, thickness=t, ctx=ctx)
    f.transv_shear_formulation = form
    f.associategeometry(xyz)
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d).diagonal()
    return f, K, M


def _layered(ctx, nen, xyz, conn, d):
    lay = W.c5_layup(5, [0.0, 45.0, -45.0, 90.0, 0.0])
    f = FM.make(conn, lay, ctx=ctx)
    f.associategeometry(xyz)
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d).diagonal()
    return f, K, M


def _case(ctx, kind):
    if kind in ("t3", "t3k", "q4"):
        nen = 4 if kind == "q4" else 3
        xyz, conn = SC.distorted_mesh(nen, 10, 8, seed=5, warp=0.4)
        d = fe.DofField(xyz.shape[0])
        d.setebc(fe.selectnode(xyz, [-np.inf, 0.02, -np.inf, np.inf, -np.inf, np.inf], 0.05), [1, 2, 3, 4, 5, 6])
        d.numberdofs()
        f, K, M = _homog(ctx, nen, xyz, conn, d, t=0.02, form=FM.AVERAGE_K if kind == "t3k" else FM.AVERAGE_B)
    else:
        nen = 3 if kind == "t3comp" else 4
        spec = dict(W.served_c5(nreq=1, seed=9)[0], nen=nen, nel=300.0, kind="explicit")
        m = W.c5_model(spec)
        xyz, conn, d = m["xyz"], m["conn"], m["dchi"]
        f, K, M = _layered(ctx, nen, xyz, conn, d)
    return f, K, M, d


def _apply(f, M, U0, precision):
    """K U0 through one undamped step from rest: V_1 = -dt/2 invM K U_0."""
    n = U0.size
    dt = 1.0
    r = explicit.run(f._model, 0.0, U0, np.zeros(n), dt, 1, np.zeros(n), np.zeros(2), precision=precision)
    return -2.0 * r["V"] * M / dt


@pytest.mark.parametrize("kind", ["t3", "t3k", "q4", "t3comp", "q4comp"])
def test_mixed_operator_matches_fp64(ctx, kind):
    print(f"[gpu] mixed operator {kind}", flush=True)
    f, K, M, d = _case(ctx, kind)
    U0 = np.random.default_rng(2).normal(0, 1e-3, d.nfreedofs)
    ref = K @ U0
    # stored values carry a relative 2^-39 (FP32 + int16 residual)
    ku = _apply(f, M, U0, MIXED)
    assert np.linalg.norm(ku - ref) <= 1e-10 * np.linalg.norm(ref)
    ku32 = _apply(f, M, U0, MIXED32)
    assert np.linalg.norm(ku32 - ref) <= 1e-10 * np.linalg.norm(ref)
    # the FP64 path through the same entry point reproduces K U0 to rounding
    k64 = _apply(f, M, U0, _lib.FS_PREC_FP64)
    assert np.linalg.norm(k64 - ref) <= 1e-13 * np.linalg.norm(ref)


def _pulse_plate(nx=32, ny=16):
    c = W.plate_c4(nx=nx, ny=ny)
    d = c["dchi"]
    sigma, peak = 0.1, 1e4
    x0 = c["x0"]
    F = W.pressure_load_t3(c["xyz"], c["conn"], d,
                           lambda X: peak * np.exp(-np.sum((X - x0) ** 2, axis=1) / (2 * sigma ** 2)))
    return c, F


def test_mixed_trajectory_within_1e6_of_oracle(ctx):
    """Plate 2 x 1 m, t 0.004, steel, T3 32 x 16, clamped, Gaussian pressure
    pulse (sigma 0.1 m, 1e4 Pa peak, 1e-4 s): 1500 steps at 0.9 of the stable
    dt; max over probes and steps of |u_mixed - u_oracle| / max |u_oracle|."""
    print("[gpu] mixed trajectory vs oracle", flush=True)
    c, F = _pulse_plate()
    d = c["dchi"]
    f, K, M = _homog(ctx, 3, c["xyz"], c["conn"], d, t=c["t"])
    nf = d.nfreedofs
    dt = explicit.stable_dt(explicit.omega_max(f, np.random.default_rng(1).random(nf), maxit=60))
    nsteps = 1500
    g = (explicit.sample_times(nsteps, dt) <= 1e-4).astype(float)
    wdofs = d.dofnums[:, 2][d.dofnums[:, 2] < nf]
    probes = wdofs[:: max(1, len(wdofs) // 40)]
    z = np.zeros(nf)
    _, _, po = O.cd_loop(K, M, 0.0, dt, nsteps, F, g, z, z, probes)
    out = {}
    for prec in (MIXED, MIXED32):
        r = explicit.run(f._model, 0.0, z, z, dt, nsteps, F, g, probes, 1, precision=prec)
        err = np.abs(r["probes"] - po).max() / np.abs(po).max()
        out[prec] = err
        print(f"[gpu] precision {prec}: max rel trajectory error {err:.3e}", flush=True)
    assert out[MIXED] <= 1e-6
    # FP32 shear on a coarse thin mesh (element size ~16 thicknesses): the
    # least favourable case for it, still within the same bar
    assert out[MIXED32] <= 1e-6
    r1 = explicit.run(f._model, 0.0, z, z, dt, 200, F, g[:201], precision=MIXED)
    r2 = explicit.run(f._model, 0.0, z, z, dt, 200, F, g[:201], precision=MIXED)
    assert np.array_equal(r1["U"], r2["U"]) and np.array_equal(r1["V"], r2["V"])


def test_mixed_batch_equals_alone(ctx):
    print("[gpu] mixed batch vs alone", flush=True)
    models = W.batch_c3(nmodels=6, seed=22, nmin=4, nmax=10)
    nsteps = 50
    rng = np.random.default_rng(8)
    V0s = [W.gaussian_velocity(m["dchi"], 1e-3, seed=k) for k, m in enumerate(models)]
    Fm = [rng.normal(0, 1.0, m["dchi"].nfreedofs) for m in models]
    nm = len(models)
    dts, alone = [], []
    for k, m in enumerate(models):
        f, K, M = _homog(ctx, 3, m["xyz"], m["conn"], m["dchi"], E=m["material"].E, nu=m["material"].nu,
                         rho=m["material"].rho, t=m["t"])
        nf = m["dchi"].nfreedofs
        dts.append(explicit.stable_dt(explicit.omega_max(f, np.random.default_rng(k).random(nf))))
        gk = np.cos(np.arange(nsteps + 1) * 0.05 * (k + 1))
        alone.append(explicit.run(f._model, 0.0, np.zeros(nf), V0s[k], dts[k], nsteps, Fm[k], gk, precision=MIXED))
        f._model.close()
    gv = np.stack([np.cos(np.arange(nsteps + 1) * 0.05 * (k + 1)) for k in range(nm)], axis=1)
    bt = B.ShellBatch(models, ctx=ctx)
    bt.stiffness(to_host=False)
    bt.mass()
    for use_graph in (False, True):
        r = bt.explicit(dts, np.zeros(nm), np.zeros(bt.nfree), np.concatenate(V0s), np.concatenate(Fm), gv, nsteps,
                        use_graph=use_graph, precision=MIXED)
        for k in range(nm):
            sl = bt.model_slice(k)
            assert np.array_equal(r["U"][sl], alone[k]["U"]), (use_graph, k)
            assert np.array_equal(r["V"][sl], alone[k]["V"]), (use_graph, k)
    bt.close()
