"""Batched models: every model of a batch gives bitwise the same stiffness,
mass and trajectory as when it runs alone (and agrees with the oracle)."""
import numpy as np
import pytest

from flexshell import batch as B, explicit, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(ctx, nmodels=10, seed=21):
    models = W.batch_c3(nmodels=nmodels, seed=seed, nmin=4, nmax=12)
    rng = np.random.default_rng(5)
    V0s = [W.gaussian_velocity(m["dchi"], 1e-3, seed=k) for k, m in enumerate(models)]
    Fm = [rng.normal(0, 1.0, m["dchi"].nfreedofs) for m in models]
    return models, V0s, Fm


def test_batch_equals_models_alone(ctx):
    print("[gpu] batch vs alone", flush=True)
    models, V0s, Fm = _setup(ctx)
    nsteps = 60
    # per-model dt from the plain path's power iteration
    from flexshell.femm import FEMMShellT3FF, SysmatAssemblerFFBlock, SysmatAssemblerSparseDiag
    dts, Ks, Ms = [], [], []
    for k, m in enumerate(models):
        f = FEMMShellT3FF(m["conn"], m["material"], thickness=m["t"], ctx=ctx)
        f.associategeometry(m["xyz"])
        nf = m["dchi"].nfreedofs
        Ks.append(f.stiffness(SysmatAssemblerFFBlock(nf), m["xyz"], None, None, m["dchi"]))
        Ms.append(f.mass(SysmatAssemblerFFBlock(SysmatAssemblerSparseDiag(), nf, nf), m["xyz"], m["dchi"]).diagonal())
        dts.append(explicit.stable_dt(explicit.omega_max(f, np.random.default_rng(k).random(nf))))
    nm = len(models)
    cd = np.array([0.0 if k % 2 else 50.0 * k for k in range(nm)])
    gv = np.stack([np.cos(np.arange(nsteps + 1) * 0.05 * (k + 1)) for k in range(nm)], axis=1)
    plain, _ = B.run_plain(models, dts, cd, [np.zeros_like(v) for v in V0s], V0s, Fm, gv, nsteps, ctx=ctx)
    bt = B.ShellBatch(models, ctx=ctx)
    Kb = bt.stiffness()
    Mb = bt.mass()
    for k in range(nm):
        sl = bt.model_slice(k)
        Kk = Kb[sl, sl]
        assert np.array_equal(Kk.indices, Ks[k].indices) and np.array_equal(Kk.data, Ks[k].data)
        assert np.array_equal(Mb[sl], Ms[k])
        assert Kb[sl, :].nnz == Kk.nnz  # block diagonal: nothing couples models
    for use_graph in (False, True):
        r = bt.explicit(dts, cd, np.zeros(bt.nfree), np.concatenate(V0s), np.concatenate(Fm), gv, nsteps,
                        use_graph=use_graph)
        for k in range(nm):
            sl = bt.model_slice(k)
            assert np.array_equal(r["U"][sl], plain[k]["U"]), (use_graph, k)
            assert np.array_equal(r["V"][sl], plain[k]["V"]), (use_graph, k)
    # the same batch in Infinity-Cache groups (small groups force several),
    # probes across group boundaries: bitwise the single batch
    bg = B.ShellBatchGroups(models, ctx=ctx, group_bytes=int(1.2 * max(B.step_bytes(m) for m in models)))
    assert len(bg.groups) >= 3
    bg.assemble_device()
    assert np.array_equal(bg.mass(), Mb)
    probes = np.array([0, bt.free_off[4] - 1, bt.free_off[4], bt.nfree - 1])
    ra0 = None
    for prec in (0, 2):
        ra = bt.explicit(dts, cd, np.zeros(bt.nfree), np.concatenate(V0s), np.concatenate(Fm), gv, nsteps,
                         probe_dofs=probes, probe_stride=3, precision=prec)
        rb = bg.explicit(dts, cd, np.zeros(bt.nfree), np.concatenate(V0s), np.concatenate(Fm), gv, nsteps,
                         probe_dofs=probes, probe_stride=3, precision=prec)
        assert np.array_equal(ra["U"], rb["U"]) and np.array_equal(ra["V"], rb["V"]), prec
        ra0 = ra if ra0 is None else ra0
        assert np.array_equal(ra["probes"], rb["probes"]), prec
    bg.close()
    # the groups dealt to two contexts run concurrently: the same results
    bg2 = B.ShellBatchGroups(models, ctx=ctx, group_bytes=int(1.2 * max(B.step_bytes(m) for m in models)), streams=2)
    bg2.assemble_device()
    bg2.mass()
    rc = bg2.explicit(dts, cd, np.zeros(bt.nfree), np.concatenate(V0s), np.concatenate(Fm), gv, nsteps,
                      probe_dofs=probes, probe_stride=3)
    assert np.array_equal(rc["U"], ra0["U"]) and np.array_equal(rc["probes"], ra0["probes"])
    bg2.close()
    # one model against the oracle
    k = 3
    Uo, Vo, _ = O.cd_loop(Ks[k], Ms[k], cd[k], dts[k], nsteps, Fm[k], gv[:, k], np.zeros_like(V0s[k]), V0s[k])
    assert np.linalg.norm(plain[k]["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo)
    bt.close()


@pytest.mark.timeout(900)
def test_batch_c3_full_size(ctx):
    """The 256-model T3 batch at its configured sizes (n per side 8..128,
    2.99 M elements, the seed and initial velocities of the benchmark): a
    seeded sample of models, including the largest and the smallest, gets
    bitwise the same K, M and 500-step trajectory batched as alone; the
    largest model's K is within 1e-12 of the oracle's assembly and its
    trajectory within 1e-10 of the oracle's loop.  Both mixed precisions on
    the same inputs stay within 1e-6 of the FP64 final state of every model
    (and of the oracle's for the largest)."""
    from flexshell import _lib, femm as FM
    print("[gpu] C3 full batch: 256 T3 models, n 8..128, 500 steps", flush=True)
    models = W.batch_c3()
    nm, nsteps = len(models), 500
    E = np.array([m["conn"].shape[0] for m in models])
    big = int(np.argmax(E))
    sample = sorted({big, int(np.argmin(E))} | set(np.random.default_rng(11).choice(nm, 6, replace=False).tolist()))
    V0s = [W.gaussian_velocity(m["dchi"], 1e-3, seed=100 + k) for k, m in enumerate(models)]
    g = np.ones((nsteps + 1, nm))
    dts, Ks, Ms, runs = [], {}, {}, {}
    for k, m in enumerate(models):
        f = FM.FEMMShellT3FF(m["conn"], m["material"], thickness=m["t"], ctx=ctx)
        f.associategeometry(m["xyz"])
        d = m["dchi"]
        nf = d.nfreedofs
        if k in sample:
            Ks[k] = f.stiffness(FM.SysmatAssemblerFFBlock(nf), m["xyz"], None, None, d)
            Ms[k] = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), m["xyz"], d).diagonal()
        else:
            f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), m["xyz"], None, None, d)
            f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), m["xyz"], d)
        dts.append(W.converged_dt(f, nf, seed=k))
        if k in sample:
            z = np.zeros(nf)
            runs[k] = explicit.run(f._model, 0.0, z, V0s[k], dts[k], nsteps, z, g[:, k])
        f._model.close()
    print(f"[gpu] C3 full batch: plain path done ({int(E.sum())} elements), sample {sample}", flush=True)
    bt = B.ShellBatch(models, ctx=ctx)
    Kb = bt.stiffness()
    Mb = bt.mass()
    for k in sample:
        a, b = int(bt.free_off[k]), int(bt.free_off[k + 1])
        lo, hi = int(Kb.indptr[a]), int(Kb.indptr[b])
        assert np.array_equal(Kb.indptr[a:b + 1] - lo, Ks[k].indptr)
        assert np.array_equal(Kb.indices[lo:hi] - a, Ks[k].indices)
        assert np.array_equal(Kb.data[lo:hi], Ks[k].data), k
        assert np.array_equal(Mb[a:b], Ms[k]), k
    del Kb
    nfree = bt.nfree
    r = bt.explicit(dts, np.zeros(nm), np.zeros(nfree), np.concatenate(V0s), np.zeros(nfree), g, nsteps)
    for k in sample:
        sl = bt.model_slice(k)
        assert np.array_equal(r["U"][sl], runs[k]["U"]), k
        assert np.array_equal(r["V"][sl], runs[k]["V"]), k
    # the dt of every model is stable: no trajectory grows.  Free vibration
    # from velocities of sigma 1e-3 m/s moves a model by about sigma times
    # the elapsed time at most (100 x that is allowed); an unstable model
    # grows without bound (1e62 in 500 steps for the one whose dt came from
    # an unconverged omega_max)
    for k in range(nm):
        uk = float(np.abs(r["U"][bt.model_slice(k)]).max())
        assert np.isfinite(uk) and uk <= 100 * 1e-3 * nsteps * dts[k], (k, uk, dts[k])
    # both mixed precisions on the same inputs (the figures the batch
    # throughput is quoted with): final-state error of every model against
    # its FP64 trajectory, max |dU| / max |U|, within 1e-6
    mixed = {}
    for prec in (_lib.FS_PREC_MIXED, _lib.FS_PREC_MIXED_F32):
        rm = bt.explicit(dts, np.zeros(nm), np.zeros(nfree), np.concatenate(V0s), np.zeros(nfree), g, nsteps,
                         precision=prec)
        err = max(float(np.abs(rm["U"][bt.model_slice(k)] - r["U"][bt.model_slice(k)]).max()
                        / np.abs(r["U"][bt.model_slice(k)]).max()) for k in range(nm))
        print(f"[gpu] C3 full batch: precision {prec} max final-state error {err:.2e}", flush=True)
        assert err <= 1e-6, (prec, err)
        mixed[prec] = rm["U"][bt.model_slice(big)].copy()
    bt.close()
    m = models[big]
    d = m["dchi"]
    nf = d.nfreedofs
    Dps, Dt = O.shell_D(O.iso_D6(m["material"].E, m["material"].nu))
    nrm, val = O.normals(3, m["xyz"], m["conn"])
    Ko = O.assemble_csr(3, m["conn"], d.dofnums, nf,
                        O.element_matrices(3, m["xyz"], m["conn"], nrm, val, Dps=Dps, Dt=Dt, t=m["t"]))
    assert np.array_equal(Ko.indices, Ks[big].indices)
    assert np.linalg.norm(Ks[big].data - Ko.data) <= 1e-12 * np.linalg.norm(Ko.data)
    z = np.zeros(nf)
    Uo, Vo, _ = O.cd_loop(Ks[big], Ms[big], 0.0, dts[big], nsteps, z, g[:, big], z, V0s[big])
    assert np.linalg.norm(runs[big]["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo)
    assert np.linalg.norm(runs[big]["V"] - Vo) <= 1e-10 * np.linalg.norm(Vo)
    for prec, Um in mixed.items():
        assert np.abs(Um - Uo).max() <= 1e-6 * np.abs(Uo).max(), prec


def test_q4_batch_equals_models_alone(ctx):
    """Q4RS models of different sizes, thicknesses and materials in one
    batch (one material group per model, so the element kernel reads each
    element's group) against each model alone (one group): bitwise the same
    stiffness, and the oracle's element matrices within 1e-12."""
    print("[gpu] Q4 batch vs alone", flush=True)
    from flexshell.femm import FEMMShellQ4RS, SysmatAssemblerFFBlock
    from flexshell.material import MatDeforElastIso
    models = []
    for k, (n, t, E) in enumerate([(6, 0.005, 70e9), (9, 0.02, 200e9), (7, 0.001, 110e9)]):
        m = W.cylinder_panel_c2(n=n, seed=40 + k, t=t, E=E)
        m["material"] = MatDeforElastIso(E, 0.3 + 0.01 * k, 2700.0)
        models.append(m)
    bt = B.ShellBatch(models, ctx=ctx)
    Kb = bt.stiffness()
    for k, m in enumerate(models):
        f = FEMMShellQ4RS(m["conn"], m["material"], thickness=m["t"], ctx=ctx)
        f.associategeometry(m["xyz"])
        nf = m["dchi"].nfreedofs
        Kk = f.stiffness(SysmatAssemblerFFBlock(nf), m["xyz"], None, None, m["dchi"])
        sl = bt.model_slice(k)
        Kbk = Kb[sl, sl]
        assert np.array_equal(Kbk.indices, Kk.indices) and np.array_equal(Kbk.data, Kk.data), k
        Emod, nu = m["material"].E, m["material"].nu
        Dps, Dt = O.shell_D(O.iso_D6(Emod, nu))
        Ko = O.element_matrices(4, m["xyz"], m["conn"], f._normals, f._normal_valid, Dps=Dps, Dt=Dt, t=m["t"])
        Kg = f.element_matrices(m["xyz"])
        err = np.linalg.norm(Kg - Ko, axis=(1, 2)) / np.linalg.norm(Ko, axis=(1, 2))
        assert err.max() < 1e-12, (k, err.max())
        f._model.close()
