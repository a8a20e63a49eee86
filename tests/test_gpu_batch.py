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
    # one model against the oracle
    k = 3
    Uo, Vo, _ = O.cd_loop(Ks[k], Ms[k], cd[k], dts[k], nsteps, Fm[k], gv[:, k], np.zeros_like(V0s[k]), V0s[k])
    assert np.linalg.norm(plain[k]["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo)
    bt.close()
