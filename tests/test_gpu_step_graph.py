"""The explicit loop's step graph is kept by the model and replayed by later
calls with the same sizes, dt and load factors: the replay computes bitwise
what the same call computes without a graph (new initial state, damping,
mass and force vectors are data, not graph parameters), and a call that
changes a graph parameter (load factors, dt, step count, precision, probes)
captures a new graph and is again bitwise equal to the ungraphed call."""
import numpy as np
import pytest

from flexshell import _lib, explicit, fe, femm as FM
from flexshell.material import MatDeforElastIso

pytestmark = pytest.mark.gpu


def _model(ctx):
    xy, conn = fe.T3block(1.0, 0.8, 14, 11)
    xyz = fe.xyz3(xy)
    xyz[:, 2] = 0.05 * np.sin(2.0 * xyz[:, 0])
    d = fe.DofField(xyz.shape[0])
    d.setebc(np.nonzero(xy[:, 0] < 1e-9)[0], [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    f = FM.FEMMShellT3FF(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
    f.associategeometry(xyz)
    nf = d.nfreedofs
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, None, None, d)
    f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, d)
    dt = 0.9 * explicit.stable_dt(explicit.omega_max(f, np.random.default_rng(1).random(nf), maxit=40))
    return f, nf, dt, xyz, d


def _same(a, b):
    return np.array_equal(a["U"], b["U"]) and np.array_equal(a["V"], b["V"]) and np.array_equal(a["probes"], b["probes"])


def test_step_graph_replay_bitwise(ctx):
    print("[gpu] step graph: replay vs ungraphed", flush=True)
    f, nf, dt, _, _ = _model(ctx)
    m = f._model
    rng = np.random.default_rng(3)
    nsteps = 300
    g = np.cos(np.arange(nsteps + 1) * 0.02)
    probes = [0, nf // 3, nf - 1]
    runs = [(1e-3 * rng.standard_normal(nf), rng.standard_normal(nf), 0.0) for _ in range(3)]
    runs.append((runs[0][0], runs[0][1], 50.0))  # damping enters as data (C, invMC)
    first = True
    for k, (V0, F, cd) in enumerate(runs):
        a = explicit.run(m, cd, np.zeros(nf), V0, dt, nsteps, F, g, probes, 7, use_graph=True)
        b = explicit.run(m, cd, np.zeros(nf), V0, dt, nsteps, F, g, probes, 7, use_graph=False)
        assert _same(a, b), k
        if first:
            assert a["graph_build_ms"] > 0.0
            first = False
        elif cd == 0.0:
            assert a["graph_build_ms"] == 0.0, k  # replayed
    # graph parameters that change: a new capture, still bitwise the ungraphed call
    V0, F = runs[0][0], runs[0][1]
    for g2, dt2, n2, prec in ((0.5 * g, dt, nsteps, _lib.FS_PREC_FP64), (g, 0.8 * dt, nsteps, _lib.FS_PREC_FP64),
                              (g[:201], dt, 200, _lib.FS_PREC_FP64), (g, dt, nsteps, _lib.FS_PREC_MIXED),
                              (g, dt, nsteps, _lib.FS_PREC_MIXED_F32), (g, dt, nsteps, _lib.FS_PREC_FP64)):
        a = explicit.run(m, 0.0, np.zeros(nf), V0, dt2, n2, F, g2, probes, 7, precision=prec, use_graph=True)
        b = explicit.run(m, 0.0, np.zeros(nf), V0, dt2, n2, F, g2, probes, 7, precision=prec, use_graph=False)
        assert a["graph_build_ms"] > 0.0
        assert _same(a, b)
        a2 = explicit.run(m, 0.0, np.zeros(nf), 2.0 * V0, dt2, n2, F, g2, probes, 7, precision=prec, use_graph=True)
        b2 = explicit.run(m, 0.0, np.zeros(nf), 2.0 * V0, dt2, n2, F, g2, probes, 7, precision=prec, use_graph=False)
        assert a2["graph_build_ms"] == 0.0
        assert _same(a2, b2)
    f._model.close()


def test_step_graph_after_new_stiffness(ctx):
    """A stiffness recomputed with other material constants changes the
    values behind the graph's pointers: the replay uses the new K."""
    print("[gpu] step graph: replay after a new stiffness", flush=True)
    f, nf, dt, xyz, d = _model(ctx)
    m = f._model
    V0 = 1e-3 * np.random.default_rng(4).standard_normal(nf)
    g = np.ones(101)
    z = np.zeros(nf)
    explicit.run(m, 0.0, z, V0, dt, 100, z, g, use_graph=True)
    f.material = MatDeforElastIso(60e9, 0.3, 2700.0)
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, None, None, d)
    a = explicit.run(m, 0.0, z, V0, dt, 100, z, g, use_graph=True)
    b = explicit.run(m, 0.0, z, V0, dt, 100, z, g, use_graph=False)
    assert _same(a, b)
    f._model.close()
