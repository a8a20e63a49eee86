"""Row products of the FP64 explicit step and the power iteration: the
node-grouped form (lanes per node, positions per lane) against the row form
and the oracle.

The forms add the same products in different orders, so the trajectories
agree to rounding (bound 1e-12 relative) and each form is bitwise
reproducible.  Meshes: a Q4 panel with simply supported edges (runs of three
free dofs next to runs of six), a T3 plate with a node permutation (runs not
in column order), and a T3 fan whose hub node has 48 neighbours (rows of 294
entries, several passes of the lanes over one node)."""
import numpy as np
import pytest

from flexshell import _lib, explicit, fe, femm as FM
from flexshell.material import MatDeforElastIso
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORMS = [(0, 0), (8, 0), (8, 4), (8, 5), (8, 6), (8, 7), (8, 8), (16, 0), (16, 2), (16, 3), (16, 4)]


def _panel_q4():
    xy, conn = fe.Q4block(1.0, 0.6, 10, 7)
    xyz = fe.xyz3(xy)
    xyz[:, 2] = 0.1 * np.sin(np.pi * xyz[:, 0])  # curved, so all six components couple
    d = fe.DofField(xyz.shape[0])
    d.setebc(np.nonzero(xy[:, 0] < 1e-9)[0], [1, 2, 3])
    d.setebc(np.nonzero(xy[:, 0] > 1 - 1e-9)[0], [1, 2, 3])
    d.numberdofs()
    return 4, xyz, conn, d


def _plate_t3_perm():
    xy, conn = fe.T3block(1.0, 1.0, 9, 8)
    xyz = fe.xyz3(xy)
    d = fe.DofField(xyz.shape[0])
    edge = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 1] < 1e-9))[0]
    d.setebc(edge, [1, 2, 3, 4, 5, 6])
    d.numberdofs(perm=np.random.default_rng(5).permutation(xyz.shape[0]))
    return 3, xyz, conn, d


def _fan_t3(m=48):
    ang = 2 * np.pi * np.arange(m) / m
    r1, r2 = 0.5, 1.0
    xyz = np.vstack([[0.0, 0.0, 0.0],
                     np.column_stack([r1 * np.cos(ang), r1 * np.sin(ang), np.zeros(m)]),
                     np.column_stack([r2 * np.cos(ang + np.pi / m), r2 * np.sin(ang + np.pi / m), np.zeros(m)])])
    conn = []
    for i in range(m):
        a, b = 1 + i, 1 + (i + 1) % m
        oa, ob = 1 + m + i, 1 + m + (i + 1) % m
        conn += [[0, a, b], [a, oa, b], [b, oa, ob]]
    d = fe.DofField(xyz.shape[0])
    d.setebc(np.arange(1 + m, 1 + 2 * m), [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    return 3, xyz, np.array(conn, dtype=np.int64), d


@pytest.mark.parametrize("mesh", [_panel_q4, _plate_t3_perm, _fan_t3], ids=["q4_ss", "t3_perm", "t3_fan"])
def test_cd_forms_agree(ctx, mesh):
    print(f"[gpu] cd forms {mesh.__name__}", flush=True)
    nen, xyz, conn, d = mesh()
    cls = FM.FEMMShellT3FF if nen == 3 else FM.FEMMShellQ4RS
    f = cls(conn, MatDeforElastIso(210e9, 0.3, 7850.0), thickness=0.01, ctx=ctx)
    f.associategeometry(xyz)
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d)
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, None, None, d)
    f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), xyz, d)
    rng = np.random.default_rng(11)
    U0, V0, Fmag = rng.normal(0, 1e-6, nf), rng.normal(0, 1e-3, nf), rng.normal(0, 1.0, nf)
    dt = 0.9 * explicit.stable_dt(explicit.omega_max(f, rng.random(nf), maxit=40))
    nsteps = 200
    g = np.cos(2 * np.pi * 500 * explicit.sample_times(nsteps, dt))
    cdamp = 50.0
    probes = np.array([0, nf // 3, nf - 1])
    Uo, Vo, _ = O.cd_loop(K, M.diagonal(), cdamp, dt, nsteps, Fmag, g, U0, V0, probes)
    v0 = rng.random(nf)
    oo = O.pwr(K, M.diagonal(), v0, maxit=30)
    ref = None
    try:
        for lanes, pj in FORMS:
            ctx.set_cd_lanes(lanes, pj)
            assert abs(explicit.omega_max(f, v0, maxit=30) - oo) < 1e-10 * oo, (lanes, pj)
            r = explicit.run(f._model, cdamp, U0, V0, dt, nsteps, Fmag, g, probes, 1)
            r2 = explicit.run(f._model, cdamp, U0, V0, dt, nsteps, Fmag, g, probes, 1)
            assert np.array_equal(r["U"], r2["U"]) and np.array_equal(r["V"], r2["V"]), (lanes, pj)
            assert np.linalg.norm(r["U"] - Uo) <= 1e-10 * np.linalg.norm(Uo), (lanes, pj)
            if ref is None:
                ref = r
            else:
                assert np.linalg.norm(r["U"] - ref["U"]) <= 1e-12 * np.linalg.norm(ref["U"]), (lanes, pj)
                assert np.linalg.norm(r["V"] - ref["V"]) <= 1e-12 * np.linalg.norm(ref["V"]), (lanes, pj)
    finally:
        ctx.set_cd_lanes(8, 0)


def test_cd_forms_rejected(ctx):
    for lanes, pj in ((4, 0), (32, 0), (8, 3), (8, 9), (16, 1), (16, 5), (-1, 0)):
        with pytest.raises(_lib.FlexShellError):
            ctx.set_cd_lanes(lanes, pj)
    ctx.set_cd_lanes(8, 0)
