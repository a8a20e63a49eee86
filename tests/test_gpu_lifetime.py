"""Context and model lifetime, and the state a failing call leaves behind.

- A context destroyed before its models stays valid until the last model is
  destroyed (Julia finalizers and Python garbage collection run in no fixed
  order).
- A call that fails after it has queued kernels (here: a degenerate Q4
  element found by the element kernel, with the assembly pipeline's element
  ranges and scatters on both streams) returns with that work finished, so
  the device buffers it hands back to the context are idle: the next models
  on the same context compute bitwise what a fresh context computes.
"""
import numpy as np
import pytest

from flexshell import _lib, fe, femm as FM
from flexshell.material import MatDeforElastIso

pytestmark = pytest.mark.gpu


def _q4_case(n=24, seed=0):
    xy, conn = fe.Q4block(1.0, 0.7, n, n)
    xyz = fe.xyz3(xy)
    xyz[:, 2] = 0.1 * np.sin(2.0 * xyz[:, 0]) * xyz[:, 1]
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], 1e-6), [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    return xyz, conn, d


def _stiffness(ctx, xyz, conn, d):
    f = FM.FEMMShellQ4RS(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
    f.associategeometry(xyz)
    nf = d.nfreedofs
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d).diagonal()
    return f, K, M


def test_context_destroyed_before_model(ctx):
    print("[gpu] lifetime: context destroyed before its model", flush=True)
    xyz, conn, d = _q4_case()
    c2 = _lib.Context(0)
    f, K2, _ = _stiffness(c2, xyz, conn, d)
    model = f._model
    c2.close()                 # the model keeps the context alive
    model.close()              # the last reference tears it down
    _, K, _ = _stiffness(ctx, xyz, conn, d)
    assert np.array_equal(K.data, K2.data)


def test_failed_call_leaves_context_usable():
    print("[gpu] lifetime: failing stiffness (degenerate Q4) with 4 pipeline ranges", flush=True)
    xyz, conn, d = _q4_case()
    ref = _lib.Context(0)
    _, K_ref, M_ref = _stiffness(ref, xyz, conn, d)
    ref.close()
    c = _lib.Context(0)
    c.set_pipeline(4)
    bad = xyz.copy()
    e = conn.shape[0] // 2
    bad[conn[e, 2]] = bad[conn[e, 1]]        # element e collapses to a segment: singular metric
    bad[conn[e, 3]] = bad[conn[e, 0]]
    fb = FM.FEMMShellQ4RS(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=c)
    fb.associategeometry(bad)
    with pytest.raises(_lib.FlexShellError, match="Singular metric"):
        fb.stiffness(FM.SysmatAssemblerFFBlock(d.nfreedofs), bad, None, None, d)
    fb._model.close()
    for _ in range(2):
        _, K, M = _stiffness(c, xyz, conn, d)
        assert np.array_equal(K.data, K_ref.data) and np.array_equal(M, M_ref)
    c.close()
