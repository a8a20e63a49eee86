"""Element-kernel launch shapes: every shape the autotuner can pick computes
bitwise the same element blocks; the autotuner records its choice."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import _lib, femm as FM, workloads as W
from flexshell.material import MatDeforElastIso

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nen,layered,form", [(3, False, 0), (3, False, 1), (3, True, 0), (4, False, 0),
                                               (4, True, 0)])
def test_all_launch_shapes_bitwise_equal(nen, layered, form):
    print(f"[gpu] launch shapes nen={nen} layered={layered} form={form}", flush=True)
    ctx = _lib.Context(0)
    xyz, conn = SC.distorted_mesh(nen, 9, 7, seed=3, warp=0.5)  # element counts not multiples of the shapes
    if layered:
        f = FM.make(conn, W.c5_layup(4, [0.0, 45.0, -45.0, 90.0]), ctx=ctx)
    else:
        f = FM.make(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
    f.transv_shear_formulation = form
    f.associategeometry(xyz)
    ref = None
    for v in range(3):
        ctx.set_tuning(nen, form, v)
        blocks = f._model.element_blocks(f._shell_params())
        assert ctx.tuning(nen, form)[0] == v
        if ref is None:
            ref = blocks
        else:
            assert np.array_equal(blocks, ref), v
    ctx.set_tuning(nen, form, -1)  # re-arm: the next evaluation tunes
    f._model.element_blocks(f._shell_params())
    v, ne, wg = ctx.tuning(nen, form)
    assert 0 <= v < 3 and ne > 0 and wg >= ne
    f._model.close()
    ctx.close()


@pytest.mark.parametrize("nen,shuffle", [(3, False), (4, False), (3, True), (4, True)])
def test_assembly_pipeline_bitwise_equal(nen, shuffle):
    """The stiffness assembled with the element kernel split into 1, 2, 5
    and 64 ranges (scatter overlapping the later ranges) is bitwise the same
    matrix, for the all-dof and the free-free block and for every launch
    shape.  With shuffled element numbering the last contributing element of
    the node pairs jumps around, so most pairs complete only in late ranges."""
    print(f"[gpu] assembly pipeline nen={nen} shuffle={shuffle}", flush=True)
    from flexshell import fe
    ctx = _lib.Context(0)
    xyz, conn = SC.distorted_mesh(nen, 23, 17, seed=5, warp=0.4)
    if shuffle:
        conn = conn[np.random.default_rng(1).permutation(conn.shape[0])]
    dch = fe.DofField(xyz.shape[0])
    dch.setebc(fe.selectnode(xyz, [-np.inf, 1e-9, -np.inf, np.inf, -np.inf, np.inf], 1e-6), [1, 2, 3, 4, 5, 6])
    dch.numberdofs()
    f = FM.make(conn, MatDeforElastIso(70e9, 0.33, 2700.0), thickness=0.01, ctx=ctx)
    f.associategeometry(xyz)
    ctx.set_assembly(0)  # element blocks through HBM + the scatter: the form the pipeline applies to
    ref = {}
    for v in range(3):
        ctx.set_tuning(nen, 0, v)
        for ranges in (1, 2, 5, 64):
            ctx.set_pipeline(ranges)
            K = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), xyz, None, None, dch)
            Kf = f.stiffness(FM.SysmatAssemblerFFBlock(dch.nfreedofs), xyz, None, None, dch)
            t = [f._model.timing(i) for i in range(3)]
            assert t[0] >= 0 and t[1] >= 0 and abs(t[0] + t[1] - t[2]) <= 1e-6 * max(1.0, t[2])
            for name, A in (("all", K), ("ff", Kf)):
                A = A.tocsr()
                if name not in ref:
                    ref[name] = A
                else:
                    R = ref[name]
                    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
                    assert np.array_equal(A.data, R.data), (name, v, ranges)
    with pytest.raises(_lib.FlexShellError):
        ctx.set_pipeline(0)
    ctx.set_pipeline(1)
    f._model.close()
    ctx.close()
