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
