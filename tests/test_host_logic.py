"""Host-side logic that needs no GPU: layered-shell normals from the layup
coordinate system and the served-workload generator."""
import numpy as np
import pytest

from flexshell import femm as FM, workloads as W
from shell_cases import distorted_mesh


@pytest.mark.parametrize("nen", [3, 4])
def test_layered_csys_normals_vectorized_equal_loop(nen):
    """The vectorized constant-csys normals of the layered FEMMs equal the
    per-element loop of associategeometry! (generic closure path)."""
    xyz, conn = distorted_mesh(nen, 6, 5, seed=11, warp=0.6)
    lay_a = W.c5_layup(3, [0.0, 45.0, 90.0])
    lay_b = W.c5_layup(4, [0.0, -45.0, 45.0, 90.0])
    # a tilted csys for the second group, so that both sum and validity vary
    c, s = np.cos(0.9), np.sin(0.9)
    lay_b.csys = np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=float)
    E = conn.shape[0]
    groups = [(lay_a, np.arange(0, E, 2)), (lay_b, np.arange(1, E, 2))]
    cls = FM.FEMMShellT3FFComp if nen == 3 else FM.FEMMShellQ4RSComp
    f = cls(conn, None, layup_groups=groups)
    fn = f._csys_normal_function()
    n_vec, v_vec = f._host_csys_normals(xyz, fn)
    n_loop, v_loop = FM._ShellFEMM._host_csys_normals(f, xyz, fn)
    assert np.allclose(n_vec, n_loop, rtol=0, atol=1e-14)
    assert np.array_equal(v_vec, v_loop)
    assert (~v_vec).any() and v_vec.any()


def test_served_workload_generator():
    specs = W.served_c5(nreq=200, seed=6, rate=10.0)
    arr = np.array([s["arrival"] for s in specs])
    assert np.all(np.diff(arr) > 0)
    nel = np.array([s["nel"] for s in specs])
    assert nel.min() >= 1e3 and nel.max() <= 5e5
    assert {s["kind"] for s in specs} == {"assembly", "modal", "explicit"}
    assert all(3 <= s["nplies"] <= 8 for s in specs)
    m = W.c5_model(dict(specs[0], nel=2000.0))
    assert 1500 <= m["conn"].shape[0] <= 2600
    lay = m["layup"]
    assert abs(lay.thickness() - specs[0]["nplies"] * 0.125e-3) < 1e-15
    assert W.served_c5(nreq=200, seed=6, rate=10.0) == specs  # seeded: the same stream every time


def test_footprint_grows_with_mesh():
    from flexshell.serve import footprint_bytes
    a = footprint_bytes(3, 1000, 530, 3000, "assembly")
    b = footprint_bytes(3, 100000, 50500, 300000, "assembly")
    c = footprint_bytes(3, 100000, 50500, 300000, "explicit")
    assert a < b < c
    # a 500k-element quad model with 3M dof stays far below one GPU's memory
    assert footprint_bytes(4, 500000, 501000, 3000000, "explicit") < 8e9
