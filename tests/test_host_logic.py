"""Host-side logic that needs no GPU: the served-workload generator."""
import numpy as np
import pytest

from flexshell import femm as FM, workloads as W
from shell_cases import distorted_mesh


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
