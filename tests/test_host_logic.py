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


def test_cache_groups_of_the_batch():
    """step_bytes counts the free-free values of a model from its
    connectivity and free dofs: on the batch's plates (clamped or simply
    supported edges) exactly the stored entries of the free-free block.  The groups
    keep model order, cover every model once and stay within the budget
    unless a model is larger than the budget alone."""
    import scipy.sparse as sp
    from flexshell import batch as B
    models = W.batch_c3(nmodels=12, seed=4, nmin=4, nmax=20)
    for m in models:
        d = m["dchi"]
        nf = d.nfreedofs
        dn = d.dofnums[m["conn"]].reshape(m["conn"].shape[0], -1)
        r = np.repeat(dn, dn.shape[1], axis=1).ravel()
        c = np.tile(dn, (1, dn.shape[1])).ravel()
        keep = (r < nf) & (c < nf)
        A = sp.coo_matrix((np.ones(keep.sum()), (r[keep], c[keep])), shape=(nf, nf)).tocsr()
        A.sum_duplicates()
        exact = A.nnz * 8 + 72 * nf
        assert B.step_bytes(m) == exact
    budget = 3 * max(B.step_bytes(m) for m in models)
    gs = B.cache_groups(models, budget)
    assert [k for g in gs for k in g] == list(range(len(models)))
    for g in gs:
        assert len(g) == 1 or sum(B.step_bytes(models[k]) for k in g) <= budget
    assert B.cache_groups(models, 1) == [[k] for k in range(len(models))]
