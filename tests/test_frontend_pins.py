"""Pin the oracle's load and numbering restatements on the CPU: each
integration rule integrates polynomial force fields of its degree exactly
(total force and first moments, which also fix the shape-function weights),
and the numbering agrees with the vectorized DofField numbering and with the
ordering properties of numberdofs!."""
import numpy as np
import pytest

import shell_cases as SC
from flexshell import fe
from oracle import oracle as O


def _poly_int(i, j, a=1.0, b=0.7):
    """Integral of x^i y^j over [0, a] x [0, b]."""
    return a ** (i + 1) / (i + 1) * b ** (j + 1) / (j + 1)


@pytest.mark.parametrize("nen,npts,deg", [(3, 1, 1), (3, 3, 2), (3, 6, 4), (4, 4, 3)])
def test_rule_exactness(nen, npts, deg):
    """Consistent loads of f_z = x^i y^j: sum of the nodal loads = the
    integral of f, sum of x_a F_a = the integral of x f (the nodal weights
    N_a reproduce x), for every monomial the rule must integrate exactly on a
    straight-sided mesh (T3: total degree; Q4 on rectangles: per variable)."""
    # the unit-width rectangle [0, 1] x [0, 0.7]; T3: interior nodes jittered
    # (straight-sided triangles), Q4: rectangles
    if nen == 3:
        xy, conn = fe.T3block(1.0, 0.7, 5, 4)
        inner = (xy[:, 0] > 1e-9) & (xy[:, 0] < 1 - 1e-9) & (xy[:, 1] > 1e-9) & (xy[:, 1] < 0.7 - 1e-9)
        xy[inner] += np.random.default_rng(3).uniform(-0.03, 0.03, (int(inner.sum()), 2))
    else:
        xy, conn = fe.Q4block(1.0, 0.7, 5, 4)
    xyz = fe.xyz3(xy)
    N = xyz.shape[0]
    d = fe.DofField(N).numberdofs()
    pts = O.load_points(nen, xyz, conn, npts)
    for i in range(deg + 1):
        for j in range(deg + 1):
            if (nen == 3 and i + j > deg) or (nen == 4 and max(i, j) > deg):
                continue
            f = np.zeros(pts.shape[:2] + (6,))
            f[..., 2] = pts[..., 0] ** i * pts[..., 1] ** j
            F = O.surface_load(nen, xyz, conn, f, d.nalldofs, d.dofnums, npts)
            Fz = F[d.dofnums[:, 2]]
            ref = _poly_int(i, j)
            assert abs(Fz.sum() - ref) < 1e-13 * max(1.0, abs(ref))
            if (nen == 3 and i + j + 1 <= deg) or (nen == 4 and i + 1 <= deg):
                assert abs((Fz * xyz[:, 0]).sum() - _poly_int(i + 1, j)) < 1e-13
    # a uniform intensity gives f * area in total, whatever the rule
    F = O.surface_load(nen, xyz, conn, [0, 0, 2.5, 0, 0, 0], d.nalldofs, d.dofnums, npts)
    assert abs(F[d.dofnums[:, 2]].sum() - 2.5 * 0.7) < 1e-13


@pytest.mark.parametrize("perm", [False, True])
def test_number_dofs_oracle(perm):
    rng = np.random.default_rng(7)
    N = 300
    fx = rng.random((N, 6)) < 0.3
    p = rng.permutation(N) if perm else None
    dn, nf = O.number_dofs(fx, p)
    d = fe.DofField(N)
    d.is_fixed[:] = fx
    d.numberdofs(p)
    assert nf == d.nfreedofs and np.array_equal(dn, d.dofnums)
    # ordering: along the numbering order of the nodes, free numbers increase
    # from 0 and fixed numbers increase from nfree
    order = np.arange(N) if p is None else p
    seq = dn[order].ravel()
    fseq = fx[order].ravel()
    assert np.array_equal(seq[~fseq], np.arange(nf))
    assert np.array_equal(seq[fseq], nf + np.arange(fseq.sum()))
