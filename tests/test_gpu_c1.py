"""The CPU-reference configuration run in full through the device path: flat
steel plate 1 x 1 m, t 0.01, T3FF 32 x 32 (2048 triangles, 1089 nodes), all
edges clamped.  As test_shell_statics.jl:84-89 and test_shell_dynamics.jl:80
use it: the free-free K and lumped M, a static solve under a uniform
pressure of 1e3 Pa with the load from distribloads, and the 10 lowest
natural frequencies from a host eigensolver on the GPU-assembled matrices.

Bars: CSR pattern bit-exact and values within 1e-12 (relative Frobenius) of
the oracle's COO assembly; M within 1e-14 per entry; load within 1e-14 of
the oracle's distribloads; displacements within 1e-10 of the oracle's solve;
frequencies within 1e-9 of those of the oracle's matrices.  Two physical
checks tie the numbers to the plate: the centre deflection and the
fundamental frequency of a clamped square Kirchhoff plate (Timoshenko:
w = 0.00126 q a^4 / D; f1 = 35.99 / (2 pi a^2) sqrt(D / (rho t))) within 3 %.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from flexshell import femm as FM, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _lowest_frequencies(K, Mdiag, k=10):
    w = sla.eigsh(K.tocsc(), k=k, M=sp.diags(Mdiag).tocsc(), sigma=0.0, which="LM", return_eigenvectors=False)
    return np.sqrt(np.sort(w)) / (2 * np.pi)


def test_c1_plate_32x32_full(ctx):
    print("[gpu] C1: T3FF plate 32x32 clamped, K, M, static solve, 10 frequencies", flush=True)
    c = W.plate_c1(n=32)
    d, xyz, conn, t = c["dchi"], c["xyz"], c["conn"], c["t"]
    mat = c["material"]
    nf = d.nfreedofs
    assert conn.shape[0] == 2048 and xyz.shape[0] == 1089
    f = FM.FEMMShellT3FF(conn, mat, thickness=t, ctx=ctx)
    f.associategeometry(xyz)
    K = f.stiffness(FM.SysmatAssemblerFFBlock(nf), xyz, None, None, d)
    M = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), xyz, d).diagonal()

    Dps, Dt = O.shell_D(O.iso_D6(mat.E, mat.nu))
    nrm, val = O.normals(3, xyz, conn)
    Ko = O.assemble_csr(3, conn, d.dofnums, nf, O.element_matrices(3, xyz, conn, nrm, val, Dps=Dps, Dt=Dt, t=t))
    assert np.array_equal(K.indptr, Ko.indptr) and np.array_equal(K.indices, Ko.indices)
    rel = np.linalg.norm(K.data - Ko.data) / np.linalg.norm(Ko.data)
    assert rel <= 1e-12, rel
    tm, rm = O.element_masses(3, xyz, conn, rho=mat.rho, t=t)
    Mo = O.lumped_diag(3, conn, d.dofnums, nf, tm, rm)
    np.testing.assert_allclose(M, Mo, rtol=1e-14, atol=0)

    q = 1e3
    force = np.array([0.0, 0.0, -q, 0.0, 0.0, 0.0])
    F = f.distribloads(xyz, d, force, npts=3)
    Fo = O.surface_load(3, xyz, conn, force, d.nalldofs, d.dofnums, 3)[:nf]
    assert np.abs(F - Fo).max() <= 1e-14 * np.abs(Fo).max()
    U = sla.spsolve(K.tocsc(), F)
    Uo = sla.spsolve(Ko.tocsc(), Fo)
    assert np.linalg.norm(U - Uo) <= 1e-10 * np.linalg.norm(Uo)
    Dflex = mat.E * t ** 3 / (12 * (1 - mat.nu ** 2))
    centre = np.argmin(np.linalg.norm(xyz - [0.5, 0.5, 0.0], axis=1))
    wc = d.scattersysvec(U)[centre, 2]
    w_kirchhoff = -0.00126 * q * 1.0 ** 4 / Dflex
    assert abs(wc / w_kirchhoff - 1) < 0.03, (wc, w_kirchhoff)

    fr = _lowest_frequencies(K, M)
    fro = _lowest_frequencies(Ko, Mo)
    print(f"[gpu] C1: frequencies {np.array2string(fr, precision=4)} Hz; max rel diff "
          f"{np.abs(fr - fro).max() / fro.max():.1e}", flush=True)
    np.testing.assert_allclose(fr, fro, rtol=1e-9, atol=0)
    f1 = 35.99 / (2 * np.pi * 1.0 ** 2) * np.sqrt(Dflex / (mat.rho * t))
    assert abs(fr[0] / f1 - 1) < 0.03, (fr[0], f1)
