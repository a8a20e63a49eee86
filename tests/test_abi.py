"""CPU checks of the C ABI library and the host-side logic (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "flexshell.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from flexshell import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/flexshell.h but not exported"
    # the Python binding attaches a signature to every declared function
    assert set(names) == set(_lib.SIGNATURES)


def test_version_string():
    from flexshell import _lib
    assert b"gfx950" in _lib.load().fs_version()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from flexshell import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_struct_layouts_match_header():
    """ctypes structures mirror the C structs (field order and sizes)."""
    from flexshell import _lib
    assert ctypes.sizeof(_lib.ShellParams) == 4 + 4 + 3 * 8 + 9 * 8 + 4 * 8 + 4 + 4 + 5 * 8
    assert ctypes.sizeof(_lib.MassParams) == 4 + 4 + 2 * 8 + 3 * 8
    assert ctypes.sizeof(_lib.CDParams) == 8 + 8 + 8 + 4 + 4 + 4 + 4 + 3 * 8 + 8


def test_dof_numbering_free_first_node_major():
    from flexshell import fe
    d = fe.DofField(4)
    d.setebc([1], [2, 3])
    d.setebc([3], [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    assert d.nfreedofs == 16
    assert list(d.dofnums[0]) == [0, 1, 2, 3, 4, 5]
    assert list(d.dofnums[1]) == [6, 16, 17, 7, 8, 9]
    assert list(d.dofnums[3]) == list(range(18, 24))
    perm = [2, 0, 3, 1]
    d.numberdofs(perm)
    assert list(d.dofnums[2]) == [0, 1, 2, 3, 4, 5]
    assert d.dofnums[1, 0] == 12


def test_structured_mesh_orientation():
    """T3block :a splits each rectangle along the (i+1, j)-(i, j+1) diagonal;
    element order is x outer, y inner (FinEtools T3blockx)."""
    from flexshell import fe
    xy, conn = fe.T3block(2.0, 1.0, 2, 1)
    assert conn.tolist() == [[0, 1, 3], [1, 4, 3], [1, 2, 4], [2, 5, 4]]
    xy, conn = fe.Q4block(2.0, 2.0, 2, 2)
    assert conn.tolist()[:2] == [[0, 1, 4, 3], [3, 4, 7, 6]]
    assert np.allclose(xy[4], [1.0, 1.0])


def test_host_material_matches_oracle():
    """Host-side plane-stress condensation and laminate matrices equal the oracle's."""
    from flexshell.material import CompositeLayup, MatDeforElastIso, Ply, lamina_material, shell_material_stiffness
    from oracle import oracle as O
    Dps, Dt = shell_material_stiffness(MatDeforElastIso(70e9, 0.33))
    Po, To = O.shell_D(O.iso_D6(70e9, 0.33))
    np.testing.assert_allclose(Dps, Po, rtol=1e-15)
    np.testing.assert_allclose(Dt, To, rtol=1e-15)
    E1, E2, nu12, G12, G13, G23 = 140e9, 10e9, 0.3, 5e9, 5e9, 10e9 / 2.6
    mat = lamina_material(1500.0, E1, E2, nu12, G12, G13, G23)
    angles = [0, 45, -45, 90, 90, -45, 45, 0]
    lay = CompositeLayup("l", [Ply(f"p{i}", mat, 0.125e-3, a) for i, a in enumerate(angles)], np.eye(3))
    A, B, D = lay.laminate_stiffnesses()
    H = lay.laminate_transverse_stiffness()
    D6 = O.lamina_D6(E1, E2, nu12, G12, G13, G23)
    Ao, Bo, Do, Ho, (md, mi) = O.laminate([(D6, 0.125e-3, a, 1500.0) for a in angles])
    for x, y in ((A, Ao), (D, Do), (H, Ho)):
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-13 * np.abs(y).max())
    assert np.abs(B).max() < 1e-12 * np.abs(A).max()
    np.testing.assert_allclose(lay.laminate_inertia(), (md, mi), rtol=1e-14)
