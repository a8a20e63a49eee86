"""Mesh input and result export (flexshell.meshio), CPU only: the Abaqus
reader on the reference's fixtures (node and element counts, set
generation), compaction, Q4toT3 orientations, node merging, edge selection
and the VTK / CSV writers."""
import os

import numpy as np

import shell_cases as SC
from flexshell import fe, meshio

INP = os.path.join(SC.GOLDEN, "inp")


def test_import_abaqus_raasch():
    m = meshio.import_abaqus(os.path.join(INP, "raasch_s4_1x9.inp"))
    assert m["xyz"].shape == (1916, 3)
    et, conn = m["fesets"][0]
    assert et == "S4" and conn.shape == (9, 4)
    # element 1 of the file: nodes 1, 11, 20011, 20001 (labels) -> positions
    lab = m["labels"]
    assert list(lab[conn[0]]) == [1, 11, 20011, 20001]
    assert list(m["elsets"]["EALL"]) == list(range(1, 10))
    assert len(m["nsets"]["BOT"]) == 91  # *Nset, generate 1, 91, 1
    xyz, c2, new = meshio.compact_nodes(m["xyz"], conn)
    assert xyz.shape[0] == 20 and c2.max() == 19
    assert np.array_equal(xyz[c2], m["xyz"][conn])
    assert (new >= 0).sum() == 20


def test_q4_to_t3_orientations():
    conn = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    t = meshio.q4_to_t3(conn)
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]]
    t = meshio.q4_to_t3(conn, "alternate")
    assert t.tolist() == [[0, 1, 3], [2, 3, 1], [4, 5, 7], [6, 7, 5]]


def test_merge_nodes():
    """Two 1 x 1 Q4 patches side by side with their own copies of the shared
    edge nodes (offset below the tolerance) merge into one 2 x 1 mesh."""
    xy1, c1 = fe.Q4block(1.0, 1.0, 1, 1)
    xy2, c2 = fe.Q4block(1.0, 1.0, 1, 1)
    xy2[:, 0] += 1.0 + 1e-9
    xyz = fe.xyz3(np.vstack([xy1, xy2]))
    conn = np.vstack([c1, c2 + 4])
    x, c = meshio.merge_nodes(xyz, conn, 1e-6)
    assert x.shape[0] == 6
    assert len(np.unique(c)) == 6
    assert fe.boundary_edges(c).shape[0] == 6  # the shared edge is interior now


def test_select_edges_all_nodes_in_box():
    xy, conn = fe.Q4block(2.0, 1.0, 4, 2)
    xyz = fe.xyz3(xy)
    e = meshio.select_edges(xyz, fe.boundary_edges(conn), [2, 2, 0, 1, 0, 0], 1e-6)
    assert e.shape[0] == 2 and np.allclose(xyz[e][:, :, 0], 2.0)


def test_vtk_and_csv_writers(tmp_path):
    xy, conn = fe.T3block(1.0, 1.0, 2, 2)
    xyz = fe.xyz3(xy)
    p = tmp_path / "m.vtk"
    meshio.vtk_write(str(p), xyz, conn, scalars=[("w", xyz[:, 0])], vectors=[("u", xyz)],
                     cell_scalars=[("e", np.arange(conn.shape[0]))])
    txt = p.read_text().split("\n")
    assert txt[4] == f"POINTS {xyz.shape[0]} double"
    i = txt.index(f"CELLS {conn.shape[0]} {4 * conn.shape[0]}")
    assert [int(v) for v in txt[i + 1].split()] == [3] + conn[0].tolist()
    assert f"CELL_TYPES {conn.shape[0]}" in txt and "SCALARS w double 1" in txt and "VECTORS u double" in txt
    q = tmp_path / "h.csv"
    meshio.csv_write(str(q), [np.arange(3), np.arange(3) * 0.5], ["t", "u"])
    d = np.loadtxt(str(q), delimiter=",", skiprows=1)
    assert d.shape == (3, 2) and d[2, 1] == 1.0
