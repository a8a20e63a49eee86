"""Mesh input and result export around the shell path (host side).

- import_abaqus: the nodes, shell elements and node / element sets of an
  Abaqus input file (FinEtools MeshImportModule.import_ABAQUS, as
  test/test_shell_statics.jl:158-160 uses it), with element connectivity
  mapped from node labels to 0-based node positions in file order.
- compact_nodes: findunconnnodes + compactnodes + renumberconn!
  (test_shell_statics.jl:162-164): drops the nodes no element uses.
- q4_to_t3: FinEtools Q4toT3 (test_shell_statics.jl:166), two triangles
  (1, 2, 3) and (1, 3, 4) per quadrilateral, or (1, 2, 4) and (3, 4, 2)
  with orientation "alternate".
- merge_nodes: FinEtools mergenodes (nodes closer than a tolerance become
  one, test_q4rs_shell_statics.jl:591).
- select_edges: selectelem on boundary edges with a box, every node of a
  selected edge inside the inflated box (test_shell_statics.jl:220).
- vtk_write / csv_write: VTK unstructured-grid (legacy ASCII) and CSV
  export of nodal fields and probe histories (vtkwrite,
  plate_expl_examples.jl:240-241, 317).
"""
import numpy as np

_SHELL_TYPES = {"S3": 3, "S3R": 3, "STRI3": 3, "S3RS": 3, "S4": 4, "S4R": 4, "S4RS": 4, "S4R5": 4}


def _keyword(line):
    """Keyword name and its parameters of an Abaqus '*' line."""
    parts = [p.strip() for p in line[1:].split(",")]
    params = {}
    for p in parts[1:]:
        if "=" in p:
            k, v = p.split("=", 1)
            params[k.strip().lower()] = v.strip()
        elif p:
            params[p.lower()] = True
    return parts[0].lower(), params


def import_abaqus(path):
    """Nodes, shell element sets and node / element sets of an Abaqus input
    file.  Returns dict(xyz=N x 3, labels=N node labels, fesets=[(type,
    conn E x nen 0-based)], elem_labels=[E labels per set], nsets={name:
    0-based node positions}, elsets={name: element labels})."""
    xyz, labels = [], []
    fesets, elem_labels = [], []
    nsets, elsets = {}, {}
    mode, cur, gen, name = None, None, False, None
    with open(path) as fh:
        lines = fh.read().splitlines()
    for raw in lines:
        line = raw.strip()
        if not line or line.startswith("**"):
            continue
        if line.startswith("*"):
            kw, params = _keyword(line)
            mode, gen = None, False
            if kw == "node":
                mode = "node"
            elif kw == "element":
                et = str(params.get("type", "")).upper()
                if et not in _SHELL_TYPES:
                    raise ValueError(f"import_abaqus: element type {et} is not a 3- or 4-node shell")
                cur = (et, [], [])
                fesets.append(cur)
                mode = "element"
            elif kw == "nset":
                name = params["nset"]
                nsets.setdefault(name, [])
                mode, gen = "nset", "generate" in params
            elif kw == "elset":
                name = params["elset"]
                elsets.setdefault(name, [])
                mode, gen = "elset", "generate" in params
            continue
        vals = [v.strip() for v in line.split(",") if v.strip()]
        if mode == "node":
            labels.append(int(vals[0]))
            xyz.append([float(v) for v in vals[1:4]] + [0.0] * (3 - len(vals[1:4])))
        elif mode == "element":
            nen = _SHELL_TYPES[cur[0]]
            cur[1].append(int(vals[0]))
            cur[2].append([int(v) for v in vals[1:1 + nen]])
        elif mode in ("nset", "elset"):
            target = nsets if mode == "nset" else elsets
            if gen:
                a, b, s = (int(v) for v in vals[:3])
                target[name].extend(range(a, b + 1, s))
            else:
                target[name].extend(int(v) for v in vals)
    labels = np.asarray(labels, dtype=np.int64)
    pos = {int(l): i for i, l in enumerate(labels)}
    out_sets = []
    for et, el, conn in fesets:
        c = np.array([[pos[n] for n in row] for row in conn], dtype=np.int64).reshape(-1, _SHELL_TYPES[et])
        out_sets.append((et, c))
        elem_labels.append(np.asarray(el, dtype=np.int64))
    return dict(xyz=np.asarray(xyz, dtype=float).reshape(-1, 3), labels=labels, fesets=out_sets,
                elem_labels=elem_labels,
                nsets={k: np.array([pos[n] for n in v if n in pos], dtype=np.int64) for k, v in nsets.items()},
                elsets={k: np.asarray(v, dtype=np.int64) for k, v in elsets.items()})


def compact_nodes(xyz, conn):
    """Drop the nodes that no element references, keeping the others in
    their order; returns (xyz, conn, new_numbering) with new_numbering[old]
    = new position or -1."""
    xyz = np.asarray(xyz)
    conn = np.asarray(conn)
    used = np.zeros(xyz.shape[0], dtype=bool)
    used[conn.ravel()] = True
    new = np.full(xyz.shape[0], -1, dtype=np.int64)
    new[used] = np.arange(int(used.sum()))
    return xyz[used].copy(), new[conn], new


def q4_to_t3(conn, orientation="default"):
    """Two triangles per quadrilateral, in element order."""
    conn = np.asarray(conn)
    if orientation == "alternate":
        a, b = (0, 1, 3), (2, 3, 1)
    else:
        a, b = (0, 1, 2), (0, 2, 3)
    out = np.empty((2 * conn.shape[0], 3), dtype=conn.dtype)
    out[0::2] = conn[:, a]
    out[1::2] = conn[:, b]
    return out


def merge_nodes(xyz, conn, tolerance):
    """FinEtools mergenodes: in node order, every node not yet merged takes
    all later nodes whose L1 distance (sum of the absolute coordinate
    differences) to it is below `tolerance`; the kept nodes stay in their
    order.  Returns (xyz, conn)."""
    xyz = np.asarray(xyz, dtype=float)
    N = xyz.shape[0]
    rep = np.full(N, -1, dtype=np.int64)
    order = np.argsort(xyz[:, 0], kind="stable")
    xs = xyz[order, 0]
    for i in range(N):
        if rep[i] >= 0:
            continue
        rep[i] = i
        # candidates: |x_j - x_i| < tolerance (a necessary condition of the L1 test)
        lo = np.searchsorted(xs, xyz[i, 0] - tolerance, side="left")
        hi = np.searchsorted(xs, xyz[i, 0] + tolerance, side="right")
        cand = order[lo:hi]
        cand = cand[(cand > i) & (rep[cand] < 0)]
        if cand.size:
            d = np.abs(xyz[cand] - xyz[i]).sum(axis=1)
            rep[cand[d < tolerance]] = i
    conn = rep[np.asarray(conn)]
    return compact_nodes(xyz, conn)[:2]


def select_edges(xyz, edges, box, inflate=0.0):
    """Edges with every node inside the box [xmin xmax ymin ymax zmin zmax]
    grown by `inflate`."""
    xyz = np.asarray(xyz)
    b = np.asarray(box, dtype=float).reshape(3, 2)
    lo, hi = b[:, 0] - inflate, b[:, 1] + inflate
    inside = np.all((xyz >= lo) & (xyz <= hi), axis=1)
    edges = np.asarray(edges)
    return edges[np.all(inside[edges], axis=1)]


def vtk_write(path, xyz, conn, scalars=(), vectors=(), cell_scalars=()):
    """Legacy ASCII VTK unstructured grid of a T3 (VTK_TRIANGLE) or Q4
    (VTK_QUAD) mesh with named nodal scalars / 3-vectors and cell scalars."""
    xyz = np.asarray(xyz, dtype=float)
    conn = np.asarray(conn, dtype=np.int64)
    nen = conn.shape[1]
    ctype = 5 if nen == 3 else 9
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nflexshell\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {xyz.shape[0]} double\n")
        np.savetxt(f, xyz, fmt="%.17g")
        f.write(f"CELLS {conn.shape[0]} {conn.shape[0] * (nen + 1)}\n")
        np.savetxt(f, np.hstack([np.full((conn.shape[0], 1), nen), conn]), fmt="%d")
        f.write(f"CELL_TYPES {conn.shape[0]}\n")
        np.savetxt(f, np.full(conn.shape[0], ctype), fmt="%d")
        if len(scalars) or len(vectors):
            f.write(f"POINT_DATA {xyz.shape[0]}\n")
            for name, v in scalars:
                f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n")
                np.savetxt(f, np.asarray(v, dtype=float).reshape(-1), fmt="%.17g")
            for name, v in vectors:
                f.write(f"VECTORS {name} double\n")
                np.savetxt(f, np.asarray(v, dtype=float).reshape(-1, 3), fmt="%.17g")
        if len(cell_scalars):
            f.write(f"CELL_DATA {conn.shape[0]}\n")
            for name, v in cell_scalars:
                f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n")
                np.savetxt(f, np.asarray(v, dtype=float).reshape(-1), fmt="%.17g")


def csv_write(path, columns, names):
    """CSV of equal-length columns (e.g. time and probe histories)."""
    data = np.column_stack([np.asarray(c, dtype=float) for c in columns])
    np.savetxt(path, data, delimiter=",", header=",".join(names), comments="", fmt="%.17g")
