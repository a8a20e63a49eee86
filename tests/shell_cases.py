"""Reference test scenarios restated as plain arrays (meshes, BCs, loads),
shared by the oracle pin tests and the GPU parity tests.  Values quoted from
the reference's test files are cited next to them."""
import os

import numpy as np

from flexshell import fe
from flexshell.material import cartesian_csys
from oracle import oracle as O

mm = 1e-3
MPa = 1e6
GPa = 1e9

# test/test_shell_dynamics.jl:113-128 (elastic modes 7..14)
FV12_REF = [1.572130183778014, 2.2424585076387427, 2.8079394352847316, 3.883763676656034, 4.039123204140305,
            6.787320617260535, 6.920636670319986, 7.127888889722697]
# test/test_shell_statics.jl:98-105 and test/test_q4rs_shell_statics.jl:94-101
SCORDELIS_T3 = dict(zip([4, 8, 10, 12, 16, 24], [66.54771615057949, 85.54615143134853, 89.85075281481419,
                                                  92.50616661644985, 95.40469210310079, 97.65376880486126]))
SCORDELIS_Q4 = dict(zip([4, 8, 10, 12, 16, 24], [97.88098976068304, 98.766679964194, 99.1305903792941,
                                                  99.33597700449404, 99.52904556664862, 99.63694998474682]))
# test/test_composite_shell_statics.jl:97, test/test_composite_shell_statics_q4rs.jl:97
BARBERO_T3 = 0.22004349767718365
BARBERO_Q4 = 0.2278585006007462
# test/test_composite_shell_dynamics.jl:432-442
NAYAK_REF = np.array([0.04571652264814249, 0.10096323319412286, 0.11467782865238098, 0.16264669289730896,
                      0.18683613787902217, 0.20655080934826295, 0.23954509827380233, 0.2476225295798577,
                      0.2718218525817157])


def fv12(n=8):
    """NAFEMS FV12 free plate (test/test_shell_dynamics.jl:26-73)."""
    L = 10.0
    xy, conn = fe.T3block(L, L, n, n)
    xy[:, 0] -= L / 2
    xy[:, 1] -= L / 2
    xyz = fe.xyz3(xy)
    dch = fe.DofField(xyz.shape[0]).numberdofs()
    return dict(nen=3, xyz=xyz, conn=conn, dchi=dch, E=200e9, nu=0.3, rho=8000.0, t=0.05, stab_c=0.2)


def scordelis(nen, n):
    """Scordelis-Lo roof (test/test_shell_statics.jl:14-89, test/test_q4rs_shell_statics.jl:14-86)."""
    R, L = 25.0, 50.0
    tol = R / n / 1000
    if nen == 3:
        xy, conn = fe.T3block(40 / 360 * 2 * np.pi, L / 2, n, n)
    else:
        xy, conn = fe.Q4block(40 / 360 * 2 * np.pi, L / 2, n, n)
    a, y = xy[:, 0].copy(), xy[:, 1].copy()
    xyz = np.stack([R * np.sin(a), y, R * (np.cos(a) - 1)], axis=1)
    dch = fe.DofField(xyz.shape[0])
    dch.setebc(fe.selectnode(xyz, [-np.inf, np.inf, 0, 0, -np.inf, np.inf], tol), [1, 3, 5])
    dch.setebc(fe.selectnode(xyz, [-np.inf, np.inf, L / 2, L / 2, -np.inf, np.inf], tol), [2, 4, 6])
    dch.setebc(fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], tol), [1, 5, 6])
    dch.numberdofs()
    s40 = np.sin(40 / 360 * 2 * np.pi) * 25
    nl = fe.selectnode(xyz, [s40, s40, L / 2, L / 2, -np.inf, np.inf], tol)[0]
    F = O.surface_load(nen, xyz, conn, [0, 0, -90, 0, 0, 0], dch.nalldofs, dch.dofnums)
    return dict(nen=nen, xyz=xyz, conn=conn, dchi=dch, E=4.32e8, nu=0.0, rho=0.0, t=0.25, stab_c=0.2, F=F,
                probe=nl)


def barbero(nen):
    """Barbero Example 3.1 cross-ply [0/90] plate (test/test_composite_shell_statics.jl:13-98)."""
    ax = ay = 2000 * mm
    nx = ny = 8
    E1, E2, G12, nu12, G23 = 133860 * MPa, 7706 * MPa, 4306 * MPa, 0.301, 2760 * MPa
    G13 = G12
    th = 10 * mm
    tol = ax / nx / 100
    D6 = O.lamina_D6(E1, E2, nu12, G12, G13, G23)
    if nen == 3:
        xy, conn = fe.T3block(ax, ay, nx, ny)
    else:
        xy, conn = fe.Q4block(ax, ay, nx, ny)
    xyz = fe.xyz3(xy)
    dch = fe.DofField(xyz.shape[0])
    dch.setebc(fe.selectnode(xyz, [ax, ax, 0, 0, -np.inf, np.inf], tol), [1, 2, 3])
    dch.setebc(fe.selectnode(xyz, [0, 0, 0, 0, -np.inf, np.inf], tol), [2])
    dch.setebc(fe.boundary_nodes(conn), [3])
    dch.numberdofs()
    bed = fe.boundary_edges(conn)

    def edges_at(y):
        sel = set(fe.selectnode(xyz, [-np.inf, np.inf, y, y, 0, 0], tol).tolist())
        return np.array([e for e in bed if e[0] in sel and e[1] in sel])

    F = fe.edge_load(xyz, edges_at(0.0), [0, 0.1 * MPa * th, 0, 0, 0, 0], dch.nalldofs, dch.dofnums)
    F += fe.edge_load(xyz, edges_at(ay), [0, -0.1 * MPa * th, 0, 0, 0, 0], dch.nalldofs, dch.dofnums)
    plies = [(D6, th / 2, 0.0, 0.0), (D6, th / 2, 90.0, 0.0)]
    return dict(nen=nen, xyz=xyz, conn=conn, dchi=dch, plies=plies, csys=np.eye(3), F=F,
                lamina=(E1, E2, nu12, G12, G13, G23), th=th)


def nayak(nplies, axes):
    """Nayak 4.3 orthotropic plate (test/test_composite_shell_dynamics.jl:341-443)."""
    ax = ay = 100 * mm
    nx = ny = 9
    E1, E2, G12, G13, nu12, G23 = 143.52 * GPa, 75.38 * GPa, 42.03 * GPa, 25.56 * GPa, 0.44, 42.65 * GPa
    rho, C11 = 1500.0, 159.85 * GPa
    th = ax / 10
    tol = ax / nx / 100
    D6 = O.lamina_D6(E1, E2, nu12, G12, G13, G23)
    xy, conn = fe.T3block(ax, ay, nx, ny)
    xyz = fe.xyz3(xy)
    dch = fe.DofField(xyz.shape[0])
    dch.setebc(fe.selectnode(xyz, [ax, ax, 0, 0, -np.inf, np.inf], tol), [1, 2, 3])
    dch.setebc(fe.selectnode(xyz, [0, 0, 0, 0, -np.inf, np.inf], tol), [2])
    dch.setebc(fe.boundary_nodes(conn), [1, 2, 3])
    dch.numberdofs()
    plies = [(D6, th / nplies, 0.0, rho)] * nplies
    return dict(nen=3, xyz=xyz, conn=conn, dchi=dch, plies=plies, csys=cartesian_csys(axes), th=th, rho=rho,
                C11=C11, lamina=(E1, E2, nu12, G12, G13, G23))


def distorted_mesh(nen, nx, ny, seed, warp=0.0, jitter=0.2):
    """Small structured mesh with jittered nodes and an optional out-of-plane
    warp (curved surface), for element-level comparisons."""
    rng = np.random.default_rng(seed)
    if nen == 3:
        xy, conn = fe.T3block(1.0, 0.7, nx, ny)
    else:
        xy, conn = fe.Q4block(1.0, 0.7, nx, ny)
    h = min(1.0 / nx, 0.7 / ny)
    xy = xy + rng.uniform(-jitter * h, jitter * h, xy.shape)
    z = warp * (xy[:, 0] ** 2 - 0.5 * xy[:, 1] ** 2 + 0.3 * xy[:, 0] * xy[:, 1])
    xyz = np.stack([xy[:, 0], xy[:, 1], z], axis=1)
    return xyz, conn

# test/test_shell_statics.jl:226-234: Raasch hook, T3FF on the Abaqus S4
# meshes split by Q4toT3, percentage of the analytical 5.02 (rtol 1e-4)
RAASCH_T3 = {"1x9": 91.7059961843231, "3x18": 95.9355786538892, "5x36": 97.19276899988246,
             "10x72": 98.38896641657374}
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def raasch(mesh):
    """Raasch hook (test/test_shell_statics.jl:148-224): E 3300, nu 0.35,
    t 2, clamped at x = 0, line load 0.05 along z on the tip edge."""
    from flexshell import meshio
    m = meshio.import_abaqus(os.path.join(GOLDEN, "inp", f"raasch_s4_{mesh}.inp"))
    xyz, conn, _ = meshio.compact_nodes(m["xyz"], m["fesets"][0][1])
    conn = meshio.q4_to_t3(conn)
    t = 2.0
    tol = t / 2
    dch = fe.DofField(xyz.shape[0])
    dch.setebc(fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], tol), [1, 2, 3, 4, 5, 6])
    dch.numberdofs()
    tip = meshio.select_edges(xyz, fe.boundary_edges(conn), [97.9615, 97.9615, -16, -16, 0, 20], tol)
    F = fe.edge_load(xyz, tip, [0, 0, 0.05, 0, 0, 0], dch.nalldofs, dch.dofnums)
    probe = fe.selectnode(xyz, [97.9615, 97.9615, -16, -16, 0, 0], tol)[0]
    return dict(nen=3, xyz=xyz, conn=conn, dchi=dch, E=3300.0, nu=0.35, rho=0.0, t=t, stab_c=0.2, F=F,
                probe=probe, ntip=len(tip))


def _csr_solve_max(K, F, d, comp, sign):
    """Solve the free-free system and return max (sign > 0) or min of the
    given nodal component (FinEtools solve_blocked! with zero prescribed
    values)."""
    import scipy.sparse.linalg as sla
    nf = d.nfreedofs
    U = sla.spsolve(K[:nf, :nf].tocsc(), F[:nf])
    v = d.scattersysvec(U)[:, comp]
    return v.max() if sign > 0 else v.min()


# test/test_composite_shell_statics.jl: T3FFComp plates of Barbero's examples
# and variants, reference values of the @test lines (default rtol of isapprox)
COMP_T3 = {
    "c1": 0.010276258430770539,     # :200, 10 pairs [0/90], max w (mm)
    "c2": 3.3161098653904544,       # :323, [+45/-45/0/0/-45/+45], max u (mm)
    "c6": -1502.0324040447633,      # :753, [0/90] 50 x 50, pressure, min w (mm)
    "c7": -1176.8346659735575,      # :844, [-45/45] 50 x 50, pressure, min w (mm)
    "c8+": 0.22004349767718365,     # :954, [90/0], edge load +, max w (mm)
    "c8-": -0.22004349767718365,    # :952, edge load -, min w (mm)
    "c9+": 0.22004349767718365,     # :1070, rotated plate and csys, max u (mm)
    "c9-": -0.22004349767718365,    # :1068, min u (mm)
    "c10w": 1.9529295928169146,     # :1181, Sacco [0/90], normalized max w
    "c10u": 0.12375916912652352,    # :1183, normalized max u
}


# test/test_composite_shell_statics_q4rs.jl: the same setups with Q4RSComp on
# Q4block meshes (pressure integrated with GaussRule(2, 2)); these values are
# printed with fewer digits, so they hold to the @test default rtol (1.5e-8)
COMP_Q4 = {
    "c1": 0.010650045488315574,     # :200
    "c2": 3.3161098653904544,       # :323
    "c6": -1503.4206439821,         # :752
    "c7": -1177.349547131,          # :843
    "c8+": 0.22785850060074,        # :953
    "c8-": -0.22785850060074,       # :951
    "c9+": 0.227858500600,          # :1069
    "c9-": -0.22785850060,          # :1067
    "c10w": 2.0134037916683,        # :1180-1181
    "c10u": 0.128716225,            # :1182-1183
}


def comp_t3(case, nen=3):
    """The setups of mcompshell1/2/6/7/8/9/10 (test_composite_shell_statics.jl:
    117-200, 234-323, 683-753, 774-844, 870-954, 978-1070, 1099-1183) as plain
    arrays: mesh, layup plies (D6, thickness, angle, rho), csys, fixed dofs,
    loads ("edges": list of (box, 6-vector per length); "pressure": 6-vector
    per area with TriRule(1)), and the quantity to compare (component, sign,
    scale)."""
    from flexshell import meshio
    mm, MPa = 1e-3, 1e6
    key = case.rstrip("+-")
    if key in ("c10w", "c10u"):
        key = "c10"
    sign = -1.0 if case.endswith("-") else 1.0
    lam = (133860 * MPa, 7706 * MPa, 0.301, 4306 * MPa, 4306 * MPa, 2760 * MPa)
    ax = ay = 2000 * mm
    nx = ny = 8
    csys = np.eye(3)
    edges, pressure, quantity = [], None, None
    fixes = []
    if key == "c1":
        th = 10 * mm
        angles = [0, 90] * 10
        pt = [th / 20] * 20
        fixes = [([ax, ax, 0, 0, -np.inf, np.inf], [1, 2, 3]), ([0, 0, 0, 0, -np.inf, np.inf], [2]),
                 ("boundary", [3])]
        edges = [([-np.inf, np.inf, 0, 0, 0, 0], [0, 0.1 * MPa * th, 0, 0, 0, 0]),
                 ([-np.inf, np.inf, ay, ay, 0, 0], [0, -0.1 * MPa * th, 0, 0, 0, 0])]
        quantity = (2, +1, 1 / mm)
    elif key == "c2":
        lam = (37880 * MPa, 9407 * MPa, 0.299, 3405 * MPa, 3405 * MPa, 3308 * MPa)
        angles = [45, -45, 0, 0, -45, 45]
        pt = [1 * mm] * 6
        fixes = [([0, 0, 0, 0, -np.inf, np.inf], [1, 2, 3]), ([ax, ax, 0, 0, -np.inf, np.inf], [2]),
                 ("boundary", [3])]
        edges = [([0, 0, -np.inf, np.inf, 0, 0], [-200 * 1e3, 0, 0, 0, 0, 0]),
                 ([ax, ax, -np.inf, np.inf, 0, 0], [200 * 1e3, 0, 0, 0, 0, 0])]
        quantity = (0, +1, 1 / mm)
    elif key in ("c6", "c7"):
        ax = ay = 1000 * mm
        nx = ny = 50
        th = 10 * mm
        angles = [0, 90] if key == "c6" else [-45, 45]
        pt = [th / 2] * 2
        fixes = [("boundary", [1, 2, 3])]
        pressure = [0, 0, -1.0 * MPa, 0, 0, 0]
        quantity = (2, -1, 1 / mm)
    elif key == "c8":
        th = 10 * mm
        angles = [90, 0]
        pt = [th / 2] * 2
        fixes = [([ax, ax, 0, 0, -np.inf, np.inf], [1, 2, 3]), ([0, 0, 0, 0, -np.inf, np.inf], [2]),
                 ("boundary", [3])]
        edges = [([0, 0, -np.inf, np.inf, 0, 0], [0.1 * MPa * sign * th, 0, 0, 0, 0, 0]),
                 ([ax, ax, -np.inf, np.inf, 0, 0], [-0.1 * MPa * sign * th, 0, 0, 0, 0, 0])]
        quantity = (2, sign, 1 / mm)
    elif key == "c9":
        th = 10 * mm
        angles = [90, 0]
        pt = [th / 2] * 2
        csys = np.array([[0.0, 0, 1], [1, 0, 0], [0, 1, 0]])
        fixes = [([0, 0, 0, 0, 0, 0], [1, 2, 3]), ([0, 0, 0, 0, ax, ax], [2]), ("boundary", [1])]
        edges = [([0, 0, -np.inf, np.inf, ax, ax], [0, 0, 0.1 * MPa * sign * th, 0, 0, 0]),
                 ([0, 0, -np.inf, np.inf, 0, 0], [0, 0, -0.1 * MPa * sign * th, 0, 0, 0])]
        quantity = (0, sign, 1 / mm)
    elif key == "c10":
        a = 250 * mm
        ax = ay = a
        E1 = 100e9
        E2 = E1 / 25
        lam = (E1, E2, 0.25, E2 * 0.5, E2 * 0.5, E2 * 0.2)
        th = a / 10
        angles = [0, 90]
        pt = [th / 2] * 2
        q = 10.0e3
        fixes = [([0, 0, -np.inf, np.inf, 0, 0], [2, 4]), ([a, a, -np.inf, np.inf, 0, 0], [2, 4]),
                 ([-np.inf, np.inf, 0, 0, 0, 0], [1, 5]), ([-np.inf, np.inf, a, a, 0, 0], [1, 5]),
                 ("boundary", [3])]
        pressure = [0, 0, q, 0, 0, 0]
        norm = (100 * E2) / (q * th * (a / th) ** 4)
        quantity = (2, +1, norm) if case == "c10w" else (0, +1, norm)
    xy, conn = (fe.T3block if nen == 3 else fe.Q4block)(ax, ay, nx, ny)
    xyz = fe.xyz3(xy)
    if key == "c9":
        xyz = xyz[:, [2, 0, 1]]  # (x, y, z) -> (z, x, y)
    tol = ax / nx / 100
    d = fe.DofField(xyz.shape[0])
    for box, comps in fixes:
        nodes = fe.boundary_nodes(conn) if isinstance(box, str) else fe.selectnode(xyz, box, tol)
        d.setebc(nodes, comps)
    d.numberdofs()
    F = np.zeros(d.nalldofs)
    bed = fe.boundary_edges(conn)
    for box, f in edges:
        F += fe.edge_load(xyz, meshio.select_edges(xyz, bed, box, tol), f, d.nalldofs, d.dofnums)
    D6 = O.lamina_D6(*lam)
    plies = [(D6, t, float(ang), 0.0) for t, ang in zip(pt, angles)]
    return dict(nen=nen, xyz=xyz, conn=conn, dchi=d, plies=plies, csys=csys, th=float(sum(pt)), F=F,
                pressure=pressure, load_npts=1 if nen == 3 else 4, quantity=quantity, lamina=lam,
                ref=(COMP_T3 if nen == 3 else COMP_Q4)[case], rtol=1e-9 if nen == 3 else 1.5e-8)


# test/test_shell_statics.jl:690-728: minimum and maximum of the nodal
# resultant fields of the barrel vault (fieldfromintegpoints, cylindrical
# output csys), rtol 0.01
BARREL_RESULTANTS = {
    "moment": [(-1520.6449167366522, 14.067403309095397), (-73.8262145426215, 425.93651541819503),
               (-0.005341121492547284, 953.0929383629322)],
    "membrane": [(-306.83173146926623, 309.5860993742647), (-1011.4832977998705, 2167.0403478574167),
                 (-687.3500043290137, 69.38703021678862)],
    "shear": [(-13.784764125688811, 21.165312065421237), (-7.4216963152916255, 25.679801383392967)],
}


def barrel_vault():
    """Barrel vault on the irregular Abaqus STRI3 mesh (test_shell_statics.jl:
    578-670): compacted, merged (thickness / 10), T3FF with stab 0.2 h^2 and
    drilling scale 0.1, load -0.625 along x per area (TriRule(3)), and the
    cylindrical output csys of the resultants (:676-688)."""
    from flexshell import meshio
    m = meshio.import_abaqus(os.path.join(GOLDEN, "inp", "barrelvault_stri3_irreg.inp"))
    xyz, conn, _ = meshio.compact_nodes(m["xyz"], m["fesets"][0][1])
    t = 3.0
    xyz, conn = meshio.merge_nodes(xyz, conn, t / 10)
    L = 50.0 * 12
    tol = t / 20
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [-np.inf, np.inf, -np.inf, np.inf, 0, 0], tol), [1, 2])
    d.setebc(fe.selectnode(xyz, [-np.inf, np.inf, -np.inf, np.inf, L / 2, L / 2], tol), [3, 4, 5])
    d.setebc(fe.selectnode(xyz, [-np.inf, np.inf, 0, 0, -np.inf, np.inf], tol), [2, 4, 6])
    d.numberdofs()

    def cylindrical(X, J=None, e=None, q=None):
        r = -np.asarray(X, dtype=float).copy()
        r[2] = 0.0
        c3 = r / np.linalg.norm(r)
        c2 = np.array([0.0, 0.0, 1.0])
        return np.stack([np.cross(c2, c3), c2, c3], axis=1)

    return dict(nen=3, xyz=xyz, conn=conn, dchi=d, E=3.0e6, nu=0.0, t=t, stab_c=0.2, drill=0.1,
                force=[-0.625, 0, 0, 0, 0, 0], csys=cylindrical)


def inverse_distance_field(xyz, conn, loc, vals):
    """Nodal field by inverse squared point-node distance weights (FinEtools
    fieldfromintegpoints, interpolation :invdistance) from point values
    (E x npts) at point locations (E x npts x 3); test-side restatement."""
    X = np.asarray(xyz)[conn]
    N = np.asarray(xyz).shape[0]
    num = np.zeros(N)
    den = np.zeros(N)
    for e in range(conn.shape[0]):
        for q in range(vals.shape[1]):
            d2 = np.sum((X[e] - loc[e, q]) ** 2, axis=1)
            d2 = np.maximum(d2, d2[d2 > 0].min() / 1e9)
            for a in range(conn.shape[1]):
                num[conn[e, a]] += vals[e, q] / d2[a]
                den[conn[e, a]] += 1.0 / d2[a]
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)


# test/test_shell_statics.jl:364-381 (T3FF) and test_q4rs_shell_statics.jl:
# 360-377 (Q4RS): twisted beam tip deflection in percent of the reference
# solution, for (thick, dir 2), (thick, dir 3), (thin, dir 2), (thin, dir 3)
# and meshes 2n x n, n = 2, 4, 8, 16 (rtol 1e-3 there)
TWISTED_T3 = [39.709921740907355, 68.87306876326497, 86.01944734315117, 95.04101960524827,
              53.10262177376127, 83.8593790803426, 94.91359387874728, 98.21549248655576,
              48.16757753755567, 79.43420077873479, 92.54464819755955, 96.85008269135751,
              50.577029703967334, 80.34160167730624, 92.48675665271801, 96.7096641005938]
TWISTED_Q4 = [57.52004303100694, 82.57318302555431, 94.20131995162048, 98.37454367262973,
              76.28227868667001, 93.90901921611645, 98.28145099724978, 99.43562910255783,
              68.51008267652144, 90.66189189880987, 97.4223162100464, 99.27913153086936,
              76.09971224380796, 93.00031411228612, 97.58071557550575, 99.2319932521005]
_TWISTED_PARAMS = [dict(t=0.32, force=1.0, dir=2, uex=0.001753248285256),
                   dict(t=0.32, force=1.0, dir=3, uex=0.005424534868469),
                   dict(t=0.0032, force=1.0e-6, dir=2, uex=0.001294),
                   dict(t=0.0032, force=1.0e-6, dir=3, uex=0.005256)]


def twisted_beam(nen, which, n):
    """Twisted cantilever (test_shell_statics.jl:289-348,
    test_q4rs_shell_statics.jl:283-342): 12 x 1.1, twisted by 90 degrees
    along its length, clamped at x = 0, a point force at the tip centre.
    which: 0..3 as in _TWISTED_PARAMS; mesh 2n x n."""
    p = _TWISTED_PARAMS[which]
    E, nu, W, L = 0.29e8, 0.22, 1.1, 12.0
    nL, nW = 2 * n, n
    tol = W / nW / 100
    xy, conn = (fe.T3block if nen == 3 else fe.Q4block)(L, W, nL, nW)
    xyz = fe.xyz3(xy)
    a = xyz[:, 0] / L * (np.pi / 2)
    y = xyz[:, 1] - W / 2
    z = xyz[:, 2]
    xyz = np.stack([xyz[:, 0], y * np.cos(a) - z * np.sin(a), y * np.sin(a) + z * np.cos(a)], axis=1)
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], tol), [1, 2, 3, 4, 5, 6])
    d.numberdofs()
    tip = fe.selectnode(xyz, [L, L, 0, 0, 0, 0], tol)
    F = np.zeros(d.nalldofs)
    F[d.dofnums[tip[0], p["dir"] - 1]] = p["force"]
    ref = (TWISTED_T3 if nen == 3 else TWISTED_Q4)[4 * which + [2, 4, 8, 16].index(n)]
    return dict(nen=nen, xyz=xyz, conn=conn, dchi=d, E=E, nu=nu, rho=0.0, t=p["t"],
                stab_c=(None if nen == 3 else 0.05), F=F, probe=tip[0], dir=p["dir"], uex=p["uex"], ref=ref)


# test/test_shell_statics.jl:521-531: LE5 Z-section cantilever (NAFEMS), min
# and max of the deflection w (default rtol of isapprox)
LE5_W = (-0.01568028415401719, 0.015401663810490641)


def le5():
    """LE5 Z-section cantilever under torsion (test_shell_statics.jl:438-519)
    on the reference's Abaqus mesh nle5xf3c.inp: T3FF, E 210e9, nu 0.3,
    t 0.1, stab 0.2 h^2, translations fixed at x = 0, point forces +-0.6e6
    along z at (10, +-1, 0)."""
    from flexshell import meshio
    m = meshio.import_abaqus(os.path.join(GOLDEN, "inp", "nle5xf3c.inp"))
    xyz, conn, _ = meshio.compact_nodes(m["xyz"], m["fesets"][0][1])
    t = 0.1
    tol = t / 1000
    d = fe.DofField(xyz.shape[0])
    d.setebc(fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], tol), [1, 2, 3])
    d.numberdofs()
    F = np.zeros(d.nalldofs)
    for y, fz in ((1.0, 0.6e6), (-1.0, -0.6e6)):
        nl = fe.selectnode(xyz, [10.0, 10.0, y, y, 0, 0], tol)
        F[d.dofnums[nl[0], 2]] += fz
    return dict(nen=3, xyz=xyz, conn=conn, dchi=d, E=210e9, nu=0.3, rho=0.0, t=t, stab_c=0.2, F=F)


# test/test_composite_shell_dynamics.jl:93-104 (and the same list for the
# layered model): 8 lowest frequencies (Hz) of a simply supported square
# plate, isotropic, homogeneous T3FF and [45/-45/45/-45] T3FFComp
PETYT_FS = [21.822774909379287, 54.52203720717488, 54.60475772077202, 86.5480437899711, 108.65349048567792,
            108.78050954575491, 138.29506871044424, 140.89370172063866]


def petyt(composite):
    """mcompshelldyn0 (test_composite_shell_dynamics.jl:14-92 and 106-190):
    1 m square plate, t = 1 mm, E 1000 MPa, nu 0.396, rho 2, 8 x 8 T3 mesh;
    boundary translations fixed (the layered model also pins a corner)."""
    ax = ay = 1.0
    nx = ny = 8
    tol = ax / nx / 100
    xy, conn = fe.T3block(ax, ay, nx, ny)
    xyz = fe.xyz3(xy)
    d = fe.DofField(xyz.shape[0])
    if composite:
        d.setebc(fe.selectnode(xyz, [ax, ax, 0, 0, -np.inf, np.inf], tol), [1, 2, 3])
        d.setebc(fe.selectnode(xyz, [0, 0, 0, 0, -np.inf, np.inf], tol), [2])
    d.setebc(fe.boundary_nodes(conn), [1, 2, 3])
    d.numberdofs()
    t = ax / 1000
    D6 = O.iso_D6(1000e6, 0.396)
    plies = [(D6, t / 4, a, 2.0) for a in (45.0, -45.0, 45.0, -45.0)]
    return dict(nen=3, xyz=xyz, conn=conn, dchi=d, E=1000e6, nu=0.396, rho=2.0, t=t, stab_c=None, plies=plies,
                csys=np.eye(3), th=t)
