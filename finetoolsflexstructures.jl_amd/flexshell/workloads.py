"""Seeded synthetic models of the benchmark configurations.

plate_c1          flat square plate 1 x 1 m, t 0.01, steel, T3 32 x 32 (2048
                  triangles), all edges clamped.
cylinder_panel_c2 cylindrical panel R 1 m, L 2 m, 60 degree arc, t 0.005,
                  E 70e9, nu 0.33, rho 2700, Q4 n x n (512 x 512 = 262,144
                  elements), interior nodes jittered by U(+-10% h) in the
                  surface parameters (seed 2), straight edges simply supported
                  (u = v = w = 0, rotations free).
plate_c4          flat plate 2 x 1 m, t 0.004, steel, T3 1024 x 512
                  (1,048,576 triangles), all edges clamped, Gaussian pressure
                  pulse (sigma 0.1 m, 1e4 Pa peak, centre seeded, 1e-4 s).
"""
import numpy as np

from . import fe
from .material import MatDeforElastIso

STEEL = dict(E=210e9, nu=0.3, rho=7850.0)


def _edges_fixed(dchi, nodes, comps):
    dchi.setebc(nodes, comps)
    return dchi


def plate_c1(n=32, L=1.0, t=0.01):
    xy, conn = fe.T3block(L, L, n, n)
    xyz = fe.xyz3(xy)
    d = fe.DofField(xyz.shape[0])
    tol = L / n / 100
    bn = np.unique(np.concatenate([fe.selectnode(xyz, [0, 0, -np.inf, np.inf, -np.inf, np.inf], tol),
                                   fe.selectnode(xyz, [L, L, -np.inf, np.inf, -np.inf, np.inf], tol),
                                   fe.selectnode(xyz, [-np.inf, np.inf, 0, 0, -np.inf, np.inf], tol),
                                   fe.selectnode(xyz, [-np.inf, np.inf, L, L, -np.inf, np.inf], tol)]))
    d.setebc(bn, [1, 2, 3, 4, 5, 6]).numberdofs()
    mat = MatDeforElastIso(STEEL["E"], STEEL["nu"], STEEL["rho"])
    return dict(nen=3, xyz=xyz, conn=conn, dchi=d, material=mat, t=t, name="plate_t3_32x32")


def cylinder_panel_c2(n=512, seed=2, R=1.0, L=2.0, arc_deg=60.0, t=0.005, E=70e9, nu=0.33, rho=2700.0,
                      jitter=0.1):
    arc = arc_deg / 180 * np.pi
    ab, conn = fe.Q4block(arc, L, n, n)
    rng = np.random.default_rng(seed)
    ha, hy = arc / n, L / n
    a, y = ab[:, 0].copy(), ab[:, 1].copy()
    tol = 1e-9
    interior = (a > tol) & (a < arc - tol) & (y > tol) & (y < L - tol)
    a[interior] += rng.uniform(-jitter * ha, jitter * ha, interior.sum())
    y[interior] += rng.uniform(-jitter * hy, jitter * hy, interior.sum())
    xyz = np.stack([R * np.sin(a), y, R * np.cos(a)], axis=1)
    d = fe.DofField(xyz.shape[0])
    straight = np.nonzero((ab[:, 0] < tol) | (ab[:, 0] > arc - tol))[0]
    d.setebc(straight, [1, 2, 3]).numberdofs()
    mat = MatDeforElastIso(E, nu, rho)
    return dict(nen=4, xyz=xyz, conn=conn, dchi=d, material=mat, t=t, name=f"cylinder_panel_q4_{n}x{n}")


def plate_c4(nx=1024, ny=512, Lx=2.0, Ly=1.0, t=0.004, seed=5):
    xy, conn = fe.T3block(Lx, Ly, nx, ny)
    xyz = fe.xyz3(xy)
    d = fe.DofField(xyz.shape[0])
    tol = min(Lx / nx, Ly / ny) / 100
    bn = np.nonzero((xyz[:, 0] < tol) | (xyz[:, 0] > Lx - tol) | (xyz[:, 1] < tol) | (xyz[:, 1] > Ly - tol))[0]
    d.setebc(bn, [1, 2, 3, 4, 5, 6]).numberdofs()
    rng = np.random.default_rng(seed)
    x0 = np.array([rng.uniform(0.25 * Lx, 0.75 * Lx), rng.uniform(0.25 * Ly, 0.75 * Ly), 0.0])
    mat = MatDeforElastIso(STEEL["E"], STEEL["nu"], STEEL["rho"])
    return dict(nen=3, xyz=xyz, conn=conn, dchi=d, material=mat, t=t, x0=x0, name=f"plate_t3_{nx}x{ny}")


def pressure_load_t3(xyz, conn, dchi, pressure_at):
    """Consistent nodal loads of a normal pressure p(x) on a flat T3 mesh in the
    xy plane, integrated with the 3-point rule (TriRule(3)); vectorized form
    of distribloads for this case.  Returns the free-dof vector."""
    X = xyz[conn]                       # E x 3 x 3
    area = 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)
    pts = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])  # N at the 3 points
    F = np.zeros(dchi.nalldofs)
    fe_ = np.zeros((conn.shape[0], 3))
    for q in range(3):
        xq = np.einsum("a,eak->ek", pts[q], X)
        p = pressure_at(xq)
        fe_ += np.outer(p * area / 3, pts[q])
    np.add.at(F, dchi.dofnums[conn, 2].ravel(), fe_.ravel())
    return F[:dchi.nfreedofs]


def gaussian_velocity(dchi, sigma, seed):
    """Seeded N(0, sigma) velocities on the translational free dofs."""
    rng = np.random.default_rng(seed)
    V = np.zeros(dchi.nfreedofs)
    trans = dchi.dofnums[:, :3][~dchi.is_fixed[:, :3]]
    V[trans] = rng.normal(0.0, sigma, trans.size)
    return V


def converged_dt(femm, nfree, seed=0, maxit=300, factor=0.9):
    """Stable time step of a workload: factor * 2 / omega_max with omega_max
    from the device power iteration (fs_model_omega_max) started from a
    seeded Gaussian vector and run for maxit iterations.  The reference's
    start (rand(n), all positive, 30 iterations,
    clamp_cyl_transient_examples.jl:45-60) has little weight in the
    oscillatory top modes: on one model of the 256-model batch it
    underestimated omega_max by 11 %, so that 0.9 of its dt was unstable
    and the trajectory grew to 1e62 in 500 steps.  A Gaussian start reached
    omega_max within 1e-4 in 300 iterations on every batch model, and within
    0.4 % (0.06 %) on the 512 x 512 cylinder (1024 x 512 plate)."""
    from . import explicit
    v0 = np.random.default_rng(seed).standard_normal(nfree)
    return explicit.stable_dt(explicit.omega_max(femm, v0, maxit=maxit), factor=factor)


def batch_c3(nmodels=256, seed=4, nmin=8, nmax=128):
    """256 independent T3 plate models: n per side U{nmin..nmax} (2 n^2
    triangles), aspect ratio U(1, 4), thickness log-U(0.002, 0.05), E
    log-U(1e9, 3e11), nu U(0.2, 0.4), rho U(1000, 8000), interior node jitter
    +-10% h, flat or cylindrically curved (seeded), each edge clamped or
    simply supported (u = v = w = 0) by seed."""
    rng = np.random.default_rng(seed)
    models = []
    for k in range(nmodels):
        n = int(rng.integers(nmin, nmax + 1))
        aspect = rng.uniform(1.0, 4.0)
        Ly = 1.0
        Lx = aspect * Ly
        t = float(np.exp(rng.uniform(np.log(0.002), np.log(0.05))))
        E = float(np.exp(rng.uniform(np.log(1e9), np.log(3e11))))
        nu = float(rng.uniform(0.2, 0.4))
        rho = float(rng.uniform(1000.0, 8000.0))
        curved = bool(rng.random() < 0.5)
        R = float(rng.uniform(2.0, 10.0) * Lx)
        edges = rng.integers(0, 2, 4)  # 0 clamped, 1 simply supported: x=0, x=Lx, y=0, y=Ly
        xy, conn = fe.T3block(Lx, Ly, n, n)
        hx, hy = Lx / n, Ly / n
        tol = 1e-9
        interior = (xy[:, 0] > tol) & (xy[:, 0] < Lx - tol) & (xy[:, 1] > tol) & (xy[:, 1] < Ly - tol)
        xy = xy.copy()
        xy[interior, 0] += rng.uniform(-0.1 * hx, 0.1 * hx, interior.sum())
        xy[interior, 1] += rng.uniform(-0.1 * hy, 0.1 * hy, interior.sum())
        if curved:
            a = (xy[:, 0] - Lx / 2) / R
            xyz = np.stack([R * np.sin(a), xy[:, 1], R * (np.cos(a) - 1)], axis=1)
        else:
            xyz = fe.xyz3(xy)
        d = fe.DofField(xyz.shape[0])
        sel = [xy[:, 0] < tol, xy[:, 0] > Lx - tol, xy[:, 1] < tol, xy[:, 1] > Ly - tol]
        for s_, kind in zip(sel, edges):
            d.setebc(np.nonzero(s_)[0], [1, 2, 3, 4, 5, 6] if kind == 0 else [1, 2, 3])
        d.numberdofs()
        models.append(dict(xyz=xyz, conn=conn, dchi=d, material=MatDeforElastIso(E, nu, rho), t=t,
                           name=f"plate{k}_n{n}"))
    return models


C5_LAMINA = dict(E1=140e9, E2=10e9, G12=5e9, nu12=0.3, rho=1600.0)
C5_KINDS = ("assembly", "modal", "explicit")


def c5_layup(nplies, angles, ply_t=0.125e-3):
    """Layered plate of orthotropic plies (E1 140e9, E2 10e9, G12 5e9, nu12
    0.3, G13 = G12, G23 = E2 / (2 (1 + 0.3)), rho 1600) in the global
    cartesian frame (csys = I, so the layup normals are the true normals of
    a flat plate in the xy plane)."""
    from .material import CompositeLayup, Ply, lamina_material
    L = C5_LAMINA
    mat = lamina_material(L["rho"], L["E1"], L["E2"], L["nu12"], L["G12"], L["G12"], L["E2"] / (2 * 1.3))
    plies = [Ply(f"ply{k}", mat, ply_t, float(a)) for k, a in enumerate(angles[:nplies])]
    return CompositeLayup("c5", plies, np.eye(3))


def served_c5(nreq=1000, seed=6, rate=20.0, emin=1e3, emax=5e5, explicit_steps=500):
    """Request specs of the served layered-shell workload: Poisson arrivals at
    `rate` requests/s (seed), kind uniform over assembly / assembly + modal
    matrices / explicit dynamics, element count log-uniform in [emin, emax],
    T3FFComp or Q4RSComp, 3..8 plies with angles uniform in {0, 45, -45, 90}
    and ply thickness 0.125 mm; flat plates Ly = 0.5 m, aspect U(1, 3),
    interior jitter +-10% h, edges clamped or simply supported (seeded).
    Explicit requests start from N(0, 1e-3 m/s) translational velocities and
    run `explicit_steps` steps at 0.9 of the stable dt."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, nreq)
    specs = []
    t = 0.0
    for k in range(nreq):
        t += gaps[k]
        nel = float(np.exp(rng.uniform(np.log(emin), np.log(emax))))
        nen = int(rng.choice([3, 4]))
        aspect = float(rng.uniform(1.0, 3.0))
        nplies = int(rng.integers(3, 9))
        angles = [float(a) for a in rng.choice([0.0, 45.0, -45.0, 90.0], 8)]
        kind = C5_KINDS[int(rng.integers(0, 3))]
        specs.append(dict(rid=k, arrival=t, kind=kind, nen=nen, nel=nel, aspect=aspect, nplies=nplies,
                          angles=angles, edges=[int(e) for e in rng.integers(0, 2, 4)],
                          mesh_seed=int(rng.integers(0, 2 ** 31)), steps=explicit_steps if kind == "explicit" else 0))
    return specs


def c5_model(spec, Ly=0.5):
    """Mesh, dof field and layup of one served request spec."""
    nen, aspect = spec["nen"], spec["aspect"]
    cells = spec["nel"] / (2.0 if nen == 3 else 1.0)
    ny = max(2, int(round(np.sqrt(cells / aspect))))
    nx = max(2, int(round(cells / ny)))
    Lx = aspect * Ly
    xy, conn = (fe.T3block if nen == 3 else fe.Q4block)(Lx, Ly, nx, ny)
    rng = np.random.default_rng(spec["mesh_seed"])
    hx, hy = Lx / nx, Ly / ny
    tol = 1e-9
    interior = (xy[:, 0] > tol) & (xy[:, 0] < Lx - tol) & (xy[:, 1] > tol) & (xy[:, 1] < Ly - tol)
    xy = xy.copy()
    xy[interior, 0] += rng.uniform(-0.1 * hx, 0.1 * hx, interior.sum())
    xy[interior, 1] += rng.uniform(-0.1 * hy, 0.1 * hy, interior.sum())
    xyz = fe.xyz3(xy)
    d = fe.DofField(xyz.shape[0])
    sel = [xy[:, 0] < tol, xy[:, 0] > Lx - tol, xy[:, 1] < tol, xy[:, 1] > Ly - tol]
    for s_, kind in zip(sel, spec["edges"]):
        d.setebc(np.nonzero(s_)[0], [1, 2, 3, 4, 5, 6] if kind == 0 else [1, 2, 3])
    d.numberdofs()
    layup = c5_layup(spec["nplies"], spec["angles"])
    V0 = gaussian_velocity(d, 1e-3, spec["mesh_seed"] + 1) if spec["kind"] == "explicit" else None
    return dict(nen=nen, xyz=xyz, conn=conn, dchi=d, layup=layup, V0=V0)
