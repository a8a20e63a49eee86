"""Materials and laminates on the host (small per-model precompute).

MatDeforElastIso / MatDeforElastOrtho 3-D stiffness (FinEtoolsDeforLinear,
ordering 11, 22, 33, 12, 13, 23), the plane-stress / transverse-shear
condensation of the shell FEMMs (FEMMShellT3FFModule.jl:323-342), and the
CompositeLayupModule (CompositeLayupModule.jl) laminate quantities that the
layered FEMMs pass to the kernels.
"""
import numpy as np


class MatDeforElastIso:
    def __init__(self, E, nu, rho=0.0):
        self.E, self.nu, self.rho = float(E), float(nu), float(rho)

    def D(self):
        E, nu = self.E, self.nu
        lam = E * nu / (1 + nu) / (1 - 2 * nu)
        mu = E / 2 / (1 + nu)
        D = np.zeros((6, 6))
        D[:3, :3] = lam
        for i in range(3):
            D[i, i] = lam + 2 * mu
        for i in range(3, 6):
            D[i, i] = mu
        return D


class MatDeforElastOrtho:
    def __init__(self, E1, E2, E3, nu12, nu13, nu23, G12, G13, G23, rho=0.0):
        self.p = (E1, E2, E3, nu12, nu13, nu23, G12, G13, G23)
        self.rho = float(rho)

    def D(self):
        E1, E2, E3, nu12, nu13, nu23, G12, G13, G23 = self.p
        C = np.array([
            [1.0 / E1, -nu12 / E1, -nu13 / E1, 0, 0, 0],
            [-nu12 / E1, 1.0 / E2, -nu23 / E2, 0, 0, 0],
            [-nu13 / E1, -nu23 / E2, 1.0 / E3, 0, 0, 0],
            [0, 0, 0, 1 / G12, 0, 0],
            [0, 0, 0, 0, 1 / G13, 0],
            [0, 0, 0, 0, 0, 1 / G23],
        ])
        return np.linalg.inv(C)


def shell_material_stiffness(material):
    """(Dps, Dt) of _shell_material_stiffness (FEMMShellT3FFModule.jl:323-342)."""
    D = material.D()
    Dps = np.zeros((3, 3))
    Dps[:2, :2] = D[:2, :2] - np.outer(D[:2, 2], D[2, :2]) / D[2, 2]
    ix = [0, 1, 3]
    for i in range(3):
        Dps[2, i] = Dps[i, 2] = D[3, ix[i]]
    Dt = np.diag([D[4, 4], D[5, 5]])
    return Dps, Dt


def lamina_material(*args):
    """CompositeLayupModule.lamina_material (:105-167): transversely isotropic
    lamina with E3 = E2, nu13 = nu12, nu23 = 0 (or isotropic for (E, nu))."""
    if len(args) == 6:
        E1, E2, nu12, G12, G13, G23 = args
        rho = 0.0
    elif len(args) == 7:
        rho, E1, E2, nu12, G12, G13, G23 = args
    elif len(args) == 2:
        return MatDeforElastIso(args[0], args[1], 0.0)
    elif len(args) == 3:
        return MatDeforElastIso(args[1], args[2], args[0])
    else:
        raise TypeError("lamina_material(E1, E2, nu12, G12, G13, G23[, rho first])")
    return MatDeforElastOrtho(E1, E2, E2, nu12, nu12, 0.0, G12, G13, G23, rho)


class Ply:
    """CompositeLayupModule.Ply (:71-97): ply stiffness in the ply frame."""

    def __init__(self, name, material, thickness, angle):
        self.name = name
        self.material = material
        self.thickness = float(thickness)
        self.angle = float(angle)
        Dps, Dts = shell_material_stiffness(material)
        self._Dps = Dps
        self._Dts = Dts


def cartesian_csys(axes):
    """3 x 3 matrix of cartesian_csys (:26-44)."""
    M = np.zeros((3, 3))
    for j in range(3):
        M[abs(axes[j]) - 1, j] = np.sign(axes[j]) * 1.0
    return M


def _tinv(m, n):
    return np.array([[m * m, n * n, 2 * (m * n)], [n * n, m * m, -2 * (m * n)], [-(m * n), (m * n), m * m - n * n]])


class CompositeLayup:
    """CompositeLayupModule.CompositeLayup (:185-225).  csys: 3 x 3 matrix
    (constant layup coordinate system); tsc: None for the Vinson-Sierakowski
    transverse shear model, else the constant shear correction factor."""

    def __init__(self, name, plies, csys, tsc=None, offset=0.0):
        self.name = name
        self.plies = list(plies)
        self.csys = np.asarray(csys, dtype=float)
        self.tsc = tsc
        self.offset = float(offset)

    def thickness(self):
        return sum(p.thickness for p in self.plies)

    def laminate_stiffnesses(self):
        """A, B, D (:246-271)."""
        A = np.zeros((3, 3))
        B = np.zeros((3, 3))
        D = np.zeros((3, 3))
        tl = self.thickness()
        zs = -tl / 2 - self.offset
        for p in self.plies:
            ze = zs + p.thickness
            ang = p.angle / 180 * np.pi
            T = _tinv(np.cos(-ang), np.sin(-ang))
            Dps = T @ (p._Dps @ T.T)
            A += (ze - zs) * Dps
            B += (ze ** 2 - zs ** 2) / 2 * Dps
            D += (ze ** 3 - zs ** 3) / 3 * Dps
            zs += p.thickness
        return A, B, D

    def laminate_transverse_stiffness(self):
        """H (:280-308)."""
        H = np.zeros((2, 2))
        tl = self.thickness()
        zs = -tl / 2 - self.offset
        for p in self.plies:
            ze = zs + p.thickness
            ang = p.angle / 180 * np.pi
            m, n = np.cos(ang), np.sin(ang)
            T = np.array([[m, -n], [n, m]])
            Dts = T.T @ (p._Dts @ T)
            if self.tsc is None:
                H += 5 / 4 * (ze - zs - 4 / 3 * (ze ** 3 - zs ** 3) / tl ** 2) * Dts
            else:
                H += self.tsc * (ze - zs) * Dts
            zs += p.thickness
        return H

    def laminate_inertia(self):
        """(mass density, moment of inertia density) (:317-330)."""
        tl = self.thickness()
        zs = -tl / 2 - self.offset
        md = 0.0
        mi = 0.0
        for p in self.plies:
            ze = zs + p.thickness
            md += (ze - zs) * p.material.rho
            mi += (ze ** 3 - zs ** 3) * p.material.rho / 3
            zs += p.thickness
        return md, mi
