"""Shell FEMMs and assemblers: the drop-in surface of the reference's API.

Mirrors FEMMShellT3FFModule, FEMMShellQ4RSModule, FEMMShellT3FFCompModule and
FEMMShellQ4RSCompModule (make, associategeometry!, stiffness, mass) and the
assembler protocol of AssemblyModule.jl / FinEtools (SysmatAssemblerSparseSymm,
SysmatAssemblerSparseCSRSymm, SysmatAssemblerFFBlock, SysmatAssemblerSparseDiag).
Every numeric result comes from libflexshell on the GPU; this module only
marshals arrays across the C ABI.

Julia's `!` suffix is dropped (associategeometry! -> associategeometry).  The
stabilisation closure stab_fun(t, h) = t^2 / (t^2 + c h^2) is given by its
coefficient c (`stab_c`), which covers every form the reference uses.
"""
import ctypes as C
import warnings

import numpy as np
import scipy.sparse as sp

from . import _lib
from .material import shell_material_stiffness

__TRANSV_SHEAR_FORMULATION_AVERAGE_B = 0
__TRANSV_SHEAR_FORMULATION_AVERAGE_K = 1
AVERAGE_B = 0
AVERAGE_K = 1


# ---------------------------------------------------------------------------
# assemblers
# ---------------------------------------------------------------------------

class SysmatAssemblerSparseSymm:
    """All-dof sparse result (FinEtools default assembler); returned as CSR."""
    block = _lib.FS_BLOCK_ALL
    diag_only = False


class SysmatAssemblerSparseCSRSymm(SysmatAssemblerSparseSymm):
    """AssemblyModule.jl:20-53: all dofs, CSR result."""


class SysmatAssemblerSparseDiag:
    """Diagonal (lumped) result over all dofs."""
    block = _lib.FS_BLOCK_ALL
    diag_only = True


class SysmatAssemblerFFBlock:
    """Free-free block (FinEtools SysmatAssemblerFFBlock(nfreedofs) or
    SysmatAssemblerFFBlock(SysmatAssemblerSparseDiag(), nf, nf))."""
    block = _lib.FS_BLOCK_FF

    def __init__(self, *args):
        self.diag_only = any(isinstance(a, SysmatAssemblerSparseDiag) for a in args)


class SysmatAssemblerHIP:
    """Keeps the result resident on the device; returns a DeviceMatrix."""

    def __init__(self, block=_lib.FS_BLOCK_FF):
        self.block = block
        self.diag_only = False


class SysmatAssemblerHIPGeneric:
    """The assembler protocol of FinEtools' SysmatAssemblerSparse (extended to
    CSR by SysmatAssemblerSparseCSRSymm, AssemblyModule.jl:28-52) for FEMMs
    whose element matrices are computed on the host: startassembly sizes the
    buffer, assemble appends one element matrix with its row and column dof
    numbers (0-based), makematrix reduces the buffer on the device
    (fs_coo_values) into a CSR matrix with the pattern of sparse(I, J, V)
    (zeros retained) and duplicates summed in assemble order.  Dof numbers
    at or beyond nrows / ncols are dropped (the free-free block when they are
    the numbers of free dofs).  The reduction plan is kept while the dof
    numbers of a new assembly equal the previous ones."""

    def __init__(self, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self._h = None
        self._key = None

    def startassembly(self, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs):
        self.kr, self.kc = int(elem_mat_nrows), int(elem_mat_ncols)
        self.nrows, self.ncols = int(row_nalldofs), int(col_nalldofs)
        self._mats = np.zeros((int(n_elem_mats), self.kc, self.kr))  # column-major element matrices
        self._dr = np.zeros((int(n_elem_mats), self.kr), dtype=np.int64)
        self._dc = np.zeros((int(n_elem_mats), self.kc), dtype=np.int64)
        self._n = 0
        return self

    def assemble(self, mat, dofnums_row, dofnums_col):
        mat = np.asarray(mat, dtype=np.float64)
        if mat.shape != (self.kr, self.kc):
            raise ValueError(f"element matrix {mat.shape}, expected {(self.kr, self.kc)}")
        if self._n >= self._mats.shape[0]:
            raise ValueError("more element matrices than startassembly announced")
        self._mats[self._n] = mat.T
        self._dr[self._n] = np.asarray(dofnums_row).ravel()
        self._dc[self._n] = np.asarray(dofnums_col).ravel()
        self._n += 1
        return self

    def makematrix(self):
        n = self._n
        dr = np.ascontiguousarray(self._dr[:n])
        dc = np.ascontiguousarray(self._dc[:n])
        key = (self.nrows, self.ncols, self.kr, self.kc, n, dr.tobytes(), dc.tobytes())
        lib = self.ctx.lib
        if self._h is None or key != self._key:
            self.close()
            h = C.c_void_p()
            self.ctx.check(lib.fs_coo_create(self.ctx.h, self.nrows, self.ncols, self.kr, self.kc, n,
                                             _lib.ptr(dr, C.c_int64), _lib.ptr(dc, C.c_int64), 0, C.byref(h)),
                           "fs_coo_create")
            self._h, self._key = h, key
            nz = C.c_int64()
            self.ctx.check(lib.fs_coo_pattern(h, C.byref(nz)), "fs_coo_pattern")
            self._rp = np.zeros(self.nrows + 1, dtype=np.int64)
            self._cv = np.zeros(nz.value, dtype=np.int64)
            self.ctx.check(lib.fs_coo_pattern_get(h, _lib.ptr(self._rp, C.c_int64), _lib.ptr(self._cv, C.c_int64), 0),
                           "fs_coo_pattern_get")
        vals = np.zeros(len(self._cv))
        mats = np.ascontiguousarray(self._mats[:n])
        self.ctx.check(lib.fs_coo_values(self._h, _lib.ptr(mats), _lib.ptr(vals)), "fs_coo_values")
        return sp.csr_matrix((vals, self._cv.copy(), self._rp.copy()), shape=(self.nrows, self.ncols))

    def close(self):
        if self._h is not None:
            self.ctx.lib.fs_coo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMatrix:
    """Handle to a matrix resident in HBM (the CSR stiffness or the lumped
    mass of one block of a model)."""

    def __init__(self, femm, block, kind):
        self.femm = femm
        self.block = block
        self.kind = kind

    @property
    def model(self):
        return self.femm._model


# ---------------------------------------------------------------------------
# FEMMs
# ---------------------------------------------------------------------------

class _ShellFEMM:
    nen = None
    default_stab_c = None
    composite = False

    def __init__(self, conn, material, thickness=None, stab_c=None, mcsys=None, ctx=None):
        self.conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int64))
        if self.conn.shape[1] != self.nen:
            raise ValueError(f"{type(self).__name__} needs {self.nen}-node elements")
        self.material = material
        self.thickness = thickness
        self.stab_c = self.default_stab_c if stab_c is None else float(stab_c)
        self.mcsys = mcsys
        self.transv_shear_formulation = AVERAGE_B
        self.drilling_stiffness_scale = 1.0
        self.threshold_angle = 30.0
        self._ctx = ctx
        self._model = None
        self._geom_key = None
        self._dofs_key = None
        self._associatedgeometry = False
        self._normals = None
        self._normal_valid = None
        self._patterns = {}

    # -- device model management --
    def _context(self):
        if self._ctx is None:
            self._ctx = _lib.default_context()
        return self._ctx

    def _ensure_model(self, geom0):
        geom0 = np.ascontiguousarray(np.asarray(geom0, dtype=np.float64))
        key = (geom0.shape, hash(geom0.tobytes()))
        if self._model is None or key != self._geom_key:
            if self._model is not None:
                self._model.close()
            self._model = _lib.Model(self._context(), self.nen, geom0, self.conn)
            self._geom_key = key
            self._dofs_key = None
            self._patterns = {}
        return self._model

    def _ensure_dofs(self, dchi):
        key = (dchi.nalldofs, dchi.nfreedofs, hash(dchi.dofnums.tobytes()))
        if key != self._dofs_key:
            self._model.set_dofs(dchi.dofnums, dchi.nalldofs, dchi.nfreedofs)
            self._dofs_key = key
            self._patterns = {}
            if self._normals is not None:
                self._model.normals_upload(self._normals, self._normal_valid)

    def _pattern(self, block):
        if block not in self._patterns:
            self._patterns[block] = self._model.pattern(block)
        return self._patterns[block]

    # -- associategeometry! --
    def associategeometry(self, geom0):
        """Nodal normals (FEMMShellT3FFModule.jl:570-616, FEMMShellQ4RSModule.jl:472-525)."""
        m = self._ensure_model(geom0)
        csys_normal = self._csys_normal_function()
        if csys_normal is None:
            self._normals, self._normal_valid = m.normals_isoparametric(self.threshold_angle)
        else:
            self._normals, self._normal_valid = self._host_csys_normals(np.asarray(geom0, dtype=float), csys_normal)
            m.normals_upload(self._normals, self._normal_valid)
        self._associatedgeometry = True
        return self

    def _csys_normal_function(self):
        return None if self.mcsys is None else (lambda X, J, e, q: np.asarray(self.mcsys(X, J, e, q))[:, 2])

    def _host_csys_normals(self, xyz, normal_of):
        """Normals from a coordinate-system closure, which cannot cross the ABI
        (the `mcsys` / `layup.csys` branch of associategeometry!)."""
        conn = self.conn
        N = xyz.shape[0]
        nrm = np.zeros((N, 3))
        valid = np.ones(N, dtype=bool)
        s = np.sin(self.threshold_angle / 180 * np.pi)
        ntol = 1 - np.sqrt(1 - s * s)
        if self.nen == 3:
            for el, (i, j, k) in enumerate(conn):
                J0 = np.stack([xyz[j] - xyz[i], xyz[k] - xyz[i]], axis=1)
                for n in (i, j, k):
                    nrm[n] += normal_of(xyz[n], J0, el, 0)
            nn = np.linalg.norm(nrm, axis=1)
            nrm[nn > 0] /= nn[nn > 0, None]
            for el, (i, j, k) in enumerate(conn):
                J0 = np.stack([xyz[j] - xyz[i], xyz[k] - xyz[i]], axis=1)
                e3 = _element_normal(J0)
                for n in (i, j, k):
                    if nrm[n] @ e3 < 1 - ntol:
                        valid[n] = False
        else:
            npc = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
            for el, c in enumerate(conn):
                X = xyz[c]
                for jj in range(4):
                    J = _q4_jac(X, *npc[jj])
                    Jac = np.linalg.norm(np.cross(J[:, 0], J[:, 1]))
                    loc = _q4_N(*npc[jj]) @ X
                    nrm[c[jj]] += Jac * normal_of(loc, J, el, jj)
            nn = np.linalg.norm(nrm, axis=1)
            nrm[nn > 0] /= nn[nn > 0, None]
            for el, c in enumerate(conn):
                X = xyz[c]
                for jj in range(4):
                    J = _q4_jac(X, *npc[jj])
                    loc = _q4_N(*npc[jj]) @ X
                    if nrm[c[jj]] @ normal_of(loc, J, el, jj) < 1 - ntol:
                        valid[c[jj]] = False
        return nrm, valid

    # -- parameters --
    def _shell_params(self):
        p = _lib.ShellParams()
        p.composite = 0
        p.formulation = int(self.transv_shear_formulation)
        p.thickness = float(self.thickness)
        p.stab_c = float(self.stab_c)
        p.drilling_stiffness_scale = float(self.drilling_stiffness_scale)
        Dps, Dt = shell_material_stiffness(self.material)
        p.Dps[:] = list(np.asarray(Dps).T.ravel())
        p.Dt[:] = list(np.asarray(Dt).T.ravel())
        self._keep = ()
        return p

    def _mass_params(self):
        p = _lib.MassParams()
        p.composite = 0
        p.rho = float(self.material.rho)
        p.thickness = float(self.thickness)
        self._mkeep = ()
        return p

    # -- stiffness / mass --
    def stiffness(self, assembler, geom0, u1=None, Rfield1=None, dchi=None):
        """stiffness(femm, assembler, geom0, u1, Rfield1, dchi); u1 and Rfield1
        are ignored by the linear shells, as in the reference."""
        if dchi is None:  # stiffness(femm, geom0, u1, Rfield1, dchi) form
            assembler, geom0, dchi = SysmatAssemblerSparseSymm(), assembler, Rfield1
        assert self._associatedgeometry, "associategeometry must be called first"
        m = self._ensure_model(geom0)
        self._ensure_dofs(dchi)
        block = assembler.block
        if isinstance(assembler, SysmatAssemblerHIP):
            m.stiffness(self._shell_params(), block)
            return DeviceMatrix(self, block, "K")
        nrows, rp, cv = self._pattern(block)
        vals = m.stiffness(self._shell_params(), block, len(cv))
        return _csr(vals, cv, rp, nrows)

    def element_matrices(self, geom0, dchi=None):
        """Element matrices (nelem x k x k, global basis): the elmat that the
        reference passes to assemble!."""
        assert self._associatedgeometry, "associategeometry must be called first"
        m = self._ensure_model(geom0)
        blocks = m.element_blocks(self._shell_params())
        nen = self.nen
        K = np.zeros((blocks.shape[0], 6 * nen, 6 * nen))
        b = 0
        for a in range(nen):
            for c in range(a, nen):
                K[:, 6 * a:6 * a + 6, 6 * c:6 * c + 6] = blocks[:, b]
                if a != c:
                    K[:, 6 * c:6 * c + 6, 6 * a:6 * a + 6] = np.transpose(blocks[:, b], (0, 2, 1))
                b += 1
        return K

    def mass(self, assembler, geom0, dchi=None):
        """Lumped mass (mass(femm, assembler, geom0, dchi))."""
        if dchi is None:
            assembler, geom0, dchi = SysmatAssemblerSparseSymm(), assembler, geom0
        assert self._associatedgeometry, "associategeometry must be called first"
        m = self._ensure_model(geom0)
        self._ensure_dofs(dchi)
        block = assembler.block
        if isinstance(assembler, SysmatAssemblerHIP):
            m.mass(self._mass_params(), block)
            return DeviceMatrix(self, block, "M")
        nrows = dchi.nfreedofs if block == _lib.FS_BLOCK_FF else dchi.nalldofs
        d = m.mass(self._mass_params(), block, nrows)
        return sp.diags(d, format="csr")

    # -- load / dof front end on the device --
    def number_dofs(self, geom0, dchi, perm=None):
        """numberdofs!(dchi[, perm]) on the device: fills dchi.dofnums and
        dchi.nfreedofs from the fixed flags set by setebc, and installs the
        numbering in the device model."""
        m = self._ensure_model(geom0)
        dn, nf = m.number_dofs(dchi.is_fixed, perm)
        dchi.dofnums, dchi.nfreedofs = dn, nf
        self._dofs_key = (dchi.nalldofs, dchi.nfreedofs, hash(dchi.dofnums.tobytes()))
        self._patterns = {}
        if self._normals is not None:
            m.normals_upload(self._normals, self._normal_valid)
        return dchi

    def distribloads(self, geom0, dchi, force, npts=None, block=_lib.FS_BLOCK_FF):
        """distribloads(FEMMBase(IntegDomain(fes, rule)), geom0, dchi,
        ForceIntensity(f), 2) on the device: the free rows (block FF, as
        SysvecAssemblerFBlock) or all rows.  npts: 1, 3 (default) or 6 for
        TriRule on T3, 4 for GaussRule(2, 2) on Q4.  force: a 6-vector, or a
        callable taking the rule's point locations (n x 3) and returning n x 6
        force intensities (the ForceIntensity function)."""
        m = self._ensure_model(geom0)
        self._ensure_dofs(dchi)
        npts = (3 if self.nen == 3 else 4) if npts is None else int(npts)
        if callable(force):
            X = m.load_points(npts).reshape(-1, 3)
            force = np.asarray(force(X), dtype=float).reshape(self.conn.shape[0], npts, 6)
        nrows = dchi.nfreedofs if block == _lib.FS_BLOCK_FF else dchi.nalldofs
        return m.surface_load(npts, force, block, nrows)

    # -- inspectintegpoints --
    def _points(self, xyz, fl, tangents=False):
        """Location (nlist x npts x 3) and, if asked, tangent matrix (nlist x
        npts x 3 x 2) of every integration point of the listed elements: the
        centroid and J0 for T3 (FEMMShellT3FFModule.jl:908-911), loc and J of
        locjac! at the Gauss points for Q4.  fl None: all elements."""
        X = xyz[self.conn] if fl is None else xyz[self.conn[fl]]
        if self.nen == 3:
            loc = X.mean(axis=1)[:, None, :]
            J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]], axis=2)[:, None] if tangents else None
            return loc, J
        g = 0.577350269189626
        pts = [(-g, -g), (-g, g), (g, -g), (g, g)]
        Nm = np.array([_q4_N(xi, eta) for xi, eta in pts])  # points x nodes
        loc = np.matmul(Nm, X)
        J = None
        if tangents:
            G = np.array([_q4_gradN(xi, eta) for xi, eta in pts])  # points x nodes x 2
            J = np.einsum("eak,pal->epkl", X, G)
        return loc, J

    def resultants(self, geom0, u, felist=None, outputcsys=None, transform_once=False):
        """All stress resultants of the listed elements on the device, as
        inspectintegpoints computes them: (nlist x npts x 8 with N11 N22 N12
        M11 M22 M12 Q1 Q2 in the output csys, point locations nlist x npts x 3).
        outputcsys: None (the FEMM's default: mcsys, or the first layup csys),
        a 3 x 3 matrix, or a callable (X, J, e, q) -> 3 x 3 evaluated per point.
        transform_once: for Q4RSComp, take the displacements into the element
        frame once instead of the reference's twice."""
        assert self._associatedgeometry, "associategeometry must be called first"
        xyz = np.asarray(geom0, dtype=float)
        m = self._ensure_model(xyz)
        vals = np.asarray(getattr(u, "values", u), dtype=float).reshape(xyz.shape[0], 6)
        fl = None if felist is None else np.asarray(felist, dtype=np.int64).ravel()
        if fl is not None and fl.size and (fl.min() < 0 or fl.max() >= self.conn.shape[0]):
            raise IndexError("inspectintegpoints: element number out of range")
        ocs = outputcsys
        if ocs is None:
            ocs = self.layup_groups[0][0].csys if self.composite else self.mcsys
        if fl is None and not callable(ocs):
            # every element, no tangents needed: the point locations come from
            # the device (the load-point kernel's TriRule(1) / GaussRule(2, 2)
            # points are these integration points, in this order)
            npts = 1 if self.nen == 3 else 4
            loc, J = m.load_points(npts).reshape(self.conn.shape[0], npts, 3), None
        else:
            loc, J = self._points(xyz, fl, tangents=callable(ocs))
        if callable(ocs):
            npts = loc.shape[1]
            enum = np.arange(self.conn.shape[0]) if fl is None else fl
            ocs = np.array([[np.asarray(ocs(loc[l, j], J[l, j], int(enum[l]), j), dtype=float)
                             for j in range(npts)] for l in range(enum.shape[0])]).reshape(-1, 3, 3)
        flags = _lib.FS_RES_LAYERED_Q4_SINGLE_T if transform_once else 0
        out, mismatch = m.resultants(self._shell_params(), vals, fl, ocs, flags)
        if mismatch:
            warnings.warn("Coordinate systems mismatched?")
        return out, loc

    _QUANTITY = {"bending": 3, "moment": 3, "bending_moment": 3, "transverse_shear": 6, "transverse": 6,
                 "shear": 6, "membrane_force": 0, "membrane": 0}

    def inspectintegpoints(self, geom0, u, *args, outputcsys=None, transform_once=False):
        """inspectintegpoints(femm, geom0, u[, dT], felist, inspector, idat[,
        quantity]; outputcsys) (FEMMShellT3FFModule.jl:850-962,
        FEMMShellQ4RSModule.jl:1061-1170, FEMMShellT3FFCompModule.jl:809-943,
        FEMMShellQ4RSCompModule.jl:1077-1209).  The resultants come from the
        device in one launch; the inspector is then called per element and
        point as idat = inspector(idat, e, conn, ecoords, out, loc) with
        out = the moments (default, :bending/:moment), the membrane forces
        (:membrane) or the shear forces (:shear, out[2] = 0).  Element numbers
        are 0-based."""
        if len(args) >= 3 and callable(args[1]):
            felist, inspector, idat, rest = args[0], args[1], args[2], args[3:]
        else:
            felist, inspector, idat, rest = args[1], args[2], args[3], args[4:]
        quantity = str(rest[0]).lstrip(":") if rest else "Cauchy"
        k0 = self._QUANTITY.get(quantity, 3)
        nk = 2 if k0 == 6 else 3
        fl = np.asarray(felist, dtype=np.int64).ravel()
        res, loc = self.resultants(geom0, u, fl, outputcsys, transform_once)
        xyz = np.asarray(geom0, dtype=float)
        for l, e in enumerate(fl):
            conn = self.conn[e]
            xe = xyz[conn]
            for j in range(res.shape[1]):
                out = np.zeros(3)
                out[:nk] = res[l, j, k0:k0 + nk]
                idat = inspector(idat, int(e), conn, xe, out, loc[l, j])
        return idat

    def fieldfromintegpoints(self, geom0, u, quantity, component, outputcsys=None, transform_once=False):
        """Nodal field of one resultant component (FinEtools
        fieldfromintegpoints with inverse-distance interpolation, as
        test_shell_statics.jl:690-728 uses it): every integration point
        contributes its value to the element's nodes with weights 1/d^2 (d
        the point-node distance, floored at 1e-9 of the smallest nonzero one
        of the element), and each node takes the weighted mean.  quantity:
        membrane, moment or shear; component: 0-based.  The resultants come
        from the device; the interpolation runs on the host."""
        k0 = self._QUANTITY.get(str(quantity).lstrip(":"), 3)
        res, loc = self.resultants(geom0, u, None, outputcsys, transform_once)
        v = res[:, :, k0 + int(component)]
        X = np.asarray(geom0, dtype=float)[self.conn]               # E x nen x 3
        d2 = np.sum((X[:, None, :, :] - loc[:, :, None, :]) ** 2, axis=3)  # E x npts x nen
        pos = np.where(d2 > 0, d2, np.inf).min(axis=2, keepdims=True)
        d2 = np.maximum(d2, pos / 1e9)
        w = 1.0 / d2
        N = np.asarray(geom0).shape[0]
        num = np.zeros(N)
        den = np.zeros(N)
        np.add.at(num, np.repeat(self.conn[:, None, :], v.shape[1], axis=1).ravel(), (w * v[:, :, None]).ravel())
        np.add.at(den, np.repeat(self.conn[:, None, :], v.shape[1], axis=1).ravel(), w.ravel())
        out = np.zeros(N)
        np.divide(num, den, out=out, where=den > 0)
        return out

    def elemfieldfromintegpoints(self, geom0, u, quantity, component, outputcsys=None, transform_once=False):
        """Element field of one resultant component: the mean over the
        element's integration points (FinEtools elemfieldfromintegpoints)."""
        k0 = self._QUANTITY.get(str(quantity).lstrip(":"), 3)
        res, _ = self.resultants(geom0, u, None, outputcsys, transform_once)
        return res[:, :, k0 + int(component)].mean(axis=1)

    def num_normals(self):
        v = self._normal_valid
        return len(v), int(np.count_nonzero(~v))


def _csr(vals, colval, rowptr, n):
    """scipy CSR view of library output: the arrays are already a valid,
    sorted CSR (int64 indices), so they are attached without the copy and
    index-dtype conversion of the validating constructor."""
    K = sp.csr_matrix((n, n), dtype=np.float64)
    K.data, K.indices, K.indptr = vals, colval, rowptr
    K.has_sorted_indices = True
    return K


def _element_normal(J0):
    e1 = J0[:, 0] / np.linalg.norm(J0[:, 0])
    n = np.cross(e1, J0[:, 1])
    return n / np.linalg.norm(n)


def _q4_N(xi, eta):
    return 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])


def _q4_gradN(xi, eta):
    return 0.25 * np.array([[-(1 - eta), -(1 - xi)], [(1 - eta), -(1 + xi)], [(1 + eta), (1 + xi)],
                            [-(1 + eta), (1 - xi)]])


def _q4_jac(X, xi, eta):
    gN = 0.25 * np.array([[-(1 - eta), -(1 - xi)], [(1 - eta), -(1 + xi)], [(1 + eta), (1 + xi)],
                          [-(1 + eta), (1 - xi)]])
    return X.T @ gN


class FEMMShellT3FF(_ShellFEMM):
    """FEMMShellT3FFModule.FEMMShellT3FF (FEMMShellT3FFModule.jl:101-112)."""
    nen = 3
    default_stab_c = 5 / 12 / 1.5  # _stab_fun, :36


class FEMMShellQ4RS(_ShellFEMM):
    """FEMMShellQ4RSModule.FEMMShellQ4RS (FEMMShellQ4RSModule.jl:75-85)."""
    nen = 4
    default_stab_c = 0.1  # _stab_fun, :32


class _CompositeMixin:
    composite = True

    def __init__(self, conn, layup, stab_c=None, ctx=None, layup_groups=None):
        """layup: a CompositeLayup for all elements, or None with layup_groups =
        [(layup, element_indices), ...] (the layup_groups field of the reference)."""
        if layup_groups is None:
            layup_groups = [(layup, np.arange(np.asarray(conn).shape[0]))]
        super().__init__(conn, None, thickness=None, stab_c=stab_c, ctx=ctx)
        self.layup_groups = layup_groups
        E = self.conn.shape[0]
        self._group = np.full(E, -1, dtype=np.int32)
        for g, (_, eset) in enumerate(layup_groups):
            self._group[np.asarray(eset)] = g
        if (self._group < 0).any():
            raise ValueError("every element must belong to a layup group")

    def associategeometry(self, geom0):
        """Nodal normals of the layered FEMMs (FEMMShellT3FFCompModule.jl:500-521,
        FEMMShellQ4RSCompModule.jl:462-503) on the device: the layup csys is
        constant per group, so its normal crosses the ABI per group."""
        m = self._ensure_model(geom0)
        gn = np.array([lg[0].csys[:, 2] for lg in self.layup_groups], dtype=float)
        self._normals, self._normal_valid = m.normals_layered(self.threshold_angle, gn, self._group)
        self._associatedgeometry = True
        return self

    def _csys_normal_function(self):
        groups = self._group
        lay = [lg[0] for lg in self.layup_groups]
        return lambda X, J, e, q: lay[groups[e]].csys[:, 2]

    def _shell_params(self):
        p = _lib.ShellParams()
        p.composite = 1
        p.formulation = int(self.transv_shear_formulation)
        p.stab_c = float(self.stab_c)
        p.drilling_stiffness_scale = float(self.drilling_stiffness_scale)
        ng = len(self.layup_groups)
        abdh = np.zeros((ng, 31))
        csys = np.zeros((ng, 9))
        tg = np.zeros(ng)
        for g, (lay, _) in enumerate(self.layup_groups):
            A, B, D = lay.laminate_stiffnesses()
            H = lay.laminate_transverse_stiffness()
            abdh[g] = np.concatenate([A.T.ravel(), B.T.ravel(), D.T.ravel(), H.T.ravel()])
            csys[g] = lay.csys.T.ravel()
            tg[g] = lay.thickness()
        self._keep = (abdh, csys, tg)  # alive for the duration of the call
        p.ngroups = ng
        p.group_ABDH = _lib.ptr(abdh)
        p.group_csys = _lib.ptr(csys)
        p.group_thickness = _lib.ptr(tg)
        p.elem_group = _lib.ptr(self._group, _lib.C.c_int32)
        return p

    def _mass_params(self):
        p = _lib.MassParams()
        p.composite = 1
        ng = len(self.layup_groups)
        inert = np.array([lay.laminate_inertia() for lay, _ in self.layup_groups], dtype=float).reshape(ng, 2)
        self._mkeep = (inert,)
        p.ngroups = ng
        p.group_inertia = _lib.ptr(inert)
        p.elem_group = _lib.ptr(self._group, _lib.C.c_int32)
        return p


class FEMMShellT3FFComp(_CompositeMixin, FEMMShellT3FF):
    """FEMMShellT3FFCompModule.FEMMShellT3FFComp (FEMMShellT3FFCompModule.jl:63-77)."""


class FEMMShellQ4RSComp(_CompositeMixin, FEMMShellQ4RS):
    """FEMMShellQ4RSCompModule.FEMMShellQ4RSComp (FEMMShellQ4RSCompModule.jl:88-98)."""


def make(conn, material_or_layup, thickness=None, stab_c=None, mcsys=None, ctx=None):
    """formul.make(IntegDomain(fes, rule, thickness), material[, stab_fun]) for the
    element family of `conn` (3 -> T3FF, 4 -> Q4RS); a CompositeLayup selects
    the layered FEMMs."""
    from .material import CompositeLayup
    nen = np.asarray(conn).shape[1]
    if isinstance(material_or_layup, CompositeLayup):
        cls = FEMMShellT3FFComp if nen == 3 else FEMMShellQ4RSComp
        return cls(conn, material_or_layup, stab_c=stab_c, ctx=ctx)
    cls = FEMMShellT3FF if nen == 3 else FEMMShellQ4RS
    return cls(conn, material_or_layup, thickness=thickness, stab_c=stab_c, mcsys=mcsys, ctx=ctx)
