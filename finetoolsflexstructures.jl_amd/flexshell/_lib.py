"""ctypes binding of libflexshell.so (the C ABI declared in include/flexshell.h).

The library is the only compute path: there is no CPU fallback.  Loading fails
loudly when the shared object is missing; creating a context fails loudly when
no HIP device is present.
"""
import ctypes as C
import os

import numpy as np

FS_OK = 0
FS_T3 = 3
FS_Q4 = 4
FS_BLOCK_ALL = 0
FS_BLOCK_FF = 1
FS_SHEAR_AVERAGE_B = 0
FS_SHEAR_AVERAGE_K = 1
FS_PREC_FP64 = 0
FS_PREC_MIXED = 1
FS_PREC_MIXED_F32 = 2
FS_OCSYS_ISOPARAMETRIC = 0
FS_OCSYS_MATRIX = 1
FS_OCSYS_PER_POINT = 2
FS_RES_LAYERED_Q4_SINGLE_T = 1
FS_LOAD_UNIFORM = 0
FS_LOAD_PER_POINT = 1

_STATUS = {
    -1: "invalid argument",
    -2: "invalid call sequence",
    -3: "device error",
    -4: "out of memory",
    -5: "numerical failure",
    -6: "unsupported input",
}

LIB_PATH = os.environ.get("FLEXSHELL_LIB",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)), "libflexshell.so"))

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class ShellParams(C.Structure):
    _fields_ = [
        ("composite", C.c_int32),
        ("formulation", C.c_int32),
        ("thickness", C.c_double),
        ("stab_c", C.c_double),
        ("drilling_stiffness_scale", C.c_double),
        ("Dps", C.c_double * 9),
        ("Dt", C.c_double * 4),
        ("ngroups", C.c_int32),
        ("reserved", C.c_int32),
        ("group_ABDH", _dp),
        ("group_csys", _dp),
        ("group_thickness", _dp),
        ("elem_group", _i32p),
        ("group_homog", _dp),
    ]


class MassParams(C.Structure):
    _fields_ = [
        ("composite", C.c_int32),
        ("ngroups", C.c_int32),
        ("rho", C.c_double),
        ("thickness", C.c_double),
        ("group_inertia", _dp),
        ("elem_group", _i32p),
        ("group_rho_t", _dp),
    ]


class CDParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("cdamp", C.c_double),
        ("nsteps", C.c_int64),
        ("precision", C.c_int32),
        ("probe_stride", C.c_int32),
        ("ngroups", C.c_int32),
        ("use_graph", C.c_int32),
        ("row_group", _i32p),
        ("group_dt", _dp),
        ("group_cdamp", _dp),
        ("dt_mass", C.c_double),
    ]


# every symbol include/flexshell.h declares, with its ctypes signature
SIGNATURES = {
    "fs_version": (C.c_char_p, []),
    "fs_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "fs_ctx_destroy": (None, [C.c_void_p]),
    "fs_last_error": (C.c_char_p, [C.c_void_p]),
    "fs_model_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, _dp, C.c_int64, _i64p, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "fs_model_destroy": (None, [C.c_void_p]),
    "fs_model_set_dofs": (C.c_int, [C.c_void_p, _i64p, C.c_int64, C.c_int64, C.c_int]),
    "fs_model_normals_isoparametric": (C.c_int, [C.c_void_p, C.c_double, _dp, _u8p]),
    "fs_model_normals_layered": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _dp, _i32p, _dp, _u8p]),
    "fs_model_normals_upload": (C.c_int, [C.c_void_p, _dp, _u8p]),
    "fs_model_pattern": (C.c_int, [C.c_void_p, C.c_int, _i64p, _i64p]),
    "fs_model_pattern_get": (C.c_int, [C.c_void_p, C.c_int, _i64p, _i64p, C.c_int]),
    "fs_model_stiffness": (C.c_int, [C.c_void_p, C.POINTER(ShellParams), C.c_int, _dp]),
    "fs_model_element_blocks": (C.c_int, [C.c_void_p, C.POINTER(ShellParams), _dp]),
    "fs_model_mass": (C.c_int, [C.c_void_p, C.POINTER(MassParams), C.c_int, _dp]),
    "fs_model_explicit": (C.c_int, [C.c_void_p, C.POINTER(CDParams), _dp, _dp, _dp, _dp, _i64p, C.c_int32,
                                    _dp, _dp, _dp]),
    "fs_model_omega_max": (C.c_int, [C.c_void_p, _dp, C.c_int32, _dp]),
    "fs_model_omega_max_noalloc": (C.c_int, [C.c_void_p, _dp, C.c_int32, _dp]),
    "fs_model_number_dofs": (C.c_int, [C.c_void_p, _u8p, _i64p, _i64p, _i64p]),
    "fs_model_surface_load": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int, _dp]),
    "fs_model_load_points": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "fs_model_resultants": (C.c_int, [C.c_void_p, C.POINTER(ShellParams), _dp, C.c_int64, _i32p, C.c_int32, _dp,
                                      C.c_int32, _dp, _i32p]),
    "fs_model_last_timing": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "fs_model_mixed_info": (C.c_int, [C.c_void_p, _i64p]),
    "fs_ctx_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _i32p]),
    "fs_ctx_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int32]),
    "fs_ctx_set_pipeline": (C.c_int, [C.c_void_p, C.c_int]),
    "fs_ctx_set_assembly": (C.c_int, [C.c_void_p, C.c_int]),
    "fs_ctx_set_cd_lanes": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fs_coo_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int64, _i64p, _i64p,
                                C.c_int, C.POINTER(C.c_void_p)]),
    "fs_coo_pattern": (C.c_int, [C.c_void_p, _i64p]),
    "fs_coo_pattern_get": (C.c_int, [C.c_void_p, _i64p, _i64p, C.c_int]),
    "fs_coo_values": (C.c_int, [C.c_void_p, _dp, _dp]),
    "fs_coo_destroy": (None, [C.c_void_p]),
}

_lib = None


def load():
    """Load libflexshell.so and attach signatures; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libflexshell.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
            "(hipcc for gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype=C.c_double):
    """Pointer to the data of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class FlexShellError(RuntimeError):
    pass


class Context:
    """One HIP device + stream (fs_ctx)."""

    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        st = lib.fs_ctx_create(int(device), C.byref(h))
        if st != FS_OK:
            raise FlexShellError(f"fs_ctx_create(device={device}) failed: {_STATUS.get(st, st)} "
                                 "(a HIP device is required)")
        self.lib = lib
        self.h = h
        self.device = int(device)

    def check(self, st, what):
        if st != FS_OK:
            msg = self.lib.fs_last_error(self.h).decode()
            raise FlexShellError(f"{what}: {_STATUS.get(st, st)}: {msg}")

    def tuning(self, nen, formulation=FS_SHEAR_AVERAGE_B):
        """(variant, elements per workgroup, workgroup size) of the element
        kernel chosen by the autotuner (variant -1: not tuned yet)."""
        out = np.zeros(3, dtype=np.int32)
        self.check(self.lib.fs_ctx_tuning(self.h, int(nen), int(formulation), ptr(out, C.c_int32)), "fs_ctx_tuning")
        return tuple(int(v) for v in out)

    def set_tuning(self, nen, formulation, variant):
        self.check(self.lib.fs_ctx_set_tuning(self.h, int(nen), int(formulation), int(variant)),
                   "fs_ctx_set_tuning")

    def set_pipeline(self, ranges):
        """Element ranges of the assembly pipeline (1: no overlap of the
        element kernel and the scatter)."""
        self.check(self.lib.fs_ctx_set_pipeline(self.h, int(ranges)), "fs_ctx_set_pipeline")

    def set_assembly(self, mode):
        """Stiffness assembly form: 1 fused element tiles (default), 0 element
        blocks through HBM and the pair scatter."""
        self.check(self.lib.fs_ctx_set_assembly(self.h, int(mode)), "fs_ctx_set_assembly")

    def set_cd_lanes(self, lanes_per_node, positions_per_lane=0):
        """Row products of the FP64 explicit step: 0 = four lanes per row,
        8 / 16 = lanes per node over the node's rows together, each lane
        loading `positions_per_lane` values of a row at once (0: chosen from
        the longest row)."""
        self.check(self.lib.fs_ctx_set_cd_lanes(self.h, int(lanes_per_node), int(positions_per_lane)),
                   "fs_ctx_set_cd_lanes")

    def close(self):
        if self.h:
            self.lib.fs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("FLEXSHELL_DEVICE", "0")))
    return _default_ctx


class Model:
    """A mesh resident on the device (fs_model)."""

    def __init__(self, ctx, nen, xyz, conn, index_base=0):
        self.ctx = ctx
        xyz = np.asarray(xyz, dtype=np.float64)
        N = xyz.shape[0]
        xcm = np.ascontiguousarray(xyz.T)  # N x 3 column-major == 3 planes
        conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int64))
        E = conn.shape[0]
        if conn.shape[1] != nen:
            raise ValueError(f"connectivity has {conn.shape[1]} columns, expected {nen}")
        h = C.c_void_p()
        ctx.check(ctx.lib.fs_model_create(ctx.h, int(nen), N, ptr(xcm), E, ptr(conn, C.c_int64),
                                          int(index_base), C.byref(h)), "fs_model_create")
        self.h = h
        self.nen = nen
        self.N = N
        self.E = E
        self.nall = None
        self.nfree = None

    def set_dofs(self, dofnums, nalldofs, nfreedofs, index_base=0):
        d = np.asfortranarray(np.asarray(dofnums, dtype=np.int64))
        flat = np.ascontiguousarray(d.T)  # column-major N x 6
        self.ctx.check(self.ctx.lib.fs_model_set_dofs(self.h, ptr(flat, C.c_int64), int(nalldofs), int(nfreedofs),
                                                      int(index_base)), "fs_model_set_dofs")
        self.nall, self.nfree = int(nalldofs), int(nfreedofs)

    def normals_isoparametric(self, threshold_deg=30.0):
        out = np.zeros((3, self.N))
        valid = np.zeros(self.N, dtype=np.uint8)
        self.ctx.check(self.ctx.lib.fs_model_normals_isoparametric(self.h, float(threshold_deg), ptr(out),
                                                                    ptr(valid, C.c_uint8)),
                       "fs_model_normals_isoparametric")
        return out.T.copy(), valid.astype(bool)

    def normals_layered(self, threshold_deg, group_normals, elem_group=None):
        """Normals of a layered shell from the per-group layup csys normals
        (group_normals: ngroups x 3; elem_group: nelem ints or None)."""
        gn = np.ascontiguousarray(np.asarray(group_normals, dtype=np.float64).reshape(-1, 3))
        eg = None if elem_group is None else np.ascontiguousarray(np.asarray(elem_group, dtype=np.int32))
        out = np.zeros((3, self.N))
        valid = np.zeros(self.N, dtype=np.uint8)
        self.ctx.check(self.ctx.lib.fs_model_normals_layered(self.h, float(threshold_deg), int(gn.shape[0]), ptr(gn),
                                                              ptr(eg, C.c_int32), ptr(out), ptr(valid, C.c_uint8)),
                       "fs_model_normals_layered")
        return out.T.copy(), valid.astype(bool)

    def normals_upload(self, normals, valid):
        n = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).T)
        v = np.ascontiguousarray(np.asarray(valid, dtype=np.uint8))
        self.ctx.check(self.ctx.lib.fs_model_normals_upload(self.h, ptr(n), ptr(v, C.c_uint8)),
                       "fs_model_normals_upload")

    def pattern(self, block):
        nr, nz = C.c_int64(), C.c_int64()
        self.ctx.check(self.ctx.lib.fs_model_pattern(self.h, int(block), C.byref(nr), C.byref(nz)),
                       "fs_model_pattern")
        rp = np.zeros(nr.value + 1, dtype=np.int64)
        cv = np.zeros(nz.value, dtype=np.int64)
        self.ctx.check(self.ctx.lib.fs_model_pattern_get(self.h, int(block), ptr(rp, C.c_int64),
                                                         ptr(cv, C.c_int64), 0), "fs_model_pattern_get")
        return nr.value, rp, cv

    def stiffness(self, params, block, nnz=None):
        """CSR values of the block; nnz None keeps them on the device only."""
        out = None if nnz is None else np.zeros(nnz)
        self.ctx.check(self.ctx.lib.fs_model_stiffness(self.h, C.byref(params), int(block), ptr(out)),
                       "fs_model_stiffness")
        return out

    def element_blocks(self, params):
        nblk = self.nen * (self.nen + 1) // 2
        out = np.zeros((self.E, nblk, 6, 6))
        self.ctx.check(self.ctx.lib.fs_model_element_blocks(self.h, C.byref(params), ptr(out)),
                       "fs_model_element_blocks")
        return out

    def number_dofs(self, is_fixed, perm=None):
        """numberdofs! on the device: (dofnums N x 6, nfree); the numbering is
        installed in the model as set_dofs would."""
        fx = np.ascontiguousarray(np.asarray(is_fixed, dtype=np.uint8).reshape(self.N, 6))
        pm = None if perm is None else np.ascontiguousarray(np.asarray(perm, dtype=np.int64))
        cm = np.zeros(6 * self.N, dtype=np.int64)
        nf = C.c_int64(0)
        self.ctx.check(self.ctx.lib.fs_model_number_dofs(self.h, ptr(fx, C.c_uint8), ptr(pm, C.c_int64),
                                                         ptr(cm, C.c_int64), C.byref(nf)), "fs_model_number_dofs")
        self.nall, self.nfree = 6 * self.N, int(nf.value)
        return cm.reshape(6, self.N).T.copy(), int(nf.value)

    def load_points(self, npts):
        out = np.empty((self.E, npts, 3))  # every entry is written
        self.ctx.check(self.ctx.lib.fs_model_load_points(self.h, int(npts), ptr(out)), "fs_model_load_points")
        return out

    def surface_load(self, npts, force, block, nrows):
        """Consistent nodal loads (distribloads) over the block's rows; force:
        a 6-vector or E x npts x 6 values at the rule's points."""
        f = np.ascontiguousarray(np.asarray(force, dtype=np.float64))
        kind = FS_LOAD_UNIFORM if f.size == 6 else FS_LOAD_PER_POINT
        if kind == FS_LOAD_PER_POINT and f.size != self.E * npts * 6:
            raise ValueError("force must be a 6-vector or nelem x npts x 6")
        out = np.zeros(nrows)
        self.ctx.check(self.ctx.lib.fs_model_surface_load(self.h, int(npts), int(kind), ptr(f), int(block), ptr(out)),
                       "fs_model_surface_load")
        return out

    def resultants(self, params, u, felist=None, ocs=None, flags=0):
        """Stress resultants N11 N22 N12 M11 M22 M12 Q1 Q2 at the integration
        points (nlist x npts x 8) and the csys-mismatch flag.  u: N x 6 nodal
        values; felist: 0-based element numbers or None (all); ocs: None
        (isoparametric), one 3 x 3 matrix, or nlist x npts x 3 x 3."""
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(self.N, 6))
        fl = None if felist is None else np.ascontiguousarray(np.asarray(felist, dtype=np.int32).ravel())
        nlist = self.E if fl is None else fl.shape[0]
        npts = 1 if self.nen == 3 else 4
        if ocs is None:
            kind, oc = FS_OCSYS_ISOPARAMETRIC, None
        else:
            m = np.asarray(ocs, dtype=np.float64)
            if m.size == 9:
                kind = FS_OCSYS_MATRIX
            elif m.size == 9 * nlist * npts:
                kind = FS_OCSYS_PER_POINT
            else:
                raise ValueError("ocs must be one 3 x 3 matrix or nlist x npts x 3 x 3")
            oc = np.ascontiguousarray(np.swapaxes(m.reshape(-1, 3, 3), 1, 2))  # column-major matrices
        out = np.empty((nlist, npts, 8))  # every entry is written
        mis = C.c_int32(0)
        self.ctx.check(self.ctx.lib.fs_model_resultants(self.h, C.byref(params), ptr(u), int(nlist),
                                                        ptr(fl, C.c_int32), int(kind), ptr(oc), int(flags), ptr(out),
                                                        C.byref(mis)), "fs_model_resultants")
        return out, bool(mis.value)

    def mass(self, params, block, nrows=None):
        """Lumped-mass diagonal of the block; nrows None keeps it on the device only."""
        out = None if nrows is None else np.zeros(nrows)
        self.ctx.check(self.ctx.lib.fs_model_mass(self.h, C.byref(params), int(block), ptr(out)), "fs_model_mass")
        return out

    def timing(self, which):
        v = C.c_double()
        self.ctx.check(self.ctx.lib.fs_model_last_timing(self.h, int(which), C.byref(v)), "fs_model_last_timing")
        return v.value

    def mixed_info(self):
        """(FP32 value pairs, shear rows per element, neighbour entries, slices)
        of the mixed-precision operator."""
        info = np.zeros(4, dtype=np.int64)
        self.ctx.check(self.ctx.lib.fs_model_mixed_info(self.h, ptr(info, C.c_int64)), "fs_model_mixed_info")
        return tuple(int(v) for v in info)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.fs_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
