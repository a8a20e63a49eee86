"""Batches of independent homogeneous shell models assembled and stepped in
one launch per kernel.

The models of a batch (one element family) are concatenated into a single
device model: disconnected meshes, one material group per model, and a
global equation numbering that keeps every model's free dofs first (model
order) and each node's dofs consecutive.  The CSR pattern is then block
diagonal and every kernel (normals, element stiffness, scatter, mass, the
central-difference step) processes all models in one launch; the step loop
is captured in a hipGraph.  Each model keeps its own material, time step,
damping and load history, and its results are bitwise identical to running
it alone: per element and per CSR row the arithmetic and its order do not
depend on the other models in the batch.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib, explicit
from .material import shell_material_stiffness


class ShellBatch:
    """models: list of dicts with keys xyz, conn, dchi (fe.DofField after
    numberdofs), material (rho used for the mass), t, and optionally stab_c,
    drilling_stiffness_scale."""

    def __init__(self, models, ctx=None, threshold_angle=30.0):
        if not models:
            raise ValueError("empty batch")
        self.ctx = ctx or _lib.default_context()
        nen = models[0]["conn"].shape[1]
        if any(m["conn"].shape[1] != nen for m in models):
            raise ValueError("a batch holds one element family")
        self.nen = nen
        self.models = models
        nm = len(models)
        N = np.array([m["xyz"].shape[0] for m in models])
        E = np.array([m["conn"].shape[0] for m in models])
        nfree = np.array([m["dchi"].nfreedofs for m in models])
        nall = np.array([m["dchi"].nalldofs for m in models])
        self.node_off = np.concatenate([[0], np.cumsum(N)])
        self.elem_off = np.concatenate([[0], np.cumsum(E)])
        self.free_off = np.concatenate([[0], np.cumsum(nfree)])
        fixed = nall - nfree
        self.fixed_off = np.concatenate([[0], np.cumsum(fixed)])
        self.nfree = int(nfree.sum())
        self.nall = int(nall.sum())
        xyz = np.concatenate([m["xyz"] for m in models])
        conn = np.concatenate([m["conn"] + self.node_off[k] for k, m in enumerate(models)])
        dn = []
        for k, m in enumerate(models):
            d = m["dchi"].dofnums
            nf = m["dchi"].nfreedofs
            g = np.where(d < nf, d + self.free_off[k], d - nf + self.nfree + self.fixed_off[k])
            dn.append(g)
        self.dofnums = np.concatenate(dn).astype(np.int64)
        self.elem_group = np.repeat(np.arange(nm, dtype=np.int32), E)
        self.row_group = np.repeat(np.arange(nm, dtype=np.int32), nfree)
        self.model = _lib.Model(self.ctx, nen, xyz, conn)
        self.model.set_dofs(self.dofnums, self.nall, self.nfree)
        self.normals, self.normal_valid = self.model.normals_isoparametric(threshold_angle)
        homog = np.zeros((nm, 16))
        rho_t = np.zeros((nm, 2))
        default_c = 5 / 12 / 1.5 if nen == 3 else 0.1
        for k, m in enumerate(models):
            Dps, Dt = shell_material_stiffness(m["material"])
            homog[k, :9] = Dps.T.ravel()
            homog[k, 9:13] = Dt.T.ravel()
            homog[k, 13] = m["t"]
            homog[k, 14] = m.get("stab_c", default_c)
            homog[k, 15] = m.get("drilling_stiffness_scale", 1.0)
            rho_t[k] = (m["material"].rho, m["t"])
        self._homog = homog
        self._rho_t = rho_t
        self._pattern = None

    def _shell_params(self):
        p = _lib.ShellParams()
        p.composite = 0
        p.formulation = 0
        p.ngroups = len(self.models)
        p.group_homog = _lib.ptr(self._homog)
        p.elem_group = _lib.ptr(self.elem_group, C.c_int32)
        return p

    def _mass_params(self):
        p = _lib.MassParams()
        p.composite = 0
        p.ngroups = len(self.models)
        p.group_rho_t = _lib.ptr(self._rho_t)
        p.elem_group = _lib.ptr(self.elem_group, C.c_int32)
        return p

    def pattern(self):
        if self._pattern is None:
            self._pattern = self.model.pattern(_lib.FS_BLOCK_FF)
        return self._pattern

    def stiffness(self, to_host=True):
        """FF-block stiffness of all models (block-diagonal CSR)."""
        if not to_host:
            self.assemble_device()
            return None
        nrows, rp, cv = self.pattern()
        vals = self.model.stiffness(self._shell_params(), _lib.FS_BLOCK_FF, len(cv))
        from .femm import _csr
        return _csr(vals, cv, rp, nrows)

    def assemble_device(self):
        """Assemble the FF-block stiffness into HBM without a host copy."""
        self.pattern()
        ctx = self.ctx
        p = self._shell_params()
        ctx.check(ctx.lib.fs_model_stiffness(self.model.h, C.byref(p), _lib.FS_BLOCK_FF, None), "fs_model_stiffness")

    def mass(self):
        return self.model.mass(self._mass_params(), _lib.FS_BLOCK_FF, self.nfree)

    def model_slice(self, k):
        """Free-dof range of model k in the batch vectors."""
        return slice(int(self.free_off[k]), int(self.free_off[k + 1]))

    def explicit(self, dts, cdamps, U0, V0, Fmag, gvals, nsteps, use_graph=True, probe_dofs=(), probe_stride=1,
                 precision=_lib.FS_PREC_FP64):
        """Central-difference loop of all models: dts, cdamps per model; U0,
        V0, Fmag concatenated free-dof vectors; gvals (nsteps+1) x nmodels."""
        nm = len(self.models)
        dts = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
        cd = np.ascontiguousarray(np.asarray(cdamps, dtype=np.float64))
        gv = np.ascontiguousarray(np.asarray(gvals, dtype=np.float64))
        if dts.shape != (nm,) or cd.shape != (nm,) or gv.shape != (nsteps + 1, nm):
            raise ValueError("batch explicit: per-model arrays have the wrong shape")
        n = self.nfree
        U0 = np.ascontiguousarray(np.asarray(U0, dtype=np.float64))
        V0 = np.ascontiguousarray(np.asarray(V0, dtype=np.float64))
        Fm = np.ascontiguousarray(np.asarray(Fmag, dtype=np.float64))
        probes = np.ascontiguousarray(np.asarray(probe_dofs, dtype=np.int64))
        stride = max(1, int(probe_stride))
        pout = np.zeros((nsteps // stride + 1, len(probes)))
        U = np.zeros(n)
        V = np.zeros(n)
        p = _lib.CDParams()
        p.nsteps = int(nsteps)
        p.precision = int(precision)
        p.probe_stride = stride
        p.ngroups = nm
        p.use_graph = 1 if use_graph else 0
        p.row_group = _lib.ptr(self.row_group, C.c_int32)
        p.group_dt = _lib.ptr(dts)
        p.group_cdamp = _lib.ptr(cd)
        ctx = self.ctx
        ctx.check(ctx.lib.fs_model_explicit(self.model.h, C.byref(p), _lib.ptr(U0), _lib.ptr(V0), _lib.ptr(Fm),
                                            _lib.ptr(gv), _lib.ptr(probes, C.c_int64) if len(probes) else None,
                                            int(len(probes)), _lib.ptr(pout) if len(probes) else None,
                                            _lib.ptr(U), _lib.ptr(V)), "fs_model_explicit")
        return {"U": U, "V": V, "probes": pout, "ms": self.model.timing(3), "build_ms": self.model.timing(4)}

    def close(self):
        self.model.close()


# Working set of one group's explicit loop (operator values and state) that
# stays resident in the 256 MB Infinity Cache of an MI355X while the group's
# steps run.  Measured on the 256-model T3 batch (500 steps) against one
# group over all models, which streams the whole operator from HBM every
# step: groups of about 225 MB gave 1.34x (FP64) and 1.35x
# (FS_PREC_MIXED_F32) the DOF-steps/s, 180 MB 1.31x and 1.34x, 135 MB 1.23x
# and 1.24x (more, smaller launches per step).
CACHE_GROUP_BYTES = 224 << 20


def step_bytes(model, value_bytes=8):
    """Bytes one central-difference step moves for a model: its free-free
    values at value_bytes each (f_I f_J per node pair and f_I^2 per node,
    f_I the free dofs of node I: the structural union of the element dof
    pairs) plus 72 B of state per free dof (U, V, A, invMC read; U, V, A
    written)."""
    conn = np.asarray(model["conn"])
    nen = conn.shape[1]
    a, b = np.triu_indices(nen, 1)
    pairs = np.sort(np.stack([conn[:, a].ravel(), conn[:, b].ravel()], axis=1), axis=1).astype(np.int64)
    key = np.unique(pairs[:, 0] * (1 << 32) + pairs[:, 1])
    d = model["dchi"]
    f = (np.asarray(d.dofnums) < d.nfreedofs).sum(axis=1).astype(np.int64)
    nnz = int((f * f).sum() + 2 * (f[key >> 32] * f[key & 0xffffffff]).sum())
    return int(nnz * value_bytes + 72 * d.nfreedofs)


def cache_groups(models, group_bytes=CACHE_GROUP_BYTES, value_bytes=8):
    """Consecutive groups of model indices whose step_bytes sum to at most
    group_bytes (a larger model forms a group of its own)."""
    out, cur, cb = [], [], 0
    for k, m in enumerate(models):
        sb = step_bytes(m, value_bytes)
        if cur and cb + sb > group_bytes:
            out.append(cur)
            cur, cb = [], 0
        cur.append(k)
        cb += sb
    if cur:
        out.append(cur)
    return out


class ShellBatchGroups:
    """A batch run as consecutive groups (cache_groups), each a ShellBatch:
    the explicit loop runs every step of one group before the next group,
    so that a group's operator is read from the Infinity Cache rather than
    from HBM on all but its first step.  Vectors are the concatenation over
    all models in model order, as for ShellBatch, and every model's results
    are bitwise those of running it alone."""

    def __init__(self, models, ctx=None, group_bytes=CACHE_GROUP_BYTES, value_bytes=8, threshold_angle=30.0,
                 streams=1):
        """streams > 1: the groups are dealt round-robin to that many
        contexts (streams) of ctx's device, and explicit() runs the contexts
        concurrently, one host thread each, so that one group's launches
        fill the gaps of another's."""
        if not models:
            raise ValueError("empty batch")
        if streams < 1:
            raise ValueError("streams must be at least 1")
        self.ctx = ctx or _lib.default_context()
        self.ctxs = [self.ctx] + [_lib.Context(self.ctx.device) for _ in range(streams - 1)]
        self.models = models
        self.groups = cache_groups(models, group_bytes, value_bytes)
        self.batches = [ShellBatch([models[k] for k in g], ctx=self.ctxs[i % streams], threshold_angle=threshold_angle)
                        for i, g in enumerate(self.groups)]
        nfree = np.array([m["dchi"].nfreedofs for m in models])
        self.free_off = np.concatenate([[0], np.cumsum(nfree)])
        self.nfree = int(nfree.sum())
        self.group_off = np.concatenate([[0], np.cumsum([b.nfree for b in self.batches])])

    def assemble_device(self):
        for b in self.batches:
            b.assemble_device()

    def mass(self):
        return np.concatenate([b.mass() for b in self.batches])

    def model_slice(self, k):
        return slice(int(self.free_off[k]), int(self.free_off[k + 1]))

    def timing(self, which):
        """Sum over the groups of fs_model_last_timing(which)."""
        return float(sum(b.model.timing(which) for b in self.batches))

    def explicit(self, dts, cdamps, U0, V0, Fmag, gvals, nsteps, use_graph=True, probe_dofs=(), probe_stride=1,
                 precision=_lib.FS_PREC_FP64):
        """ShellBatch.explicit over the groups (in turn on each context);
        "ms" is the summed device time of the groups' step loops, "wall_ms"
        the host time of the whole call."""
        import threading
        import time
        dts = np.asarray(dts, dtype=np.float64)
        cd = np.asarray(cdamps, dtype=np.float64)
        gv = np.asarray(gvals, dtype=np.float64)
        probes = np.asarray(probe_dofs, dtype=np.int64)
        stride = max(1, int(probe_stride))
        pout = np.zeros((nsteps // stride + 1, len(probes)))
        U = np.zeros(self.nfree)
        V = np.zeros(self.nfree)
        res = [None] * len(self.groups)
        errs = []

        def run(s):
            try:
                for gi in range(s, len(self.groups), len(self.ctxs)):
                    g, b = self.groups[gi], self.batches[gi]
                    lo, hi = int(self.group_off[gi]), int(self.group_off[gi + 1])
                    mine = np.nonzero((probes >= lo) & (probes < hi))[0]
                    res[gi] = (mine, b.explicit(dts[g], cd[g], U0[lo:hi], V0[lo:hi], Fmag[lo:hi], gv[:, g], nsteps,
                                                use_graph=use_graph, probe_dofs=probes[mine] - lo,
                                                probe_stride=stride, precision=precision))
            except Exception as e:  # re-raised in the caller's thread
                errs.append(e)
        t0 = time.perf_counter()
        if len(self.ctxs) == 1:
            run(0)
        else:
            th = [threading.Thread(target=run, args=(s,)) for s in range(len(self.ctxs))]
            for t in th:
                t.start()
            for t in th:
                t.join()
        wall = (time.perf_counter() - t0) * 1e3
        if errs:
            raise errs[0]
        ms = build = 0.0
        for gi, (mine, r) in enumerate(res):
            lo, hi = int(self.group_off[gi]), int(self.group_off[gi + 1])
            U[lo:hi] = r["U"]
            V[lo:hi] = r["V"]
            if len(mine):
                pout[:, mine] = r["probes"]
            ms += r["ms"]
            build += r["build_ms"]
        return {"U": U, "V": V, "probes": pout, "ms": ms, "build_ms": build, "wall_ms": wall}

    def close(self):
        for b in self.batches:
            b.close()
        for c in self.ctxs[1:]:
            c.close()


def run_plain(models, dts, cdamps, U0s, V0s, Fmags, gvals, nsteps, ctx=None):
    """The plain path: one model per launch sequence (own device model,
    assembly, mass and step loop per model, no graph).  Returns the per-model
    results and the summed device time of the step loops (ms)."""
    from .femm import FEMMShellQ4RS, FEMMShellT3FF, SysmatAssemblerFFBlock, SysmatAssemblerSparseDiag
    ctx = ctx or _lib.default_context()
    out, total_ms = [], 0.0
    for k, m in enumerate(models):
        cls = FEMMShellT3FF if m["conn"].shape[1] == 3 else FEMMShellQ4RS
        f = cls(m["conn"], m["material"], thickness=m["t"], stab_c=m.get("stab_c"), ctx=ctx)
        f.drilling_stiffness_scale = m.get("drilling_stiffness_scale", 1.0)
        f.associategeometry(m["xyz"])
        d = m["dchi"]
        nf = d.nfreedofs
        f.stiffness(SysmatAssemblerFFBlock(nf), m["xyz"], None, None, d)
        f.mass(SysmatAssemblerFFBlock(SysmatAssemblerSparseDiag(), nf, nf), m["xyz"], d)
        r = explicit.run(f._model, cdamps[k], U0s[k], V0s[k], dts[k], nsteps, Fmags[k], gvals[:, k])
        total_ms += r["ms"]
        out.append(r)
        f._model.close()
    return out, total_ms
