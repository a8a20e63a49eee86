"""Continuous-batching request server for shell jobs on one GPU.

A request is an assembly-only, an assembly + modal-matrices or an
explicit-dynamics job on its own mesh (homogeneous or layered).  Requests
wait in a FIFO queue and are admitted whenever a slot is idle and the
request's device footprint fits in the HBM budget left by the requests in
flight, so requests that arrive while others are on the GPU start as slots
and memory free up.  Each slot owns a HIP context (its own non-blocking
stream), so the kernels of different requests overlap on the device, and the
host work of one request (pattern build, copies) overlaps the device work of
the others.  A request's mesh, matrices and state stay resident in HBM from
admission to completion.

Each request is computed exactly as it would be alone: the kernels are
deterministic, their arithmetic order does not depend on concurrent work and
a request shares no device data with the others, so results are bitwise
identical alone and in any mix.

Latency is measured from the request's arrival (submission) to its result
being on the host.
"""
import collections
import threading
import time

import numpy as np

from . import _lib, explicit, femm as FM

ASSEMBLY, MODAL, EXPLICIT = "assembly", "modal", "explicit"


def footprint_bytes(nen, nelem, nnodes, nrows, kind):
    """Upper estimate of the device bytes a request holds while it runs: mesh
    and adjacency, normals, dof map, element blocks, the CSR pattern, values
    and scatter tables of one block, the lumped mass and, for explicit jobs,
    the step vectors and power-iteration vectors."""
    nblk = nen * (nen + 1) // 2
    per_row = 6 * (9 if nen == 4 else 7)  # structured stencil (neighbour nodes incl. self) x 6
    nnz = nrows * per_row
    mesh = nnodes * (24 + 24 + 1 + 48 + 4 + 6 + 24 + 3 * 4 * 8) + nelem * nen * 12
    blocks = nelem * nblk * (36 * 8 + 4)
    csr = nnz * 12 + nrows * 16
    vec = nrows * 8 * (11 if kind == EXPLICIT else 1)
    return int(1.25 * (mesh + blocks + csr + vec)) + (4 << 20)


class Request:
    """One job.  `model` holds nen, xyz, conn, dchi and either `layup` (a
    CompositeLayup: layered FEMM) or `material` + `t` (homogeneous FEMM);
    explicit jobs also take V0 (free-dof velocities) and `steps`."""

    def __init__(self, rid, kind, model, steps=0, arrival=0.0, omega_iters=100, seed=0,
                 precision=_lib.FS_PREC_FP64):
        if kind not in (ASSEMBLY, MODAL, EXPLICIT):
            raise ValueError(f"unknown request kind {kind!r}")
        self.rid = rid
        self.kind = kind
        self.model = model
        self.steps = int(steps)
        self.arrival = float(arrival)
        self.omega_iters = int(omega_iters)
        self.seed = seed
        self.precision = int(precision)  # explicit jobs: FS_PREC_FP64, FS_PREC_MIXED or FS_PREC_MIXED_F32
        self.result = None
        self.error = None
        self.phase_ms = {}  # wall time of each phase of execute(), ms
        self.t_arrive = self.t_admit = self.t_done = None
        m = model
        d = m["dchi"]
        nrows = d.nalldofs if kind == ASSEMBLY else d.nfreedofs
        self.nelem = int(m["conn"].shape[0])
        self.footprint = footprint_bytes(int(m["conn"].shape[1]), self.nelem, int(m["xyz"].shape[0]), nrows, kind)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrive


def make_femm(model, ctx):
    if "layup" in model and model["layup"] is not None:
        return FM.make(model["conn"], model["layup"], stab_c=model.get("stab_c"), ctx=ctx)
    return FM.make(model["conn"], model["material"], thickness=model["t"], stab_c=model.get("stab_c"), ctx=ctx)


def execute(req, ctx):
    """Run one request on a context and return its host results:
    assembly  -> K (all dofs, CSR);
    modal     -> K (free-free CSR) and the lumped mass diagonal M;
    explicit  -> U, V after `steps` central-difference steps at 0.9 of the
                 stable dt from the device power iteration (FP64 or mixed
                 precision, per request), and that dt."""
    m = req.model
    phase = {}
    clock = [time.perf_counter()]

    def mark(name):
        now = time.perf_counter()
        phase[name] = phase.get(name, 0.0) + (now - clock[0]) * 1e3
        clock[0] = now

    f = make_femm(m, ctx)
    d = m["dchi"]
    nf = d.nfreedofs
    out = {}
    try:
        f.associategeometry(m["xyz"])
        mark("geometry")
        if req.kind == ASSEMBLY:
            out["K"] = f.stiffness(FM.SysmatAssemblerSparseCSRSymm(), m["xyz"], None, None, d)
            mark("stiffness")
        elif req.kind == MODAL:
            out["K"] = f.stiffness(FM.SysmatAssemblerFFBlock(nf), m["xyz"], None, None, d)
            mark("stiffness")
            out["M"] = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), m["xyz"],
                              d).diagonal()
            mark("mass")
        else:
            f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), m["xyz"], None, None, d)
            mark("stiffness")
            f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), m["xyz"], d)
            mark("mass")
            # a Gaussian start weighs the oscillatory top modes (a positive one
            # barely does, and can leave omega_max 10 % low after 30 steps)
            v0 = np.random.default_rng(req.seed).standard_normal(nf)
            om = explicit.omega_max(f, v0, maxit=req.omega_iters)
            dt = explicit.stable_dt(om)
            mark("omega")
            V0 = m["V0"] if m.get("V0") is not None else np.zeros(nf)
            r = explicit.run(f._model, 0.0, np.zeros(nf), V0, dt, req.steps, np.zeros(nf),
                             np.zeros(req.steps + 1), precision=req.precision)
            mark("explicit")
            out.update(U=r["U"], V=r["V"], dt=dt, omega_max=om, cd_ms=r["ms"])
        out["elements"] = req.nelem
        out["free_dofs"] = nf
    finally:
        if f._model is not None:
            f._model.close()
            f._model = None
        mark("release")
        req.phase_ms = phase
    return out


def device_memory(device=0):
    """(free, total) bytes of a HIP device, from the HIP runtime that
    libflexshell is linked against."""
    import ctypes as C
    _lib.load()
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    if hip.hipSetDevice(int(device)) != 0 or hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise _lib.FlexShellError(f"hipMemGetInfo failed on device {device}")
    return free.value, total.value


# device bytes a context may keep as spare buffers after its requests end
# (the buffer cache of libflexshell, 4 GiB per context)
CONTEXT_SPARE_BYTES = 4 << 30


def default_budget(device=0, fraction=0.9, nslots=0):
    """HBM budget for admission: a fraction of the device memory free now,
    less the spare buffers each of the nslots contexts may keep cached between
    its requests (memory that in-flight footprints do not count)."""
    free, _total = device_memory(device)
    return max(int(fraction * free) - nslots * CONTEXT_SPARE_BYTES, 1 << 30)


def request_context(device=0):
    """A context set up as the server's slots are: a served request
    assembles its model once, so it takes the two-kernel assembly form, which
    needs only the pair plan, while the fused tiles' host plan (tens to
    hundreds of ms on large meshes) pays off over repeated assemblies.  A
    request run alone on such a context gives bitwise what it gives in a
    server."""
    c = _lib.Context(device)
    c.set_assembly(0)
    return c


class ShellServer:
    """nslots worker threads, one HIP context each; FIFO admission by
    footprint against `hbm_budget` bytes (a request larger than the whole
    budget is admitted only when nothing else runs)."""

    def __init__(self, nslots=4, hbm_budget=None, device=0):
        self.budget = int(default_budget(device, nslots=nslots) if hbm_budget is None else hbm_budget)
        self.contexts = [request_context(device) for _ in range(nslots)]
        self.in_use = 0
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.done = []
        self.stop = False
        self.running = 0
        self.max_concurrent = 0
        self.max_inflight_bytes = 0
        self.threads = [threading.Thread(target=self._slot, args=(w,), daemon=True) for w in range(nslots)]
        for t in self.threads:
            t.start()

    def submit(self, req):
        with self.cv:
            req.t_arrive = time.perf_counter()
            self.queue.append(req)
            self.cv.notify_all()

    def _admit(self):
        if not self.queue:
            return None
        req = self.queue[0]
        if self.in_use + req.footprint <= self.budget or self.running == 0:
            self.queue.popleft()
            self.in_use += req.footprint
            self.running += 1
            self.max_inflight_bytes = max(self.max_inflight_bytes, self.in_use)
            self.max_concurrent = max(self.max_concurrent, self.running)
            req.t_admit = time.perf_counter()
            return req
        return None

    def _slot(self, w):
        ctx = self.contexts[w]
        while True:
            with self.cv:
                req = None
                while not self.stop:
                    req = self._admit()
                    if req is not None:
                        break
                    self.cv.wait()
                if req is None:
                    return
            try:
                req.result = execute(req, ctx)
            except Exception as e:  # reported per request; the slot keeps serving
                req.error = f"{type(e).__name__}: {e}"
            with self.cv:
                req.t_done = time.perf_counter()
                self.in_use -= req.footprint
                self.running -= 1
                self.done.append(req)
                self.cv.notify_all()

    def wait(self, n):
        """Block until n requests have completed."""
        with self.cv:
            while len(self.done) < n:
                self.cv.wait()

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        for c in self.contexts:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def replay(server, requests):
    """Submit requests at their arrival times (seconds from the start) and
    wait for all of them; returns the latency / throughput summary."""
    base = len(server.done)
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        server.submit(r)
    server.wait(base + len(requests))
    wall = time.perf_counter() - t0
    return summarize(requests, wall, server)


def summarize(requests, wall, server=None):
    lat = np.array([r.latency for r in requests]) * 1e3
    ok = [r for r in requests if r.error is None]
    out = {"requests": len(requests), "failed": len(requests) - len(ok), "wall_s": wall,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "throughput_req_s": len(requests) / wall,
           "elements_per_s": sum(r.nelem for r in ok) / wall}
    for kind in (ASSEMBLY, MODAL, EXPLICIT):
        lk = np.array([r.latency for r in requests if r.kind == kind]) * 1e3
        if lk.size:
            out[f"{kind}_p50_ms"] = float(np.percentile(lk, 50))
            out[f"{kind}_p99_ms"] = float(np.percentile(lk, 99))
    phases = {}
    for r in ok:
        for k, v in r.phase_ms.items():
            phases.setdefault(k, []).append(v)
    out["phase_mean_ms"] = {k: float(np.mean(v)) for k, v in phases.items()}
    wait = [r.t_admit - r.t_arrive for r in requests if r.t_admit is not None]
    if wait:
        out["queue_wait_p50_ms"] = float(np.percentile(wait, 50) * 1e3)
    if server is not None:
        out.update(slots=len(server.contexts), max_concurrent=server.max_concurrent,
                   max_inflight_gb=server.max_inflight_bytes / 1e9, budget_gb=server.budget / 1e9)
    return out
