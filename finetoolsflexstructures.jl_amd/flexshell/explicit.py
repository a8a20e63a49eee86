"""Explicit central-difference dynamics on device-resident K and M.

cd_loop mirrors _cd_loop!(M, K, ksi, U0, V0, tend, dt, force!, peek)
(examples/shells/dynamics/homogeneous/explicit/plate_expl_examples.jl:61-94)
on the FF-block stiffness and lumped mass that the FEMM left in HBM
(stiffness/mass with SysmatAssemblerFFBlock).  The force closure is given as
F(t) = g(t) * Fmag; peek is replaced by probe dofs recorded every
probe_stride steps.
"""
import ctypes as C

import numpy as np

from . import _lib


def adjust_steps(tend, dt):
    """nsteps and dt exactly as _cd_loop! derives them (:72-75)."""
    nsteps = int(round(tend / dt))
    if nsteps * dt < tend:
        dt = tend / (nsteps + 1)
    return nsteps, dt


def sample_times(nsteps, dt):
    """t_n as the loop accumulates them (t = t + dt)."""
    t = np.zeros(nsteps + 1)
    acc = 0.0
    for n in range(1, nsteps + 1):
        acc = acc + dt
        t[n] = acc
    return t


def cd_loop(femm, cdamp, U0, V0, tend, dt, Fmag, g=None, probe_dofs=(), probe_stride=1,
            precision=_lib.FS_PREC_FP64):
    """Run the central-difference loop; returns a dict with the final U, V,
    the probe history, the adjusted dt, nsteps and the device time (ms)."""
    model = femm._model
    if model is None:
        raise RuntimeError("cd_loop: the FEMM has no device model (call stiffness/mass first)")
    dt_mass = dt  # invMC is formed before the adjustment (:70-71)
    nsteps, dt = adjust_steps(tend, dt)
    t = sample_times(nsteps, dt)
    gv = np.ones(nsteps + 1) if g is None else np.array([float(g(tt)) for tt in t])
    return run(model, cdamp, U0, V0, dt, nsteps, Fmag, gv, probe_dofs, probe_stride, precision, dt_mass=dt_mass)


def run(model, cdamp, U0, V0, dt, nsteps, Fmag, gvals, probe_dofs=(), probe_stride=1,
        precision=_lib.FS_PREC_FP64, dt_mass=None, use_graph=False):
    """fs_model_explicit with an explicit step count and sampled g values;
    invMC is formed with dt_mass (None: dt).  use_graph: the steps run as one
    captured graph, which the model keeps and replays on later calls with the
    same sizes, dt and load factors (result key "graph_build_ms": host time
    spent capturing, 0 on a replay)."""
    n = model.nfree
    U0 = np.ascontiguousarray(np.asarray(U0, dtype=np.float64))
    V0 = np.ascontiguousarray(np.asarray(V0, dtype=np.float64))
    Fmag = np.ascontiguousarray(np.asarray(Fmag, dtype=np.float64))
    gvals = np.ascontiguousarray(np.asarray(gvals, dtype=np.float64))
    if U0.shape != (n,) or V0.shape != (n,) or Fmag.shape != (n,) or gvals.shape != (nsteps + 1,):
        raise ValueError("cd_loop: vector sizes do not match the free-dof block")
    probes = np.ascontiguousarray(np.asarray(probe_dofs, dtype=np.int64))
    stride = max(1, int(probe_stride))
    nrec = nsteps // stride + 1
    pout = np.zeros((nrec, len(probes)))
    U = np.zeros(n)
    V = np.zeros(n)
    p = _lib.CDParams()
    p.dt = float(dt)
    p.cdamp = float(cdamp)
    p.nsteps = int(nsteps)
    p.precision = int(precision)
    p.probe_stride = stride
    p.dt_mass = float(dt if dt_mass is None else dt_mass)
    p.use_graph = 1 if use_graph else 0
    ctx = model.ctx
    ctx.check(ctx.lib.fs_model_explicit(model.h, C.byref(p), _lib.ptr(U0), _lib.ptr(V0), _lib.ptr(Fmag),
                                        _lib.ptr(gvals), _lib.ptr(probes, C.c_int64) if len(probes) else None,
                                        int(len(probes)), _lib.ptr(pout) if len(probes) else None,
                                        _lib.ptr(U), _lib.ptr(V)), "fs_model_explicit")
    return {"U": U, "V": V, "probes": pout, "dt": dt, "nsteps": nsteps, "ms": model.timing(3),
            "build_ms": model.timing(4), "graph_build_ms": model.timing(6)}


def omega_max(femm, v0, maxit=30):
    """pwr (clamp_cyl_transient_examples.jl:45-60) on the device-resident FF K and M."""
    model = femm._model
    v0 = np.ascontiguousarray(np.asarray(v0, dtype=np.float64))
    out = C.c_double()
    model.ctx.check(model.ctx.lib.fs_model_omega_max(model.h, _lib.ptr(v0), int(maxit), C.byref(out)),
                    "fs_model_omega_max")
    return out.value


def stable_dt(omega, ksi=0.0, factor=0.9):
    """dt = factor * 2 / omega * (sqrt(1 + ksi^2) - ksi) (clamp_cyl_transient_examples.jl:218)."""
    return float(factor * 2 / omega * (np.sqrt(1 + ksi ** 2) - ksi))
