"""Benchmark of the shell hot path on one MI355X.

Workload (single large Q4RS model, FP64, plain form): cylindrical panel R 1 m,
L 2 m, 60 degree arc, t 0.005 m, E 70e9, nu 0.33, rho 2700, Q4 512 x 512
(262,144 elements, 263,169 nodes), interior nodes jittered +-10% h (seed 2),
straight edges simply supported.  Timed region: `--steps` FP64 assemblies of
the free-free stiffness into CSR (pattern and scatter plan prebuilt, inputs
resident in HBM); then the central-difference loop (`--cd-steps`, default
1000) at 0.9 x the stable dt from seeded Gaussian initial velocities.

Prints one JSON line.  value = FP64 elements assembled per second;
dof_steps_per_s reports the explicit loop.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "finetoolsflexstructures.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6  # MI355X FP64 vector (and FP64 MFMA), dense


def assembly_bytes(E, N, nen, nnz):
    """Algorithmic bytes of one FP64 assembly into CSR values (pattern and
    slot map prebuilt): conn + slot map + coords + normals + valid + values."""
    k = 6 * nen
    return E * (4 * nen + 4 * k * k) + N * (24 + 24 + 1) + nnz * 8


def cd_step_bytes(n, nnz, nnb, nnodes):
    """Algorithmic bytes of one fused FP64 CD step in the node-grouped form: K
    values, one 20-byte record per node with free dofs (at most nnodes), one
    8-byte run record per neighbour slot (nnb), u gathered once, U V A invM
    Fmag reads, U V A writes."""
    return nnz * 8 + nnodes * 20 + nnb * 8 + n * 72


def resultant_bytes(E, N, nen, npts):
    """Algorithmic HBM bytes of one stress-resultant launch over all elements:
    connectivity 4 nen E, per node coordinates 24 + normal and flag 25 + the
    six displacements 48, and 8 resultants of 8 B per integration point."""
    return 4 * nen * E + (24 + 25 + 48) * N + 64 * npts * E


def _omp_threads(n):
    """Set the OpenMP thread count of the oracle library (libgomp, shared with
    the oracle's C code); returns the count in effect."""
    import ctypes as C
    gomp = C.CDLL("libgomp.so.1")
    if n:
        gomp.omp_set_num_threads(int(n))
    return int(gomp.omp_get_max_threads())


def _cpu_assembly_rate(nen, xyz, conn, dofnums, nrows, material, t, nsample):
    """Elements/s of the oracle's element matrices + COO -> CSR assembly over
    the first nsample elements."""
    from oracle import oracle as O
    Dps, Dt = O.shell_D(O.iso_D6(material.E, material.nu))
    nrm, val = O.normals(nen, xyz, conn)
    sub = conn[:nsample]
    t0 = time.perf_counter()
    Ke = O.element_matrices(nen, xyz, sub, nrm, val, Dps=Dps, Dt=Dt, t=t)
    O.assemble_csr(nen, sub, dofnums, nrows, Ke)
    return nsample / (time.perf_counter() - t0)


def cpu_baseline(case, nsample, K_host, Mdiag, dt, cd_steps_cpu):
    """The oracle (C restatement of the reference's algorithm: dense
    transforms, COO -> CSR with duplicates summed in input order, row CSR
    products; OpenMP) timed on this host, on all cores and on one core:
    - the 32 x 32 plate in full: K and M assembled, static solve under 1e3 Pa
      (scipy spsolve) and the 10 lowest frequencies (scipy eigsh,
      shift-invert), as the reference's statics/dynamics tests drive them;
    - the 512 x 512 cylinder: element matrices + assembly of the first
      nsample elements (rate extrapolated to the whole model) and
      cd_steps_cpu central-difference steps on the full K;
    - the 1024 x 512 plate: element matrices + assembly of a 2 x nsample
      element sample."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from flexshell import workloads as W
    from oracle import oracle as O
    all_cores = _omp_threads(0)
    out = {"unit": "elements/s", "kind": "port", "cores": all_cores}
    n = case["dchi"].nfreedofs
    c1 = W.plate_c1(n=32)
    c4 = W.plate_c4()
    for label, nth in (("all_cores", all_cores), ("one_core", 1)):
        _omp_threads(nth)
        r = {"threads": nth}
        # C1 in full
        d1 = c1["dchi"]
        mat1 = c1["material"]
        t0 = time.perf_counter()
        Dps, Dt = O.shell_D(O.iso_D6(mat1.E, mat1.nu))
        nrm, val = O.normals(3, c1["xyz"], c1["conn"])
        Ke = O.element_matrices(3, c1["xyz"], c1["conn"], nrm, val, Dps=Dps, Dt=Dt, t=c1["t"])
        K1 = O.assemble_csr(3, c1["conn"], d1.dofnums, d1.nfreedofs, Ke)
        tm, rm = O.element_masses(3, c1["xyz"], c1["conn"], rho=mat1.rho, t=c1["t"])
        M1 = O.lumped_diag(3, c1["conn"], d1.dofnums, d1.nfreedofs, tm, rm)
        t_asm = time.perf_counter() - t0
        F1 = O.surface_load(3, c1["xyz"], c1["conn"], [0, 0, -1e3, 0, 0, 0], d1.nalldofs, d1.dofnums, 3)[
            :d1.nfreedofs]
        t0 = time.perf_counter()
        sla.spsolve(K1.tocsc(), F1)
        w = sla.eigsh(K1.tocsc(), k=10, M=sp.diags(M1).tocsc(), sigma=0.0, which="LM", return_eigenvectors=False)
        t_solve = time.perf_counter() - t0
        r["plate_32x32"] = {"elements": int(c1["conn"].shape[0]), "assembly_s": t_asm,
                            "elements_per_s": c1["conn"].shape[0] / t_asm,
                            "solve_and_10_frequencies_s": t_solve,
                            "f1_hz": float(np.sqrt(np.sort(w)[0]) / (2 * np.pi))}
        # C2 sample and steps
        ns = nsample if nth > 1 else max(2000, nsample // 8)
        r["cylinder_512_elements_per_s"] = _cpu_assembly_rate(4, case["xyz"], case["conn"], case["dchi"].dofnums,
                                                              n, case["material"], case["t"], ns)
        r["cylinder_512_sample_elements"] = ns
        steps = cd_steps_cpu if nth > 1 else max(2, cd_steps_cpu // 5)
        t0 = time.perf_counter()
        O.cd_loop(K_host, Mdiag, 0.0, dt, steps, np.zeros(n), np.zeros(steps + 1), np.zeros(n), np.zeros(n))
        r["cylinder_512_dof_steps_per_s"] = n * steps / (time.perf_counter() - t0)
        r["cylinder_512_cd_steps"] = steps
        # C4 sample
        ns4 = 2 * ns
        d4 = c4["dchi"]
        r["plate_1024x512_elements_per_s"] = _cpu_assembly_rate(3, c4["xyz"], c4["conn"], d4.dofnums, d4.nfreedofs,
                                                                c4["material"], c4["t"], ns4)
        r["plate_1024x512_sample_elements"] = ns4
        out[label] = r
    _omp_threads(all_cores)
    a = out["all_cores"]
    out["value"] = a["cylinder_512_elements_per_s"]
    out["dof_steps_per_s"] = a["cylinder_512_dof_steps_per_s"]
    out["sample"] = (f"oracle C restatement (OpenMP) on {all_cores} threads and on 1: the 32x32 plate in full "
                     f"(assembly, static solve, 10 frequencies); 512x512 cylinder element matrices + COO->CSR "
                     f"assembly of the first {nsample} elements ({out['one_core']['cylinder_512_sample_elements']} "
                     f"on 1 thread), rate extrapolated, and {cd_steps_cpu} CD steps on the full K "
                     f"({out['one_core']['cylinder_512_cd_steps']} on 1 thread); 1024x512 plate assembly of "
                     f"{2 * nsample} elements")
    return out


def time_assembly(ctx, model, params, steps, warmup):
    """FP64 assembly of the free-free K into CSR values (pattern and plans
    prebuilt, inputs resident in HBM), timed with HIP events on the library's
    stream: the fused tile form (the default) and, for reference, the
    two-kernel form (element blocks through HBM + pair scatter).  Returns
    per-form device times (ms) of the whole assembly and of its two kernels,
    and the wall time per fused assembly."""
    import numpy as np

    from flexshell import _lib

    def assemble():
        ctx.check(ctx.lib.fs_model_stiffness(model.h, _lib.C.byref(params), _lib.FS_BLOCK_FF, None),
                  "fs_model_stiffness")
    out = {}
    for mode, name in ((0, "two_kernel"), (1, "fused")):
        ctx.set_assembly(mode)
        for _ in range(max(1, warmup)):
            assemble()
        t = []
        t0 = time.perf_counter()
        for _ in range(steps):
            assemble()
            t.append((model.timing(0), model.timing(1), model.timing(2)))
        wall = (time.perf_counter() - t0) / steps * 1e3
        t = np.array(t)
        out[name] = {"device_ms": float(t[:, 2].mean()), "kernel1_ms": float(t[:, 0].mean()),
                     "kernel2_ms": float(t[:, 1].mean()), "wall_ms": wall}
    ctx.set_assembly(1)
    return out


def bench_assembly_plate(steps, warmup):
    """FP64 assembly of the 1024 x 512 T3 plate (1,048,576 elements): elements
    per second and the HBM roofline of the fused form."""
    from flexshell import _lib, femm as FM, workloads as W
    ctx = _lib.default_context()
    c = W.plate_c4()
    d = c["dchi"]
    f = FM.FEMMShellT3FF(c["conn"], c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(c["xyz"])
    f._ensure_dofs(d)
    nnz = len(f._pattern(_lib.FS_BLOCK_FF)[2])
    E, N = c["conn"].shape[0], c["xyz"].shape[0]
    t = time_assembly(ctx, f._model, f._shell_params(), steps, warmup)
    b = assembly_bytes(E, N, 3, nnz)
    dev = t["fused"]["device_ms"]
    ach = b / (dev / 1e3) / 1e9
    f._model.close()
    return {"workload": "plate_t3_1024x512 steel t0.004 clamped, FF-block K into CSR", "elements": E, "nodes": N,
            "nnz": nnz, "elements_per_s": E / (dev / 1e3), "device_ms": dev,
            "kernels_ms": {"fused_element_kernel": t["fused"]["kernel1_ms"],
                           "pair_finalize_kernel": t["fused"]["kernel2_ms"]},
            "two_kernel_ms": {"elem_kernel": t["two_kernel"]["kernel1_ms"],
                              "scatter_kernel": t["two_kernel"]["kernel2_ms"],
                              "total": t["two_kernel"]["device_ms"]},
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": b,
                         "kernel": "assembly: fused element kernel (T3, 32 elements per tile) + pair_finalize_kernel"}}


def bench_batch(nmodels, nsteps):
    """256 independent T3 plates (seeded): the plain path (one model per launch
    sequence, FP64, no graph) against the batched path: all models in one
    launch per step (hipGraph-captured loop), and the same in Infinity-Cache
    groups (batch.ShellBatchGroups: each group's 500 steps run before the
    next group's, so that its operator stays in the 256 MB cache).  All are
    timed with HIP events around the step loops; per-model dt from the device
    power iteration."""
    import numpy as np

    from flexshell import _lib, batch as B, explicit, femm as FM, workloads as W
    ctx = _lib.default_context()
    models = W.batch_c3(nmodels=nmodels)
    nm = len(models)
    V0s = [W.gaussian_velocity(m["dchi"], 1e-3, seed=100 + k) for k, m in enumerate(models)]
    zeros = [np.zeros(m["dchi"].nfreedofs) for m in models]
    g = np.ones((nsteps + 1, nm))
    dts, plain_cd_ms, plain_asm_ms, U_plain = [], 0.0, 0.0, {}
    keep = {0, nm // 2, nm - 1}
    for k, m in enumerate(models):
        f = FM.FEMMShellT3FF(m["conn"], m["material"], thickness=m["t"], ctx=ctx)
        f.associategeometry(m["xyz"])
        d = m["dchi"]
        nf = d.nfreedofs
        f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), m["xyz"], None, None, d)
        plain_asm_ms += f._model.timing(2)
        f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), m["xyz"], d)
        dts.append(W.converged_dt(f, nf, seed=k))
        r = explicit.run(f._model, 0.0, zeros[k], V0s[k], dts[k], nsteps, zeros[k], g[:, k])
        plain_cd_ms += r["ms"]
        if k in keep:
            U_plain[k] = r["U"]
        f._model.close()
    dts = np.array(dts)
    cd0 = np.zeros(nm)
    V0 = np.concatenate(V0s)
    bt = B.ShellBatch(models, ctx=ctx)
    bt.assemble_device()
    batch_asm_ms = bt.model.timing(2)
    bt.mass()
    nfree = bt.nfree
    E = int(sum(m["conn"].shape[0] for m in models))
    r1 = bt.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g, nsteps)
    bt.close()
    plain = nfree * nsteps / (plain_cd_ms / 1e3)
    one = nfree * nsteps / (r1["ms"] / 1e3)
    out = {"models": nm, "elements": E, "free_dofs": nfree, "steps": nsteps,
           "plain_dof_steps_per_s": plain, "one_group_dof_steps_per_s": one, "one_group_over_plain": one / plain}
    ref = None
    for name, prec, vb in (("batched", _lib.FS_PREC_FP64, 8), ("batched_mixed", _lib.FS_PREC_MIXED, 6),
                           ("batched_mixed_f32", _lib.FS_PREC_MIXED_F32, 6)):
        bg = B.ShellBatchGroups(models, ctx=ctx, value_bytes=vb)
        bg.assemble_device()
        bg.mass()
        if prec != _lib.FS_PREC_FP64:
            bg.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g[:11], 10, precision=prec)  # operators
        bg.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g[:21], 20, precision=prec)  # warm
        rm = bg.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g, nsteps, precision=prec)
        rate = nfree * nsteps / (rm["ms"] / 1e3)
        out.update({f"{name}_dof_steps_per_s": rate, f"{name}_over_plain": rate / plain,
                    f"{name}_groups": len(bg.groups)})
        if ref is None:
            ref = rm
            out["max_abs_final_U"] = float(np.abs(rm["U"]).max())  # every model's dt is stable
            out["bitwise_equal_to_plain"] = bool(all(np.array_equal(rm["U"][bg.model_slice(k)], U_plain[k])
                                                     for k in keep))
            out["bitwise_equal_to_one_group"] = bool(np.array_equal(rm["U"], r1["U"]))
        else:
            errs = [float(np.abs(rm["U"][bg.model_slice(k)] - ref["U"][bg.model_slice(k)]).max()
                          / np.abs(ref["U"][bg.model_slice(k)]).max()) for k in range(nm)]
            out[f"{name}_over_batched"] = rate / out["batched_dof_steps_per_s"]
            out[f"{name}_max_rel_err_final_U"] = max(errs)
        bg.close()
    # wall clock, two contexts: groups of half the size dealt to two streams,
    # so that one group's host work and launch gaps overlap another's steps
    for name, prec, vb in (("batched", _lib.FS_PREC_FP64, 8), ("batched_mixed_f32", _lib.FS_PREC_MIXED_F32, 6)):
        bg = B.ShellBatchGroups(models, ctx=ctx, group_bytes=B.CACHE_GROUP_BYTES // 2, value_bytes=vb, streams=2)
        bg.assemble_device()
        bg.mass()
        bg.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g, nsteps, precision=prec)  # operators, graphs
        rm = bg.explicit(dts, cd0, np.zeros(nfree), V0, np.zeros(nfree), g, nsteps, precision=prec)
        rate = nfree * nsteps / (rm["wall_ms"] / 1e3)
        out.update({f"{name}_2streams_wall_dof_steps_per_s": rate, f"{name}_2streams_wall_over_plain": rate / plain})
        if prec == _lib.FS_PREC_FP64:
            out["bitwise_equal_2streams"] = bool(np.array_equal(rm["U"], r1["U"]))
        bg.close()
    return {**out,
            "plain_elements_per_s": E / (plain_asm_ms / 1e3), "batched_elements_per_s": E / (batch_asm_ms / 1e3),
            "note": "plain = one model per launch sequence (FP64, no graph); one_group = all models per launch, "
                    "hipGraph step loop, FP64, the operator streamed from HBM every step; batched = the same in "
                    "consecutive groups of at most 224 MB of operator and state (8-byte values; 6-byte for the "
                    "mixed forms), each group's steps run before the next, so the operator is read from the "
                    "256 MB Infinity Cache; batched_mixed = membrane+bending stiffness stored as FP32 + int16 "
                    "residual (FP64 products), FP64 shear factor rows, drilling, mass, state; batched_mixed_f32 = "
                    "the shear in the slices too; error = max over models of max|dU| / max|U_fp64| after the "
                    "steps (Gaussian initial velocities); *_2streams_wall = groups of half the size on two contexts "
                    "run concurrently, rate by host wall clock of the whole step loop (replay of the captured "
                    "graphs)"}


def mixed_step_bytes(N, E, nen, kmb_pairs, nsr, nnb, fmag_frac):
    """Algorithmic bytes of one mixed-precision step: the slices (FP32 value
    and int16 residual, 6 bytes per stored value), the
    neighbour lists, FP64 shear factor rows with their incidence records and
    connectivity (nsr > 0 only), drilling blocks, and the FP64 state over 6 N
    positions (u gathered once; U V A invMC read, Fmag
    read on the steps with a nonzero load factor (fraction fmag_frac), U V A
    written)."""
    n6 = 6 * N
    shear = (E * nsr * 6 * nen * 8 + E * nen * 4 * 2 + (N + 1) * 4) if nsr > 0 else 0
    return (kmb_pairs * 12 + nnb * 4 + (N + 1) * 4 + shear + N * 48
            + n6 * (8 + 32 + 8 * fmag_frac + 24))


def bench_mixed(nsteps, nx=1024, ny=512):
    """Flat steel plate 2 x 1 m, t 0.004, T3 nx x ny, clamped, Gaussian
    pressure pulse (sigma 0.1 m, 1e4 Pa peak, centre seeded, 1e-4 s): nsteps
    central-difference steps at 0.9 of the stable dt in FP64 (CSR) and in the
    two mixed precisions (FP64 shear factors; shear in the FP32 slices);
    trajectory error of the mixed runs against the FP64 run (which the tests
    pin to the oracle) over 64 probe dofs every 10 steps."""
    import numpy as np

    from flexshell import _lib, explicit, femm as FM, workloads as W
    t0 = time.perf_counter()
    ctx = _lib.default_context()
    c = W.plate_c4(nx=nx, ny=ny)
    d = c["dchi"]
    x0 = c["x0"]
    F = W.pressure_load_t3(c["xyz"], c["conn"], d,
                           lambda X: 1e4 * np.exp(-np.sum((X - x0) ** 2, axis=1) / (2 * 0.1 ** 2)))
    f = FM.FEMMShellT3FF(c["conn"], c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(c["xyz"])
    nf = d.nfreedofs
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), c["xyz"], None, None, d)
    f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), c["xyz"], d)
    dt = W.converged_dt(f, nf, seed=7)
    g = (explicit.sample_times(nsteps, dt) <= 1e-4).astype(float)
    wd = d.dofnums[:, 2][d.dofnums[:, 2] < nf]
    probes = wd[np.linspace(0, len(wd) - 1, 64).astype(int)]
    z = np.zeros(nf)
    print(f"[bench] mixed: setup {time.perf_counter() - t0:.1f}s, {nf} free dofs, dt {dt:.3e}", file=sys.stderr,
          flush=True)
    runs, info = {}, {}
    for name, prec in (("fp64", _lib.FS_PREC_FP64), ("mixed", _lib.FS_PREC_MIXED),
                       ("mixed_f32", _lib.FS_PREC_MIXED_F32)):
        explicit.run(f._model, 0.0, z, z, dt, 10, F, g[:11], precision=prec)  # warm (and build the operator)
        r = explicit.run(f._model, 0.0, z, z, dt, nsteps, F, g, probes, 10, precision=prec)
        runs[name] = r
        info[name] = f._model.mixed_info()
        print(f"[bench] mixed: {name} {r['ms'] / nsteps * 1e3:.1f} us/step", file=sys.stderr, flush=True)
    ref = runs["fp64"]["probes"]
    E, N = c["conn"].shape[0], c["xyz"].shape[0]
    out = {"workload": f"plate_t3_{nx}x{ny} steel t0.004 clamped, Gaussian pressure pulse", "free_dofs": nf,
           "steps": nsteps, "dt": dt}
    for name, r in runs.items():
        out[f"{name}_dof_steps_per_s"] = nf * nsteps / (r["ms"] / 1e3)
        out[f"{name}_us_per_step"] = r["ms"] / nsteps * 1e3
        if name != "fp64":
            out[f"{name}_max_rel_traj_err"] = float(np.abs(r["probes"] - ref).max() / np.abs(ref).max())
    for name in ("mixed", "mixed_f32"):
        out[f"{name}_over_fp64"] = out[f"{name}_dof_steps_per_s"] / out["fp64_dof_steps_per_s"]
        kmb_pairs, nsr, nnb, _ = info[name]
        b_mx = mixed_step_bytes(N, E, 3, kmb_pairs, nsr, nnb, float(np.mean(g[1:] != 0)))
        ach = b_mx / (runs[name]["ms"] / nsteps / 1e3) / 1e9
        out[f"roofline_{name}"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": ach / HBM_PEAK_GBS,
                                   "kernel": f"cd_mixed_step_kernel<3, {nsr}>",
                                   "algorithmic_bytes": b_mx, "bytes_per_dof": b_mx / nf}
    f._model.close()
    return out


def bench_serve(nreq, rate, slots, plain=True, steps=500, plain_requests=300, saturation_requests=300):
    """Served layered-shell workload (seed 6): Poisson arrivals at `rate`
    requests/s of assembly / assembly + modal matrices / explicit requests on
    T3FFComp or Q4RSComp plates of 1e3..5e5 elements (log-uniform), 3..8
    orthotropic plies; explicit requests step in FS_PREC_MIXED_F32 (48-bit
    stiffness values, FP64 products and state).  The request stream is
    replayed through the server with `slots` concurrent slots (continuous
    batching, admission by HBM footprint) and, for comparison, its first
    `plain_requests` requests through one slot (the plain GPU path: one
    request at a time in arrival order; when the offered rate exceeds what
    one slot sustains its queue grows, so its latencies depend on the prefix
    length).  Saturation: the first `saturation_requests` requests all
    offered at once, so the server runs at capacity; its throughput is the
    sustainable rate.  Meshes are generated before the clock starts; latency
    is arrival -> result on the host."""
    import numpy as np

    from flexshell import _lib, serve as S, workloads as W
    t0 = time.perf_counter()
    specs = W.served_c5(nreq=nreq, seed=6, rate=rate, explicit_steps=steps)
    models = [W.c5_model(s) for s in specs]
    print(f"[bench] serve: {nreq} requests generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    prec = _lib.FS_PREC_MIXED_F32

    def stream(at_once=False):
        return [S.Request(s["rid"], s["kind"], m, steps=s["steps"], arrival=0.0 if at_once else s["arrival"],
                          seed=s["rid"], precision=prec)
                for s, m in zip(specs, models)]

    out = {"requests": nreq, "rate_req_s": rate, "explicit_steps": steps, "explicit_precision": "mixed_f32",
           "elements_total": int(sum(m["conn"].shape[0] for m in models)),
           "kinds": {k: sum(s["kind"] == k for s in specs) for k in ("assembly", "modal", "explicit")}}
    runs = [("served", slots, nreq, False)]
    if saturation_requests:
        runs.append(("saturation", slots, min(saturation_requests, nreq), True))
    if plain:
        runs.append(("plain", 1, min(plain_requests, nreq), False))
    for name, ns, nr, at_once in runs:
        reqs = stream(at_once)[:nr]
        with S.ShellServer(nslots=ns) as srv:
            # warm the kernels of both element families on every slot before the clock starts
            warm = [S.Request(-1, "modal", W.c5_model(dict(specs[0], nel=1000.0, nen=nen, kind="modal")))
                    for nen in (3, 4) for _ in range(ns)]
            S.replay(srv, warm)
            summ = S.replay(srv, reqs)
        out[name] = summ
        if name == "served":
            served_reqs = reqs
        print(f"[bench] serve {name}: {summ}", file=sys.stderr, flush=True)
        if summ["failed"]:
            print("[bench] serve failures: " + "; ".join(r.error for r in reqs if r.error)[:2000], file=sys.stderr)
    if saturation_requests:
        out["saturation_req_s"] = out["saturation"]["throughput_req_s"]
        out["saturation_elements_per_s"] = out["saturation"]["elements_per_s"]
    if plain:
        # served latencies of the same prefix, for a like-for-like ratio
        pre = [r for r in served_reqs if r.rid < out["plain"]["requests"]]
        lat = np.array([r.latency for r in pre]) * 1e3
        out["served_prefix_p50_ms"] = float(np.percentile(lat, 50))
        out["served_prefix_p99_ms"] = float(np.percentile(lat, 99))
        out["plain_over_served_p50"] = out["plain"]["p50_ms"] / out["served_prefix_p50_ms"]
        out["plain_over_served_p99"] = out["plain"]["p99_ms"] / out["served_prefix_p99_ms"]
    return out


def copy_probe(nbytes):
    """One device-to-device copy of nbytes (a plain copy kernel of known size):
    the byte counters of this dispatch calibrate the profiled traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2 * nbytes


def hip_copy_GBs(nbytes=2 << 30, reps=5):
    """Device-to-device hipMemcpyAsync of nbytes (read + write), timed with HIP
    events on the null stream through the HIP runtime itself."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()

    def ok(status, what):
        if status != 0:
            raise RuntimeError(f"{what} returned {status}")

    try:
        ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)), "hipMalloc")
        ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)), "hipMalloc")
        ok(hip.hipMemset(src, 1, C.c_size_t(nbytes)), "hipMemset")
        ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        best = None
        for _ in range(reps + 1):
            ok(hip.hipEventRecord(e0, None), "hipEventRecord")
            ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None), "hipMemcpyAsync")  # device to device
            ok(hip.hipEventRecord(e1, None), "hipEventRecord")
            ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
            ms = C.c_float()
            ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
            best = ms.value if best is None else min(best, ms.value)
        return 2 * nbytes / (best / 1e3) / 1e9
    finally:
        for ev in (e0, e1):
            if ev.value:
                hip.hipEventDestroy(ev)
        for ptr in (src, dst):
            if ptr.value:
                hip.hipFree(ptr)


def torch_stream_GBs(nbytes, reps):
    """Copy (read + write) and sum (read) of nbytes of FP64 on torch's stream,
    timed with torch's events."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in (("copy_GBs", lambda: b.copy_(a), 2 * nbytes), ("read_GBs", lambda: a.sum(), nbytes)):
        fn()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = None
        for _ in range(reps):
            ev0.record()
            fn()
            ev1.record()
            ev1.synchronize()
            ms = ev0.elapsed_time(ev1)
            best = ms if best is None else min(best, ms)
        out[name] = moved / (best / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return out


def hbm_measured(nbytes=2 << 30, reps=5):
    """What plain streaming kernels reach on this device: a device-to-device
    copy (read + write) and a sum (read) of nbytes of FP64 (torch_stream_GBs),
    a practical ceiling beside the published 8 TB/s.  torch runs in a child
    process: it ships its own HIP runtime, and starting it in this process,
    after libflexshell has initialised the system one, has failed on some
    machines and left the process without a device.  If the child fails, the
    copy is timed here through the HIP runtime (hip_copy_GBs, smaller if the
    allocation fails) and the read rate is not measured; if that fails too,
    both are null."""
    import subprocess
    code = f"import json, bench; print(json.dumps(bench.torch_stream_GBs({int(nbytes)}, {int(reps)})))"
    try:
        r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=600)
        if r.returncode == 0:
            out = json.loads(r.stdout.strip().splitlines()[-1])
            out["source"] = "torch copy_ and sum in a child process"
            return out
        print(f"[bench] streaming kernels: torch child failed ({r.stderr.strip().splitlines()[-1:]}); HIP copy only",
              file=sys.stderr, flush=True)
    except Exception as exc:
        print(f"[bench] streaming kernels: torch child failed ({exc}); HIP copy only", file=sys.stderr, flush=True)
    for nb in (nbytes, nbytes // 8):
        try:
            return {"copy_GBs": hip_copy_GBs(nb, reps), "read_GBs": None, "source": f"hipMemcpyAsync of {nb} bytes"}
        except Exception as exc:
            print(f"[bench] streaming kernels: HIP copy of {nb} bytes failed ({exc})", file=sys.stderr, flush=True)
    return {"copy_GBs": None, "read_GBs": None, "source": "not measured"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cd-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plate-assembly", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-models", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=5000)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--no-serve-plain", action="store_true")
    ap.add_argument("--pipeline", type=int, default=1,
                    help="element ranges of the assembly pipeline (1: element kernel, then scatter)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: 3 assemblies, 50 CD steps, 50 steps of each explicit "
                         "precision on the plate, then a 1 GiB device copy that calibrates the byte counters")
    args = ap.parse_args()
    if args.profile:
        args.steps, args.warmup, args.cd_steps, args.no_cpu, args.no_batch = 3, 1, 50, True, True
        args.no_serve, args.mixed_steps = True, 50
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shard design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from flexshell import _lib, explicit, femm as FM, workloads as W

    t_setup = time.perf_counter()
    case = W.cylinder_panel_c2(n=args.n)
    d = case["dchi"]
    ctx = _lib.default_context()
    f = FM.FEMMShellQ4RS(case["conn"], case["material"], thickness=case["t"], ctx=ctx)
    f.associategeometry(case["xyz"])
    nf = d.nfreedofs
    f._ensure_dofs(d)
    nrows, rp, cv = f._pattern(_lib.FS_BLOCK_FF)
    nnz = len(cv)
    params = f._shell_params()
    model = f._model
    E, N = case["conn"].shape[0], case["xyz"].shape[0]
    print(f"[bench] {case['name']}: E={E} N={N} nfree={nf} nnz={nnz} setup {time.perf_counter() - t_setup:.1f}s",
          file=sys.stderr, flush=True)

    ctx.set_pipeline(args.pipeline)
    ta = time_assembly(ctx, model, params, args.steps, args.warmup)
    dev_ms = ta["fused"]["device_ms"]
    ms_per_step = ta["fused"]["wall_ms"]
    elem_ms = [ta["fused"]["kernel1_ms"]]
    print(f"[bench] assembly (fused tiles): device {dev_ms:.3f} ms = element kernel {ta['fused']['kernel1_ms']:.3f} + "
          f"finalize {ta['fused']['kernel2_ms']:.3f}; wall {ms_per_step:.3f} ms; two-kernel form "
          f"{ta['two_kernel']['device_ms']:.3f} ms", file=sys.stderr, flush=True)

    # explicit loop on the resident K and M
    Kd = f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), case["xyz"], None, None, d)
    Mdiag = f.mass(FM.SysmatAssemblerFFBlock(FM.SysmatAssemblerSparseDiag(), nf, nf), case["xyz"], d).diagonal()
    rng = np.random.default_rng(3)
    dt = W.converged_dt(f, nf, seed=3)
    om = 0.9 * 2 / dt
    V0 = W.gaussian_velocity(d, 1e-3, seed=3)
    ncd = args.cd_steps
    explicit.run(model, 0.0, np.zeros(nf), V0, dt, min(ncd, 20), np.zeros(nf), np.zeros(min(ncd, 20) + 1))
    t0 = time.perf_counter()
    r = explicit.run(model, 0.0, np.zeros(nf), V0, dt, ncd, np.zeros(nf), np.zeros(ncd + 1), [0], 100)
    cd_wall = time.perf_counter() - t0
    cd_ms = r["ms"]
    dof_steps = nf * ncd / (cd_ms / 1e3)
    print(f"[bench] CD: {ncd} steps, device {cd_ms:.2f} ms ({cd_ms / ncd * 1e3:.1f} us/step), wall {cd_wall:.3f}s, "
          f"omega_max {om:.4e}, dt {dt:.3e}", file=sys.stderr, flush=True)
    assert np.all(np.isfinite(r["U"])), "non-finite state in the explicit loop"

    # stress resultants (inspectintegpoints, all three quantities) of the
    # final CD state at the 4 Gauss points of every element
    Ufull = d.scattersysvec(r["U"])
    for _ in range(2):
        f.resultants(case["xyz"], Ufull)
    res_ms = []
    t0 = time.perf_counter()
    nres = 5
    for _ in range(nres):
        f.resultants(case["xyz"], Ufull)
        res_ms.append(model.timing(5))
    res_wall = (time.perf_counter() - t0) / nres * 1e3
    res_dev = float(np.mean(res_ms))
    b_res = resultant_bytes(E, N, 4, 4)
    ach_res = b_res / (res_dev / 1e3) / 1e9
    print(f"[bench] resultants: kernel {res_dev * 1e3:.1f} us, call {res_wall:.1f} ms (copies of u, the results and the point "
          f"locations)", file=sys.stderr, flush=True)

    b_asm = assembly_bytes(E, N, 4, nnz)
    b_cd = cd_step_bytes(nf, nnz, model.mixed_info()[2], N)
    ach = b_asm / (dev_ms / 1e3) / 1e9
    ach_cd = b_cd / (cd_ms / ncd / 1e3) / 1e9
    # HBM bytes per assembly (fused element kernel + finalize) from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    # (512 x 512 only), scaled by the copy-kernel calibration
    # and the FP64 operations the fused element kernel issues per launch
    # (SQ_INSTS_VALU_*_F64 of the same summary), against its live duration
    traffic, elem_flops = None, None
    tf = os.path.join(ROOT, "profiles", "traffic_q4_cylinder_512x512.json")
    if args.n == 512 and os.path.exists(tf):
        tj = json.load(open(tf))
        traffic = tj.get("fused_assembly_bytes_per_launch")
        elem_flops = tj.get("fused_elem_kernel", {}).get("fp64_ops_issued")
    elem_s = ta["fused"]["kernel1_ms"] / 1e3

    out = {
        "metric": "FP64 shell elements assembled/s + DOF-steps/s; p50/p99 latency vs plain GPU",
        "value": E / (dev_ms / 1e3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh jitter and initial velocities)",
        "config": {"workload": case["name"] + " Q4RS cylindrical panel R1 L2 60deg t0.005, FF-block K into CSR "
                                              "+ central-difference loop",
                   "elements": E, "nodes": N, "free_dofs": nf, "nnz": nnz, "cd_steps": ncd},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "assembly: fused element kernel (Q4, 12 elements per tile) + pair_finalize_kernel, "
                               "HIP events on the library stream around both",
                     "algorithmic_bytes": b_asm, "device_ms": dev_ms},
        "dof_steps_per_s": dof_steps,
        "cd_us_per_step": cd_ms / ncd * 1e3,
        "roofline_cd": {"bound": "hbm", "achieved": ach_cd, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach_cd / HBM_PEAK_GBS, "kernel": "cd_node_step_kernel (8 lanes per node)",
                        "algorithmic_bytes": b_cd},
        "roofline_elem": (None if elem_flops is None else
                          {"bound": "fp64-valu", "achieved": elem_flops / elem_s / 1e12, "peak": FP64_PEAK_TFS,
                           "unit": "TFLOP/s", "frac": elem_flops / elem_s / 1e12 / FP64_PEAK_TFS,
                           "kernel": "fused Q4 element kernel (FP64 operations issued, from the committed "
                                     "counters, over its live duration)",
                           "fp64_ops_per_launch": elem_flops}),
        "kernels_ms": {"fused_element_kernel": ta["fused"]["kernel1_ms"],
                       "pair_finalize_kernel": ta["fused"]["kernel2_ms"],
                       "two_kernel_form": {"elem_kernel": ta["two_kernel"]["kernel1_ms"],
                                           "scatter_kernel": ta["two_kernel"]["kernel2_ms"],
                                           "total": ta["two_kernel"]["device_ms"]}},
        "elem_launch_shape": dict(zip(("variant", "elements_per_workgroup", "workgroup"), ctx.tuning(4, 0))),
        "resultants": {"kernel_us": res_dev * 1e3, "points_per_s": 4 * E / (res_dev / 1e3),
                       "call_ms": res_wall,
                       "roofline": {"bound": "hbm", "achieved": ach_res, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": ach_res / HBM_PEAK_GBS, "kernel": "resultants_kernel<4, false, false>",
                                    "algorithmic_bytes": b_res}},
    }
    if not args.no_plate_assembly:
        t0 = time.perf_counter()
        out["assembly_plate_t3_1024x512"] = bench_assembly_plate(args.steps, args.warmup)
        print(f"[bench] plate assembly section {time.perf_counter() - t0:.1f}s: {out['assembly_plate_t3_1024x512']}",
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        import scipy.sparse as sp
        Kh = sp.csr_matrix((np.zeros(nnz), cv, rp), shape=(nf, nf))
        Kh.data = f._model.stiffness(params, _lib.FS_BLOCK_FF, nnz)
        out["cpu_baseline"] = cpu_baseline(case, min(args.cpu_sample, E), Kh, Mdiag, dt, 20)
    if not args.no_batch:
        t0 = time.perf_counter()
        out["batch"] = bench_batch(args.batch_models, args.batch_steps)
        print(f"[bench] batch section {time.perf_counter() - t0:.1f}s: {out['batch']}", file=sys.stderr, flush=True)
    if not args.no_mixed:
        t0 = time.perf_counter()
        out["mixed"] = bench_mixed(args.mixed_steps)
        print(f"[bench] mixed section {time.perf_counter() - t0:.1f}s: {out['mixed']}", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["serve"] = bench_serve(args.serve_requests, args.serve_rate, args.serve_slots,
                                   plain=not args.no_serve_plain, steps=args.serve_steps)
    if args.profile:
        out["copy_probe_bytes"] = copy_probe(1 << 30)
    else:
        hm = hbm_measured()
        out["hbm_measured"] = hm
        if hm["copy_GBs"]:
            out["roofline"]["frac_of_copy"] = ach / hm["copy_GBs"]
        if hm["read_GBs"]:
            out["roofline_cd"]["frac_of_read"] = ach_cd / hm["read_GBs"]
        print(f"[bench] streaming kernels: copy {hm['copy_GBs']} GB/s, read {hm['read_GBs']} GB/s",
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
