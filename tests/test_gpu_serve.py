"""Served requests: each result is bitwise the same alone and in a mix of
concurrently running requests (several slots, overlapping on the device), and
one served layered request agrees with the oracle."""
import numpy as np
import pytest

from flexshell import serve as S, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _requests(specs):
    return [S.Request(s["rid"], s["kind"], W.c5_model(s), steps=s["steps"], arrival=s["arrival"], seed=s["rid"])
            for s in specs]


def _same(a, b):
    if a.keys() != b.keys():
        return False
    for k in a:
        x, y = a[k], b[k]
        if hasattr(x, "indptr"):
            if not (np.array_equal(x.indptr, y.indptr) and np.array_equal(x.indices, y.indices)
                    and np.array_equal(x.data, y.data)):
                return False
        elif isinstance(x, np.ndarray):
            if not np.array_equal(x, y):
                return False
        elif k != "cd_ms" and x != y:
            return False
    return True


def test_served_results_identical_alone_and_in_mix(ctx):
    print("[gpu] serve: alone vs mix", flush=True)
    specs = W.served_c5(nreq=12, seed=6, rate=400.0, emin=1e3, emax=3e4, explicit_steps=40)
    c1 = S.request_context()
    alone = []
    for r in _requests(specs):
        alone.append(S.execute(r, c1))
    c1.close()
    mixed = _requests(specs)
    with S.ShellServer(nslots=4, hbm_budget=8 << 30) as srv:
        summary = S.replay(srv, mixed)
        assert summary["failed"] == 0, [r.error for r in mixed if r.error]
        assert srv.max_concurrent >= 2
    for r, a in zip(mixed, alone):
        assert _same(r.result, a), (r.rid, r.kind)
    print(f"[gpu] serve summary {summary}", flush=True)


def test_served_admission_respects_budget(ctx):
    print("[gpu] serve: admission by footprint", flush=True)
    specs = W.served_c5(nreq=6, seed=7, rate=1000.0, emin=2e3, emax=4e3, explicit_steps=10)
    reqs = _requests(specs)
    fp = sorted(r.footprint for r in reqs)
    budget = fp[-1]  # the largest request fits alone; pairs only where they fit together
    with S.ShellServer(nslots=3, hbm_budget=budget) as srv:
        summary = S.replay(srv, reqs)
        assert summary["failed"] == 0
        assert srv.max_inflight_bytes <= budget
    # a request larger than the whole budget still runs (alone)
    big = _requests(specs[:1])
    with S.ShellServer(nslots=2, hbm_budget=big[0].footprint // 2) as srv:
        summary = S.replay(srv, big)
        assert summary["failed"] == 0 and srv.max_concurrent == 1


def test_served_layered_request_matches_oracle(ctx):
    print("[gpu] serve: layered modal request vs oracle", flush=True)
    spec = dict(W.served_c5(nreq=3, seed=8)[0], kind="modal", nel=1500.0, nen=4, nplies=5,
                angles=[0.0, 45.0, -45.0, 90.0, 45.0, 0.0, 0.0, 0.0], steps=0)
    m = W.c5_model(spec)
    r = S.Request(0, "modal", m)
    res = S.execute(r, ctx)
    L = W.C5_LAMINA
    D6 = O.lamina_D6(L["E1"], L["E2"], L["nu12"], L["G12"], L["G12"], L["E2"] / 2.6)
    plies = [(D6, 0.125e-3, a, L["rho"]) for a in spec["angles"][:spec["nplies"]]]
    A, B, D, H, (md, mi) = O.laminate(plies)
    d = m["dchi"]
    nrm, val = O.csys_normals(4, m["xyz"], m["conn"], np.eye(3))
    Ke = O.element_matrices(4, m["xyz"], m["conn"], nrm, val, groups=[(A, B, D, H, np.eye(3), 5 * 0.125e-3)])
    nf = d.nfreedofs
    Kff = O.assemble_csr(4, m["conn"], d.dofnums, nf, Ke)
    K = res["K"]
    assert np.array_equal(K.indptr, Kff.indptr) and np.array_equal(K.indices, Kff.indices)
    assert np.linalg.norm(K.data - Kff.data) <= 1e-12 * np.linalg.norm(Kff.data)
    tm, rm = O.element_masses(4, m["xyz"], m["conn"], m0=md, m1=mi)
    Mo = O.lumped_diag(4, m["conn"], d.dofnums, nf, tm, rm)
    assert np.allclose(res["M"], Mo, rtol=1e-13, atol=0)


def test_served_large_requests_alone_and_in_mix(ctx):
    """Requests of 1e5 to 2.5e5 elements (layered T3 and Q4, assembly and
    explicit dynamics in FS_PREC_MIXED_F32) among small ones on three slots:
    each result bitwise the same as alone."""
    from flexshell import _lib
    print("[gpu] serve: large requests alone vs mix", flush=True)
    specs = W.served_c5(nreq=6, seed=12, rate=1000.0, emin=1e3, emax=3e3, explicit_steps=60)
    big = [dict(specs[0], rid=100, kind="assembly", nen=4, nel=2.5e5),
           dict(specs[1], rid=101, kind="explicit", nen=3, nel=1.2e5, steps=60),
           dict(specs[2], rid=102, kind="explicit", nen=4, nel=1.0e5, steps=60)]
    specs = [big[0]] + specs[:2] + [big[1]] + specs[2:] + [big[2]]
    for k, sp in enumerate(specs):
        sp["arrival"] = 0.002 * k

    def reqs():
        out = []
        for sp in specs:
            prec = _lib.FS_PREC_MIXED_F32 if sp["kind"] == "explicit" else _lib.FS_PREC_FP64
            out.append(S.Request(sp["rid"], sp["kind"], W.c5_model(sp), steps=sp["steps"], arrival=sp["arrival"],
                                 seed=sp["rid"], precision=prec))
        return out
    c1 = S.request_context()
    alone = [S.execute(r, c1) for r in reqs()]
    c1.close()
    assert max(a["elements"] for a in alone) >= 2.4e5
    mixed = reqs()
    with S.ShellServer(nslots=3, hbm_budget=64 << 30) as srv:
        summary = S.replay(srv, mixed)
        assert summary["failed"] == 0, [r.error for r in mixed if r.error]
        assert srv.max_concurrent >= 2
    for r, a in zip(mixed, alone):
        assert _same(r.result, a), (r.rid, r.kind)
