"""The allocation-free stable-dt entry point (fs_model_omega_max_noalloc),
called through the C ABI: bitwise the same omega_max as the allocating
fs_model_omega_max on two calls in a row, the second on a smaller model (the
per-device workspace keeps the first call's contents, which must not leak
into the second), and a clear refusal above the workspace's 2^20 rows."""
import ctypes as C

import numpy as np
import pytest

from flexshell import _lib, explicit, femm as FM, workloads as W

pytestmark = pytest.mark.gpu


def _plate(ctx, nx, ny):
    c = W.plate_c4(nx=nx, ny=ny)
    d = c["dchi"]
    f = FM.FEMMShellT3FF(c["conn"], c["material"], thickness=c["t"], ctx=ctx)
    f.associategeometry(c["xyz"])
    f.stiffness(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), c["xyz"], None, None, d)
    f.mass(FM.SysmatAssemblerHIP(_lib.FS_BLOCK_FF), c["xyz"], d)
    return f, d.nfreedofs


def _noalloc(f, v0, maxit):
    lib = _lib.load()
    out = C.c_double(0.0)
    st = lib.fs_model_omega_max_noalloc(f._model.h, _lib.ptr(v0), int(maxit), C.byref(out))
    return st, out.value


def test_omega_max_noalloc_matches_allocating(ctx):
    for nx, ny in ((96, 48), (24, 12)):  # second call on a smaller model
        print(f"[gpu] omega_max without allocation: plate {nx}x{ny}", flush=True)
        f, nf = _plate(ctx, nx, ny)
        v0 = np.random.default_rng(nx).random(nf)
        ref = explicit.omega_max(f, v0, maxit=30)
        st, om = _noalloc(f, v0, 30)
        assert st == _lib.FS_OK, ctx.lib.fs_last_error(ctx.h)
        assert om == ref, (om, ref)
        f._model.close()


def test_omega_max_noalloc_refuses_oversize(ctx):
    print("[gpu] omega_max without allocation: model above the workspace", flush=True)
    f, nf = _plate(ctx, 600, 300)
    assert nf > (1 << 20)
    st, _ = _noalloc(f, np.ones(nf), 5)
    assert st == -6  # FS_ERR_UNSUPPORTED
    assert b"exceed the workspace" in ctx.lib.fs_last_error(ctx.h)
    f._model.close()
