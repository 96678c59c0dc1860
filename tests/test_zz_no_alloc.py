"""kf_cfl_all_noalloc (the allocation-free variant of kf_cfl_all, with its
scratch in a static device array) gives kf_cfl_all's result through the C
ABI, on two calls in a row, the second on a smaller grid, and rejects a
grid whose scratch exceeds the static array."""
import ctypes as C

import numpy as np
import pytest

from helpers import axes_for

pytestmark = pytest.mark.gpu


def _velocities(kf, ctx, g, seed, amp):
    rng = np.random.default_rng(seed)
    fft = kf.FFT(g, kf.KF_FP64)
    F = []
    for _ in range(3):
        u = ctx.to_device(amp * rng.standard_normal(g.rshape))
        Fu = g.cfield()
        fft.r2c(u, Fu)
        F.append(Fu)
    fft.close()
    return F


def _both(kf, ctx, g, F, dt):
    lib = ctx.lib
    a, b = C.c_double(0), C.c_double(0)
    kf.api.check(lib.kf_cfl_all(g.h, dt, F[0].ptr, F[1].ptr, F[2].ptr, C.byref(a)))
    kf.api.check(lib.kf_cfl_all_noalloc(g.h, dt, F[0].ptr, F[1].ptr, F[2].ptr, C.byref(b)))
    return a.value, b.value


def test_cfl_noalloc_matches_cfl_all():
    import kielflow_amd as kf
    ctx = kf.Context(0)
    print("cfl noalloc: 96^3 sist then 48^3 hyst", flush=True)
    results = []
    for kind, n, seed in (("sist", 96, 1), ("hyst", 48, 2)):
        axes, _ = axes_for(kind, n)
        g = kf.Grid(ctx, axes)
        F = _velocities(kf, ctx, g, seed, 20.0)
        a, b = _both(kf, ctx, g, F, 1e-3)
        print(f"  {kind} {n}: kf_cfl_all {a:.17e} noalloc {b:.17e}", flush=True)
        assert a < 0.01                   # set by the velocities, not by the cap
        assert b == pytest.approx(a, rel=1e-13, abs=0)
        results.append(b)
        # the spacing table is formed on the device (kf_cfl_all forms it on
        # the host), so the two agree to rounding: rel 1e-13 (FP64 parity
        # bar 1e-11), on an equidistant grid as well
        axes_e, _ = axes_for("eqst", n)
        ge = kf.Grid(ctx, axes_e)
        Fe = _velocities(kf, ctx, ge, seed + 10, 20.0)
        ae, be = _both(kf, ctx, ge, Fe, 1e-3)
        assert ae < 0.01
        assert be == pytest.approx(ae, rel=1e-13, abs=0)
        for f in F + Fe:
            f.free()
        g.close()
        ge.close()
    assert results[0] != results[1]
    # a grid whose scratch exceeds the static array is refused
    big = kf.Grid(ctx, kf.cube(192, 16.0))
    z = big.cfield()
    out = C.c_double(0)
    code = ctx.lib.kf_cfl_all_noalloc(big.h, 1e-3, z.ptr, z.ptr, z.ptr, C.byref(out))
    assert code == kf.api.KF_ERR_ARG
    assert b"static scratch" in ctx.lib.kf_last_error()
    z.free()
    big.close()
    ctx.close()
