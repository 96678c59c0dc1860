"""The kielflow-style C++ driver over the reference-interface adaptors
(kielflow_amd/host) reproduces the Python/C-ABI path step for step."""
import os
import subprocess

import numpy as np
import pytest

from helpers import axes_for, grain_radius, rel_l2

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kielflow_amd", "host",
                      "kf_driver")


@pytest.mark.parametrize("kind,solver", [("eqst", "newton"), ("hyst", "newton"), ("eqst", "nlj")])
def test_driver_matches_c_abi_path(ctx, tmp_path, kind, solver):
    import kielflow_amd as kf
    n, steps, M, Q = 16, 3, 1.0, -0.1
    print("driver", kind, solver, flush=True)
    out = tmp_path / "state.bin"
    r = subprocess.run([DRIVER, str(n), str(steps), str(out), kind, solver, str(M), str(Q)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    nc = n * n * (n // 2 + 1)
    got = [raw[2 * a * nc:2 * (a + 1) * nc].view(np.complex128).reshape(n, n, n // 2 + 1) for a in range(4)]
    phi_d = raw[8 * nc:].reshape(n, n, n)
    # the same run through the Python binding of the C ABI
    axes, _ = axes_for(kind, n)
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=Q, a=grain_radius(kind, n))
    g = kf.Grid(ctx, axes)
    mode = {"newton": kf.POISSON_NEWTON_MG, "nlj": kf.POISSON_NLJ}[solver]
    ps = kf.Poisson(g, kf.KF_FP64, mode=mode, tol=1e-10, max_it=100 if solver == "newton" else 2000)
    rhs = kf.Rhs(g, ps, M=M, tau=0.1, theta=100.0, sv_on=False, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    F = [g.cfield() for _ in range(4)]
    fft = kf.FFT(g, kf.KF_FP64)
    fft.r2c(ctx.to_device(ln_n), F[3])
    rk = kf.Rk4(rhs, 1e-3)
    rk.step(F, steps)
    ref = [f.get() for f in F]
    for a in range(4):
        assert rel_l2(got[a], ref[a]) <= 1e-12, (a, rel_l2(got[a], ref[a]))
    assert rel_l2(phi_d, phi.get()) <= 1e-12
    for o in (rk, fft, rhs, ps, g):
        o.close()
