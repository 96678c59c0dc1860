"""kielflow's files on the GPU path: the stored form of a state
(kielflow.cpp:156-182, 290-327), save_all / save_slice / load (IO.cpp),
and the main program's loop with its output cadence run by the C++ driver
from an HDF5 input (kielflow.cpp:147-331) against the same loop through the
Python binding."""
import os
import subprocess

import numpy as np
import pytest

from helpers import axes_for, initial_state, rel_l2

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kielflow_amd", "host",
                      "kf_driver")


def _state(ctx, kind, n, seed):
    import kielflow_amd as kf
    S, nd, mask, og = initial_state(kind, n, seed=seed, Q=-0.2)
    rng = np.random.default_rng(seed + 1)
    for a in range(3):
        S[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
    axes, _ = axes_for(kind, n)
    g = kf.Grid(ctx, axes)
    F = [ctx.to_device(x) for x in S]
    phi = ctx.to_device(0.1 * rng.standard_normal(og.rshape))
    return g, F, phi, S, og


@pytest.mark.parametrize("kind", ["eqst", "sist"])
def test_state_file_round_trip(ctx, tmp_path, kind):
    import kielflow_amd as kf
    from kielflow_amd import h5io
    print("h5 round trip", kind, flush=True)
    n = 16
    g, F, phi, S, og = _state(ctx, kind, n, 3)
    params = {"M": 1.2, "tau": 0.1, "theta": 100.0, "mu": 0.0, "beta": 0.0}
    path = str(tmp_path / "state.h5")
    h5io.save_all(path, g, 0.5, params, F, phi)
    axes, t, p, fields = h5io.load_all(path)
    assert t == 0.5 and p == params and [a.n for a in axes] == [n] * 3
    # the stored fields: U and the linear density in real space, Ph as given
    for a in range(3):
        assert rel_l2(fields[a], og.c2r(S[a])) <= 1e-14
    assert rel_l2(fields[3], np.exp(og.c2r(S[3]))) <= 1e-14
    assert np.array_equal(fields[4], phi.get())
    # back to the spectral state with ln n
    G = [g.cfield() for _ in range(4)]
    phi2 = g.rfield()
    t2, p2 = h5io.load_state(path, g, G, phi2)
    assert (t2, p2) == (0.5, params)
    for a in range(4):
        assert rel_l2(G[a].get(), S[a]) <= 1e-13, a
    assert np.array_equal(phi2.get(), phi.get())
    # a slice is the z plane at index_at(0) of the same fields
    sp = h5io.save_slice(7, g, 0.5, params, F, phi, str(tmp_path))
    assert os.path.basename(sp) == "slice_0007.h5"
    k = kf.axis_index_at(g.axes[2], 0.0)
    assert k == n // 2
    for i, name in enumerate(h5io.FIELDS):
        assert np.array_equal(h5io.load_slice(name, sp), fields[i][:, :, k]), name
    # a non-positive density is refused on load
    bad = fields.copy()
    bad[3, 0, 0, 0] = 0.0
    with pytest.raises(kf.KfError, match="positive"):
        g.state_from_real(bad, G)
    for f in F + G + [phi, phi2]:
        f.free()
    g.close()


def _main_program_fields(axes):
    """nd and the barrier of kielflow.cpp:186-189, 227-229 on the grid."""
    import kielflow_amd as kf
    x = [np.array([kf.axis_values(a, i)[0] for i in range(a.n)]) for a in axes]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    a = 0.15
    Q = -11498.5 * 8.1720e-06
    nd = Q / (np.pi ** 1.5 * a ** 3) * np.exp(-(X / a) ** 2 - (Y / a) ** 2 - (Z / a) ** 2)
    bar = np.exp(-(X / 0.15) ** 2 - (Y / 0.15) ** 2 - (Z / 0.15) ** 2)
    return nd, bar


def test_driver_main_loop_from_a_file(ctx, tmp_path):
    import kielflow_amd as kf
    from kielflow_amd import h5io
    print("h5 driver main loop", flush=True)
    n, steps, dt = 16, 12, 0.01
    axes, _ = axes_for("eqst", n)
    g = kf.Grid(ctx, axes)
    S, _, _, og = initial_state("eqst", n, seed=9, amp=1e-2)
    F = [ctx.to_device(x) for x in S]
    phi = g.rfield()
    params = {"M": 1.0, "tau": 0.1, "theta": 100.0, "mu": 0.0, "beta": 0.0}
    inp = str(tmp_path / "fields.h5")
    h5io.save_all(inp, g, 0.25, params, F, phi)
    d_drv = tmp_path / "drv"
    d_py = tmp_path / "py"
    d_drv.mkdir()
    d_py.mkdir()
    r = subprocess.run([DRIVER, "file", inp, str(steps), str(d_drv), "newton", "5", "10", str(dt)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    # the same loop through the Python binding
    nd, bar = _main_program_fields(axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-10, max_it=100)
    rhs = kf.Rhs(g, ps, M=1.0, tau=0.1, theta=100.0, sv_on=True, pen_on=False)
    G = [g.cfield() for _ in range(4)]
    phi2 = g.rfield()
    t0, p0 = h5io.load_state(inp, g, G, phi2)
    rhs.set_fields(phi2, ctx.to_device(nd), ctx.to_device(bar))
    rk = kf.Rk4(rhs, dt)
    t, written = h5io.time_loop(rk, g, G, phi2, steps, t0, dt, p0, str(d_py), slice_every=5, backup_every=10)
    names = sorted(os.path.basename(w) for w in written)
    assert names == ["backup_at_10.h5", "slice_0005.h5", "slice_0010.h5"], names
    assert sorted(os.listdir(d_drv)) == names
    assert abs(t - (0.25 + steps * dt)) < 1e-15
    for name in names:
        a, b = str(d_drv / name), str(d_py / name)
        step = int(name.split("_")[-1].split(".")[0])
        # the time before the increment (kielflow.cpp:303, 324, 329)
        assert abs(h5io.load_time(a) - (0.25 + step * dt)) < 1e-12
        assert abs(h5io.load_time(b) - h5io.load_time(a)) < 1e-12
        assert h5io.load_parameters(a) == params
        for f in h5io.FIELDS:
            if name.startswith("slice"):
                x, y = h5io.load_slice(f, a), h5io.load_slice(f, b)
            else:
                x, y = h5io.load_field_real(a, f), h5io.load_field_real(b, f)
            if np.abs(y).max() == 0:
                assert np.abs(x).max() == 0, (name, f)
            else:
                assert rel_l2(x, y) <= 1e-11, (name, f, rel_l2(x, y))
    for o in (rk, rhs, ps, g):
        o.close()
