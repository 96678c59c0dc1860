"""Every choice of the FFT plan's autotuner (the column-block variant of the
x and y passes, padded or unpadded scratch rows) gives the same numbers: an
RHS evaluation (rhs_standard::solve) with the defaults (KF_FFT_AUTOTUNE=0)
and with the other end of every choice (KF_FFT_AUTOTUNE=alt) agree bit for
bit, the potential included."""
import os

import numpy as np
import pytest

from helpers import initial_state

pytestmark = pytest.mark.gpu


def _eval(ctx, kind, n, prec, general, setting, env="KF_FFT_AUTOTUNE"):
    import kielflow_amd as kf
    old = os.environ.get(env)
    if setting is None:
        os.environ.pop(env, None)
    else:
        os.environ[env] = setting
    try:
        state, nd, mask, og = initial_state(kind, n, seed=3, Q=-0.2)
        rng = np.random.default_rng(7)
        for a in range(3):
            state[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
        g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
        p = kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED
        ps = kf.Poisson(g, p, tol=1e-12 if prec == "fp64" else 1e-10, max_it=80)
        rhs = kf.Rhs(g, ps, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True, precision=p)
        if general:
            rhs.force_general(True)
        phi = g.rfield()
        rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
        dF = [ctx.to_device(s) for s in state]
        rhs.solve(0.0, dF)
        assert rhs.poisson_failures() == 0
        return [f.get() for f in dF] + [phi.get()]
    finally:
        if old is None:
            os.environ.pop(env, None)
        else:
            os.environ[env] = old


@pytest.mark.parametrize("kind,n,prec,general", [
    ("eqst", 24, "fp64", False),
    ("eqst", 24, "fp64", True),
    ("eqst", 32, "mixed", False),
    ("eqst", 48, "mixed", False),
    ("eqst", 96, "mixed", False),
    ("sist", 16, "fp64", False),
])
def test_autotuner_choices_give_identical_results(ctx, kind, n, prec, general):
    print("fft variants", kind, n, prec, general, flush=True)
    a = _eval(ctx, kind, n, prec, general, "0")
    b = _eval(ctx, kind, n, prec, general, "alt")
    names = ["FUx", "FUy", "FUz", "Fn", "phi"]
    for name, x, y in zip(names, a, b):
        assert np.array_equal(x, y), (name, float(np.abs(x - y).max()))


@pytest.mark.parametrize("n,prec", [(24, "fp64"), (32, "fp64"), (48, "mixed"), (96, "mixed"), (64, "fp64")])
def test_fused_penalization_pass_gives_identical_results(ctx, n, prec):
    """The products, penalization blend and forward z transform of a field in
    one z pass give the numbers of the separate passes (KF_PEN_UNFUSED) bit
    for bit, on radix-2 and radix-3 lengths."""
    print("fused penalization", n, prec, flush=True)
    a = _eval(ctx, "eqst", n, prec, False, None, env="KF_PEN_UNFUSED")
    b = _eval(ctx, "eqst", n, prec, False, "1", env="KF_PEN_UNFUSED")
    names = ["FUx", "FUy", "FUz", "Fn", "phi"]
    for name, x, y in zip(names, a, b):
        assert np.array_equal(x, y), (name, float(np.abs(x - y).max()))


@pytest.mark.parametrize("kind,n,prec", [("sist", 16, "fp64"), ("sist", 24, "mixed"), ("hyst", 32, "fp64"),
                                         ("hyst", 48, "mixed"), ("sist", 64, "fp64")])
def test_fused_metric_pass_gives_identical_results(ctx, monkeypatch, kind, n, prec):
    """A derivative along a stretched axis with its C2R, 1/S' multiply and
    R2C in one z pass gives the numbers of the separate z passes
    (KF_PROF_UNFUSED) bit for bit, on radix-2 and radix-3 lengths (one
    x/y/z round trip per derivative on both sides)."""
    print("fused metric pass", kind, n, prec, flush=True)
    monkeypatch.setenv("KF_XY_UNSHARED", "1")
    a = _eval(ctx, kind, n, prec, False, None, env="KF_PROF_UNFUSED")
    b = _eval(ctx, kind, n, prec, False, "1", env="KF_PROF_UNFUSED")
    names = ["FUx", "FUy", "FUz", "Fn", "phi"]
    for name, x, y in zip(names, a, b):
        assert np.array_equal(x, y), (name, float(np.abs(x - y).max()))


@pytest.mark.parametrize("kind,n,prec,tol", [("sist", 24, "fp64", 1e-13), ("hyst", 32, "fp64", 1e-13),
                                             ("sist", 48, "mixed", 1e-5)])
def test_shared_xy_passes_match_one_round_trip_per_derivative(ctx, kind, n, prec, tol):
    """The three derivatives of a field from one staging of its inverse x
    and y passes (i k_y after the x pass, i k_z at the z load) against one
    x/y/z round trip per derivative (KF_XY_UNSHARED, i k_b at the x load):
    the same operator, rounded in another order.  The potential does not
    depend on the derivatives and stays bit for bit."""
    print("shared xy passes", kind, n, prec, flush=True)
    a = _eval(ctx, kind, n, prec, False, None, env="KF_XY_UNSHARED")
    b = _eval(ctx, kind, n, prec, False, "1", env="KF_XY_UNSHARED")
    for name, x, y in zip(["FUx", "FUy", "FUz", "Fn"], a, b):
        d = float(np.linalg.norm(x - y) / np.linalg.norm(y))
        assert d <= tol, (name, d)
        # the orders do round differently: the switch took effect
        assert d > 0, name
    assert np.array_equal(a[4], b[4])


def _rk_run(ctx, n, prec, steps, unfused):
    import kielflow_amd as kf
    old = os.environ.get("KF_PEN_UNFUSED")
    if unfused:
        os.environ["KF_PEN_UNFUSED"] = "1"
    else:
        os.environ.pop("KF_PEN_UNFUSED", None)
    try:
        state, nd, mask, og = initial_state("eqst", n, seed=4, Q=-0.2)
        rng = np.random.default_rng(8)
        for a in range(3):
            state[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
        g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
        p = kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED
        ps = kf.Poisson(g, p, tol=1e-12 if prec == "fp64" else 1e-10, max_it=80)
        rhs = kf.Rhs(g, ps, M=1.2, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=p)
        phi = g.rfield()
        rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
        dF = [ctx.to_device(s) for s in state]
        rk = kf.Rk4(rhs, 1e-3)
        rk.step(dF, steps)          # one call: the final update also writes the next stage input
        rk.step(dF, 1)              # and a call of one step
        assert rhs.poisson_failures() == 0
        out = [f.get() for f in dF] + [phi.get()]
        for o in (rk, rhs, ps, g):
            o.close()
        return out
    finally:
        if old is None:
            os.environ.pop("KF_PEN_UNFUSED", None)
        else:
            os.environ["KF_PEN_UNFUSED"] = old


@pytest.mark.parametrize("n,prec", [(32, "fp64"), (48, "mixed")])
def test_fused_rk_update_gives_identical_steps(ctx, n, prec):
    """The RK4 update run by the RHS's last store gives the state of the
    separate update kernels bit for bit, over several steps."""
    print("fused rk update", n, prec, flush=True)
    a = _rk_run(ctx, n, prec, 7, False)
    b = _rk_run(ctx, n, prec, 7, True)
    names = ["FUx", "FUy", "FUz", "Fn", "phi"]
    for name, x, y in zip(names, a, b):
        assert np.array_equal(x, y), (name, float(np.abs(x - y).max()))
