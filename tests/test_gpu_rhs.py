"""GPU RHS (rhs_standard::solve) and RK4 (Runge_kutta_O4::solve) against the
oracle on the same seeded inputs.  FP64: per-field rel-L2 <= 1e-11;
mixed precision: <= 1e-5."""
import numpy as np
import pytest

from helpers import initial_state, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
MIXED_TOL = 1e-5


def _perturb(og, state, seed=5, amp=1e-2):
    """Add a seeded band-limited velocity so every term of the RHS acts."""
    rng = np.random.default_rng(seed)
    out = [s.copy() for s in state]
    for a in range(3):
        u = amp * rng.standard_normal(og.rshape)
        out[a] = og.dealias23(og.r2c(u))
    return out


def _build(ctx, oracle, kind, n, sv, pen, M, Q, prec, general=False, seed=0):
    import kielflow_amd as kf
    state, nd, mask, og = initial_state(kind, n, seed=seed, Q=Q)
    state = _perturb(og, state)
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    ps = kf.Poisson(g, kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED, tol=1e-13 if prec == "fp64" else 1e-10,
                    max_it=80)
    rhs = kf.Rhs(g, ps, M=M, tau=0.1, theta=100.0, sv_on=sv, pen_on=pen,
                 precision=kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED)
    if general:
        rhs.force_general(True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    orhs = oracle.Rhs(og, M=M, tau=0.1, theta=100.0, sv_on=sv, pen_on=pen,
                      poisson_mode=oracle.POISSON_FAS, poisson_tol=1e-14, poisson_max=400)
    orhs.set_fields(np.zeros(og.rshape), nd, mask)
    return g, ps, rhs, phi, orhs, og, state


CASES = [
    ("eqst", 32, False, True, 1.0, -0.1, False),
    ("eqst", 32, False, True, 1.0, -0.1, True),
    ("eqst", 24, True, False, 1.2, -0.2, False),
    ("eqst", 48, True, True, 1.5, -0.3, False),
    ("sist", 24, True, True, 1.2, -0.2, False),
    ("hyst", 16, False, True, 1.0, -0.1, False),
]


@pytest.mark.parametrize("kind,n,sv,pen,M,Q,general", CASES)
def test_rhs_fp64_matches_oracle(ctx, oracle, kind, n, sv, pen, M, Q, general):
    print("rhs fp64", kind, n, sv, pen, general, flush=True)
    g, ps, rhs, phi, orhs, og, state = _build(ctx, oracle, kind, n, sv, pen, M, Q, "fp64", general)
    dF = [ctx.to_device(s) for s in state]
    rhs.solve(0.0, dF)
    ref, st = orhs.eval(0.0, state)
    assert st == 0
    for a in range(4):
        e = rel_l2(dF[a].get(), ref[a])
        assert e < FP64_TOL, (a, e)
    assert rel_l2(phi.get(), orhs.phi()) < FP64_TOL
    assert rhs.poisson_failures() == 0


def test_rhs_mixed_matches_oracle(ctx, oracle):
    print("rhs mixed", flush=True)
    g, ps, rhs, phi, orhs, og, state = _build(ctx, oracle, "eqst", 32, True, True, 1.5, -0.3, "mixed")
    dF = [ctx.to_device(s) for s in state]
    rhs.solve(0.0, dF)
    ref, st = orhs.eval(0.0, state)
    for a in range(4):
        e = rel_l2(dF[a].get(), ref[a])
        assert e < MIXED_TOL, (a, e)


@pytest.mark.parametrize("kind,n,steps,sv", [("eqst", 16, 20, False), ("eqst", 32, 4, False),
                                             ("sist", 16, 4, True)])
def test_rk4_fp64_matches_oracle(ctx, oracle, kind, n, steps, sv):
    import kielflow_amd as kf
    print("rk4 fp64", kind, n, steps, flush=True)
    g, ps, rhs, phi, orhs, og, state = _build(ctx, oracle, kind, n, sv, True, 1.0, -0.1, "fp64")
    dF = [ctx.to_device(s) for s in state]
    rk = kf.Rk4(rhs, 1e-3)
    ost = [s.copy() for s in state]
    t = 0.0
    rk.step(dF, steps)
    for _ in range(steps):
        t, s = orhs.rk4(1e-3, t, ost)
        assert s == 0
    for a in range(4):
        e = rel_l2(dF[a].get(), ost[a])
        assert e < FP64_TOL, (a, e)
    assert rel_l2(phi.get(), orhs.phi()) < FP64_TOL
    assert rk.t == pytest.approx(steps * 1e-3)
    assert rhs.poisson_failures() == 0


def test_rk4_mixed_matches_oracle(ctx, oracle):
    import kielflow_amd as kf
    print("rk4 mixed", flush=True)
    g, ps, rhs, phi, orhs, og, state = _build(ctx, oracle, "eqst", 32, False, True, 1.5, -0.3, "mixed")
    dF = [ctx.to_device(s) for s in state]
    rk = kf.Rk4(rhs, 1e-3)
    ost = [s.copy() for s in state]
    t = 0.0
    rk.step(dF, 4)
    for _ in range(4):
        t, s = orhs.rk4(1e-3, t, ost)
    for a in range(4):
        e = rel_l2(dF[a].get(), ost[a])
        assert e < MIXED_TOL, (a, e)


def test_spectral_operators_match_oracle(ctx, oracle):
    import kielflow_amd as kf
    for kind, n in (("eqst", 24), ("sist", 24), ("hyst", 16)):
        print("ops", kind, n, flush=True)
        state, nd, mask, og = initial_state(kind, n)
        g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
        rng = np.random.default_rng(1)
        X = og.r2c(rng.standard_normal(og.rshape))
        for d in range(3):
            dX = ctx.to_device(X)
            out = g.cfield()
            g.deriv(d, dX, out)
            assert rel_l2(out.get(), og.deriv(d, X)) < 1e-13
        dX = ctx.to_device(X)
        g.dealias23(dX)
        assert np.array_equal(dX.get() == 0, og.dealias23(X) == 0)
        assert rel_l2(dX.get(), og.dealias23(X)) == 0.0
        dX = ctx.to_device(X)
        g.filter_phi(dX)
        assert rel_l2(dX.get(), og.filter_phi(X)) < 1e-14


def test_cfl_matches_oracle(ctx, oracle):
    import kielflow_amd as kf
    state, nd, mask, og = initial_state("sist", 16)
    state = _perturb(og, state, amp=0.5)
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    d = [ctx.to_device(s) for s in state[:3]]
    assert g.cfl(0.01, *d) == pytest.approx(og.cfl(0.01, *state[:3]), rel=1e-13)


def test_config0_full_run_matches_oracle(ctx, oracle):
    """32^3 EqSt, M = 1, Q = -0.1, u = 0, seed-0 noise, SV off, penalization
    on, dt = 1e-3: the full 100 steps, FP64, checked every 25 steps."""
    import kielflow_amd as kf
    print("config0 100 steps", flush=True)
    state, nd, mask, og = initial_state("eqst", 32, seed=0, Q=-0.1)
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=80)
    rhs = kf.Rhs(g, ps, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    orhs = oracle.Rhs(og, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=True,
                      poisson_mode=oracle.POISSON_FAS, poisson_tol=1e-14, poisson_max=400)
    orhs.set_fields(np.zeros(og.rshape), nd, mask)
    dF = [ctx.to_device(s) for s in state]
    ost = [s.copy() for s in state]
    rk = kf.Rk4(rhs, 1e-3)
    t = 0.0
    for block in range(4):
        rk.step(dF, 25)
        for _ in range(25):
            t, s = orhs.rk4(1e-3, t, ost)
            assert s == 0
        for a in range(4):
            e = rel_l2(dF[a].get(), ost[a])
            assert e < FP64_TOL, (25 * (block + 1), a, e)
        assert rel_l2(phi.get(), orhs.phi()) < FP64_TOL
    assert rhs.poisson_failures() == 0


def test_config1_sist_with_cfl_dt_matches_oracle(ctx, oracle):
    """Sine-stretched axes (S'(0) = 1/3), M = 1.2, Q = -0.2, spectral
    viscosity on, dt from cfl_all every step (CFL.cpp:29-82), FP64: three
    steps at 64^3 against the oracle."""
    import kielflow_amd as kf
    print("config1 sist 64", flush=True)
    state, nd, mask, og = initial_state("sist", 64, seed=0, Q=-0.2)
    state = _perturb(og, state)
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=80)
    rhs = kf.Rhs(g, ps, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    orhs = oracle.Rhs(og, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True,
                      poisson_mode=oracle.POISSON_FAS, poisson_tol=1e-14, poisson_max=400)
    orhs.set_fields(np.zeros(og.rshape), nd, mask)
    dF = [ctx.to_device(s) for s in state]
    ost = [s.copy() for s in state]
    rk = kf.Rk4(rhs, 1e-3)
    dt, dt_o, t = 1e-3, 1e-3, 0.0
    for step in range(3):
        dt = g.cfl(dt, dF[0], dF[1], dF[2])
        dt_o = og.cfl(dt_o, ost[0], ost[1], ost[2])
        assert dt == pytest.approx(dt_o, rel=1e-12)
        rk.set_dt(dt)
        rk.step(dF, 1)
        t, s = orhs.rk4(dt_o, t, ost)
        assert s == 0
    for a in range(4):
        e = rel_l2(dF[a].get(), ost[a])
        assert e < FP64_TOL, (a, e)
    assert rel_l2(phi.get(), orhs.phi()) < FP64_TOL


def test_mixed_newton_work_matches_fp64(ctx, monkeypatch):
    """Mixed RK4 forms rho from the FP64 stage input of ln n: at 128^3 its
    Newton solves need no more corrections than the FP64 run's, and fewer
    than with rho from the FP32 stage input (KF_RHO_STATE=0), whose rounding
    reaches the Newton residual through the potential history; every stage
    solves once."""
    import kielflow_amd as kf
    print("mixed newton counts", flush=True)
    n, steps = 128, 30
    axes = kf.cube(n, 16.0)
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.3, a=2.0 * 16.0 / n)
    g = kf.Grid(ctx, axes)
    fft = kf.FFT(g, kf.KF_FP64)
    per = {}
    for name, P, rho_state in (("mixed", kf.KF_MIXED, None), ("mixed_fp32_rho", kf.KF_MIXED, "0"),
                               ("fp64", kf.KF_FP64, None)):
        if rho_state is not None:
            monkeypatch.setenv("KF_RHO_STATE", rho_state)
        ps = kf.Poisson(g, P, tol=1e-10, max_it=60)
        rhs = kf.Rhs(g, ps, M=1.5, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=P)
        monkeypatch.delenv("KF_RHO_STATE", raising=False)
        phi = g.rfield()
        rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
        F = [g.cfield() for _ in range(4)]
        fft.r2c(ctx.to_device(ln_n), F[3])
        rk = kf.Rk4(rhs, 1e-3)
        rk.step(F, 10)
        s0, c0 = ps.newton_counts()
        rk.step(F, steps)
        s1, c1 = ps.newton_counts()
        assert s1 - s0 == 4 * steps
        assert rhs.poisson_failures() == 0
        per[name] = (c1 - c0) / (s1 - s0)
        for o in (rk, rhs, ps):
            o.close()
    print("corrections per solve", per, flush=True)
    assert per["mixed"] <= per["fp64"] + 0.25, per
    assert per["mixed"] <= 0.7 * per["mixed_fp32_rho"], per
    fft.close()
    g.close()
