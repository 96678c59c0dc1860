"""Discrete adjoint of the RHS, the RK4 step and the grain drag objective:
dot-product tests against central finite differences of the forward
operators, and a checkpointed run against finite differences of J."""
import numpy as np
import pytest

from helpers import cell_volumes, initial_state

pytestmark = pytest.mark.gpu


def _setup(ctx, kind, n, pen=True, sv=False, M=1.0, Q=-0.2, tol=1e-13):
    import kielflow_amd as kf
    state, nd, mask, og = initial_state(kind, n, seed=2, Q=Q)
    rng = np.random.default_rng(9)
    for a in range(3):
        state[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    ps = kf.Poisson(g, kf.KF_FP64, tol=tol, max_it=100)
    rhs = kf.Rhs(g, ps, M=M, tau=0.1, theta=100.0, sv_on=sv, pen_on=pen)
    phi = g.rfield()
    d_nd = ctx.to_device(nd)
    rhs.set_fields(phi, d_nd, ctx.to_device(mask))
    return g, ps, rhs, phi, state, nd, mask, og


def _ip(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag)) if np.iscomplexobj(a) else float(np.sum(a * b))


def _rhs_host(ctx, g, rhs, state):
    d = [ctx.to_device(s) for s in state]
    rhs.solve(0.0, d)
    return [x.get() for x in d]


@pytest.mark.parametrize("kind,n,pen,sv", [("eqst", 16, True, False), ("eqst", 16, False, True),
                                           ("sist", 16, True, True), ("hyst", 16, True, False)])
def test_rhs_adjoint_dot_product(ctx, kind, n, pen, sv):
    import kielflow_amd as kf
    print("rhs adjoint", kind, n, pen, sv, flush=True)
    g, ps, rhs, phi, S, nd, mask, og = _setup(ctx, kind, n, pen, sv)
    rng = np.random.default_rng(4)
    v = [og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-3 for _ in range(4)]
    w = [rng.standard_normal(og.cshape) + 1j * rng.standard_normal(og.cshape) for _ in range(4)]
    adj = kf.Adjoint(rhs, 1e-3)
    adj.reset()
    Fb = [g.cfield() for _ in range(4)]
    adj.rhs_adjoint([ctx.to_device(s) for s in S], [ctx.to_device(x) for x in w], Fb)
    Fb = [f.get() for f in Fb]
    Mbar = adj.accumulators()
    ndbar = g.rfield()
    adj.accumulators(ndbar)
    ndbar = ndbar.get()
    # state direction
    h = 1e-2
    fp = _rhs_host(ctx, g, rhs, [s + h * x for s, x in zip(S, v)])
    fm = _rhs_host(ctx, g, rhs, [s - h * x for s, x in zip(S, v)])
    lhs = sum(_ip((p - m) / (2 * h), ww) for p, m, ww in zip(fp, fm, w))
    rhs_ = sum(_ip(x, f) for x, f in zip(v, Fb))
    assert abs(lhs - rhs_) <= 1e-6 * abs(lhs), (lhs, rhs_)
    # parameter M
    hm = 1e-4
    rhs.set_parameters(M=1.0 + hm)
    fp = _rhs_host(ctx, g, rhs, S)
    rhs.set_parameters(M=1.0 - hm)
    fm = _rhs_host(ctx, g, rhs, S)
    rhs.set_parameters(M=1.0)
    dM = sum(_ip((p - m) / (2 * hm), ww) for p, m, ww in zip(fp, fm, w))
    assert abs(dM - Mbar) <= 1e-6 * max(abs(dM), 1e-12), (dM, Mbar)
    # grain density direction
    dnd = rng.standard_normal(og.rshape) * 1e-2
    hq = 1e-2
    rhs.set_fields(None, ctx.to_device(nd + hq * dnd), None)
    fp = _rhs_host(ctx, g, rhs, S)
    rhs.set_fields(None, ctx.to_device(nd - hq * dnd), None)
    fm = _rhs_host(ctx, g, rhs, S)
    rhs.set_fields(None, ctx.to_device(nd), None)
    dn = sum(_ip((p - m) / (2 * hq), ww) for p, m, ww in zip(fp, fm, w))
    assert abs(dn - _ip(dnd, ndbar)) <= 1e-6 * abs(dn), (dn, _ip(dnd, ndbar))


def _rhs_adjoint_host(ctx, kind, n, unfused):
    import os
    import kielflow_amd as kf
    old = os.environ.get("KF_PROF_UNFUSED")
    if unfused:
        os.environ["KF_PROF_UNFUSED"] = "1"
    else:
        os.environ.pop("KF_PROF_UNFUSED", None)
    try:
        g, ps, rhs, phi, S, nd, mask, og = _setup(ctx, kind, n, True, True)
        rng = np.random.default_rng(5)
        w = [rng.standard_normal(og.cshape) + 1j * rng.standard_normal(og.cshape) for _ in range(4)]
        adj = kf.Adjoint(rhs, 1e-3)
        adj.reset()
        Fb = [g.cfield() for _ in range(4)]
        adj.rhs_adjoint([ctx.to_device(s) for s in S], [ctx.to_device(x) for x in w], Fb)
        ndbar = g.rfield()
        Mbar = adj.accumulators(ndbar)
        out = [f.get() for f in Fb] + [ndbar.get()], Mbar
        for o in (adj, rhs, ps, g):
            o.close()
        return out
    finally:
        if old is None:
            os.environ.pop("KF_PROF_UNFUSED", None)
        else:
            os.environ["KF_PROF_UNFUSED"] = old


@pytest.mark.parametrize("kind,n", [("sist", 16), ("hyst", 24)])
def test_fused_metric_pass_in_the_adjoint_gives_identical_results(ctx, monkeypatch, kind, n):
    """The transposed stretched-axis derivatives with C2R, dS_b multiply and
    R2C in one z pass give the adjoint RHS of the separate z passes
    (KF_PROF_UNFUSED) bit for bit (the forward derivatives it recomputes
    with one x/y/z round trip each on both sides)."""
    print("adjoint fused metric pass", kind, n, flush=True)
    monkeypatch.setenv("KF_XY_UNSHARED", "1")
    (a, ma), (b, mb) = _rhs_adjoint_host(ctx, kind, n, False), _rhs_adjoint_host(ctx, kind, n, True)
    assert ma == mb
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), (i, float(np.abs(x - y).max()))


def _step_host(ctx, rhs, y, dt=1e-3):
    import kielflow_amd as kf
    d = [ctx.to_device(s) for s in y]
    rk = kf.Rk4(rhs, dt)
    rk.step(d, 1)
    out = [x.get() for x in d]
    rk.close()
    return out


@pytest.mark.parametrize("kind", ["eqst", "hyst"])
def test_rk4_step_adjoint_dot_product(ctx, kind):
    import kielflow_amd as kf
    print("step adjoint", kind, flush=True)
    g, ps, rhs, phi, y0, nd, mask, og = _setup(ctx, kind, 16)
    rng = np.random.default_rng(5)
    v = [og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-3 for _ in range(4)]
    w = [rng.standard_normal(og.cshape) + 1j * rng.standard_normal(og.cshape) for _ in range(4)]
    adj = kf.Adjoint(rhs, 1e-3)
    adj.reset()
    yb = [ctx.to_device(x) for x in w]
    adj.step_adjoint([ctx.to_device(s) for s in y0], yb)
    yb = [x.get() for x in yb]
    h = 1e-2
    fp = _step_host(ctx, rhs, [s + h * x for s, x in zip(y0, v)])
    fm = _step_host(ctx, rhs, [s - h * x for s, x in zip(y0, v)])
    lhs = sum(_ip((p - m) / (2 * h), ww) for p, m, ww in zip(fp, fm, w))
    rr = sum(_ip(x, f) for x, f in zip(v, yb))
    assert abs(lhs - rr) <= 1e-7 * abs(lhs), (lhs, rr)


def _dV(og):
    return cell_volumes(og)


@pytest.mark.parametrize("kind", ["eqst", "sist"])
def test_objective_adjoint(ctx, kind):
    import kielflow_amd as kf
    g, ps, rhs, phi, F, nd, mask, og = _setup(ctx, kind, 16)
    dV = ctx.to_device(_dV(og))
    adj = kf.Adjoint(rhs, 1e-3)
    adj.reset()
    dF = [ctx.to_device(s) for s in F]
    J = adj.objective(dF, dV)
    Fb = [g.cfield() for _ in range(4)]
    adj.objective_adjoint(dF, dV, Fb)
    Fb = [f.get() for f in Fb]
    rng = np.random.default_rng(6)
    v = og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-2
    h = 1e-2
    Jp = adj.objective([dF[0], dF[1], dF[2], ctx.to_device(F[3] + h * v)], dV)
    Jm = adj.objective([dF[0], dF[1], dF[2], ctx.to_device(F[3] - h * v)], dV)
    fd = (Jp - Jm) / (2 * h)
    assert abs(fd - _ip(v, Fb[3])) <= 1e-6 * abs(fd), (fd, _ip(v, Fb[3]), J)


def _tangent_dJ(og, F0, nd, mask, steps, dV, dM=0.0, dnd=None, dFn=None, dt=1e-3):
    """dJ along (dM, dnd, dFn0) from the oracle's FP64 tangent-linear run."""
    from oracle import oracle as O
    r = O.Rhs(og, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=True, poisson_mode=O.POISSON_FAS,
              poisson_tol=1e-14, poisson_max=400)
    r.set_fields(np.zeros(og.rshape), nd, mask)
    y = [np.array(f, dtype=np.complex128) for f in F0]
    ty = [np.zeros(og.cshape, complex) for _ in range(3)] + [
        np.zeros(og.cshape, complex) if dFn is None else np.array(dFn, dtype=np.complex128)]
    for _ in range(steps):
        assert r.rk4_tl(dt, y, ty, dM, dnd) == 0
    return r.objective_tl(y[3], ty[3], dV, dnd)


def test_checkpointed_run_gradients(ctx):
    """dJ/dM, dJ/dQ and dJ/dFn0 of a short run: the kept trajectory, kept
    segments and per-step recomputation agree, and match central
    differences and the CPU tangent run."""
    import kielflow_amd as kf
    print("adjoint run", flush=True)
    kind, n, steps, k = "hyst", 16, 6, 4
    g, ps, rhs, phi, F0, nd, mask, og = _setup(ctx, kind, n, Q=-0.2)
    G = nd / -0.2
    dVh = _dV(og)
    dV = ctx.to_device(dVh)
    dG = ctx.to_device(G)
    adj = kf.Adjoint(rhs, 1e-3)
    phi0 = phi.get()

    def run_J(M=1.0, Q=-0.2, F=None):
        rhs.set_parameters(M=M)
        rhs.set_fields(None, ctx.to_device(Q * G), None)
        phi.upload(phi0)
        y = [ctx.to_device(s) for s in (F0 if F is None else F)]
        rk = kf.Rk4(rhs, 1e-3)
        rk.step(y, steps)
        rk.close()
        J = adj.objective(y, dV)
        rhs.set_parameters(M=1.0)
        rhs.set_fields(None, ctx.to_device(-0.2 * G), None)
        return J

    def run_adj():
        phi.upload(phi0)
        y = [ctx.to_device(s) for s in F0]
        dFn0 = g.cfield()
        J, dM, dQ = adj.run(y, steps, k, dV, dG, dFn0)
        return J, dM, dQ, dFn0.get()

    J, dM, dQ, dFn0 = run_adj()
    assert adj.run_info()[0] == kf.api.KF_ADJ_TRAJECTORY
    # room for one segment only, then for nothing
    for budget, mode in ((k * adj.step_capture_bytes(), kf.api.KF_ADJ_SEGMENTS), (0, kf.api.KF_ADJ_RECOMPUTE)):
        adj.set_capture_budget(budget)
        J2, dM2, dQ2, dFn02 = run_adj()
        assert adj.run_info()[0] == mode
        print(f"mode {mode}: dM {dM2:.12e}/{dM:.12e} dQ {dQ2:.12e}/{dQ:.12e}", flush=True)
        assert abs(J2 - J) <= 1e-9 * abs(J)
        assert abs(dM2 - dM) <= 1e-9 * abs(dM) and abs(dQ2 - dQ) <= 1e-9 * abs(dQ)
        assert np.abs(dFn02 - dFn0).max() <= 1e-9 * np.abs(dFn0).max()
    # segments whose recomputation starts each solve from the potential the
    # forward sweep kept reproduce that sweep: the kept trajectory's numbers
    adj.set_capture_budget(k * adj.step_capture_bytes())
    for keep, tol in ((True, 1e-12), (False, 1e-9)):
        adj.set_keep_potentials(keep)
        J3, dM3, dQ3, dFn03 = run_adj()
        assert adj.run_info()[0] == kf.api.KF_ADJ_SEGMENTS
        print(f"segments keep={keep}: dM {dM3:.15e}/{dM:.15e}", flush=True)
        assert abs(J3 - J) <= tol * abs(J)
        assert abs(dM3 - dM) <= tol * abs(dM) and abs(dQ3 - dQ) <= tol * abs(dQ)
        assert np.abs(dFn03 - dFn0).max() <= tol * np.abs(dFn0).max()
    adj.set_keep_potentials(True)
    J0 = run_J()
    assert abs(J - J0) <= 1e-8 * abs(J0)
    hM, hQ = 1e-3, 1e-4
    fdM = (run_J(M=1.0 + hM) - run_J(M=1.0 - hM)) / (2 * hM)
    fdQ = (run_J(Q=-0.2 + hQ) - run_J(Q=-0.2 - hQ)) / (2 * hQ)
    print(f"J={J:.6e} dM={dM:.10e} fd={fdM:.10e} dQ={dQ:.10e} fd={fdQ:.10e}", flush=True)
    assert abs(fdM - dM) <= 1e-4 * abs(fdM), (fdM, dM)
    assert abs(fdQ - dQ) <= 1e-4 * abs(fdQ), (fdQ, dQ)
    rng = np.random.default_rng(8)
    v = og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-2
    hv = 1e-3
    Fp = [F0[0], F0[1], F0[2], F0[3] + hv * v]
    Fm = [F0[0], F0[1], F0[2], F0[3] - hv * v]
    fdv = (run_J(F=Fp) - run_J(F=Fm)) / (2 * hv)
    assert abs(fdv - _ip(v, dFn0)) <= 1e-4 * abs(fdv), (fdv, _ip(v, dFn0))
    # the CPU FP64 tangent-linear run along the same three directions
    Jt, tM = _tangent_dJ(og, F0, nd, mask, steps, dVh, dM=1.0)
    _, tQ = _tangent_dJ(og, F0, nd, mask, steps, dVh, dnd=G)
    _, tv = _tangent_dJ(og, F0, nd, mask, steps, dVh, dFn=v)
    print(f"tangent: J={Jt:.10e} dM={tM:.12e}/{dM:.12e} dQ={tQ:.12e}/{dQ:.12e} v={tv:.12e}/{_ip(v, dFn0):.12e}",
          flush=True)
    assert abs(Jt - J) <= 1e-8 * abs(J)
    assert abs(tM - dM) <= 1e-7 * abs(tM)
    assert abs(tQ - dQ) <= 1e-7 * abs(tQ)
    assert abs(tv - _ip(v, dFn0)) <= 1e-7 * abs(tv)
