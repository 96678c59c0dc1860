"""The five workloads at their own sizes (or, where the FP64 oracle cannot
keep up, at their own sizes over fewer steps), each checked against the
oracle or against a size-independent property:

  config 0  32^3 EqSt, 100 steps, FP64           test_gpu_rhs.test_config0_full_run_matches_oracle
  config 1  128^3 SiSt, CFL dt, FP64             one step against the oracle
  config 2  256^3 EqSt, mixed vs FP64            100 steps, per-field rel-L2 <= 1e-5 every 25 steps
  config 3  64^3 HySt adjoint, checkpointed      20 steps with segment recomputation against the
                                                 oracle's tangent-linear run (golden fixture)
  config 4  boxes of {32, 48, 64, 96}^3, batch   graph replays; 32^3 and 48^3 boxes against the
                                                 oracle, every box against FP64 solo runs
"""
import json
import os

import numpy as np
import pytest

from helpers import axes_for, cell_volumes, grain_radius, initial_state, rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FP64_TOL = 1e-11
MIXED_TOL = 1e-5


def _perturb(og, state, seed=5, amp=1e-2):
    """A seeded band-limited velocity so that every term of the RHS acts."""
    rng = np.random.default_rng(seed)
    out = [s.copy() for s in state]
    for a in range(3):
        out[a] = og.dealias23(og.r2c(amp * rng.standard_normal(og.rshape)))
    return out


def test_config1_128_sist_step_with_cfl_dt_matches_oracle(ctx, oracle):
    """128^3 sine-stretched axes (S'(0) = 1/3), M = 1.2, Q = -0.2, spectral
    viscosity and penalization on, dt from cfl_all: one FP64 RK4 step
    against the oracle at rel-L2 <= 1e-11 per field (the oracle needs tens
    of seconds per step at this size)."""
    import kielflow_amd as kf
    n = 128
    print("config1 128^3 sist: setup", flush=True)
    state, nd, mask, og = initial_state("sist", n, seed=0, Q=-0.2)
    state = _perturb(og, state)
    g = kf.Grid(ctx, [kf.axis(ax.type, ax.N, ax.l0, ax.L, ax.m) for ax in og.axes])
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=80)
    rhs = kf.Rhs(g, ps, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    dF = [ctx.to_device(s) for s in state]
    dt = g.cfl(1e-3, dF[0], dF[1], dF[2])
    dt_o = og.cfl(1e-3, state[0], state[1], state[2])
    assert dt == pytest.approx(dt_o, rel=1e-12)
    rk = kf.Rk4(rhs, dt)
    rk.step(dF, 1)
    got = [d.get() for d in dF]
    print(f"config1: GPU step done (dt = {dt:.6e}); oracle step", flush=True)
    orhs = oracle.Rhs(og, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True,
                      poisson_mode=oracle.POISSON_FAS, poisson_tol=1e-14, poisson_max=600)
    orhs.set_fields(np.zeros(og.rshape), nd, mask)
    ost = [s.copy() for s in state]
    _, s = orhs.rk4(dt_o, 0.0, ost)
    assert s == 0
    for a in range(4):
        e = rel_l2(got[a], ost[a])
        print(f"config1 field {a}: rel-L2 {e:.3e}", flush=True)
        assert e < FP64_TOL, (a, e)
    assert rel_l2(phi.get(), orhs.phi()) < FP64_TOL
    assert rhs.poisson_failures() == 0
    for o in (rk, rhs, ps, g):
        o.close()


def test_config2_256_mixed_against_fp64_over_100_steps(ctx):
    """256^3 EqSt, M = 1.5, Q = -0.3: the mixed-precision run stays within
    rel-L2 1e-5 of the FP64 run on every field and on phi, checked every 25
    of 100 steps (the FP64 path is itself pinned to the oracle)."""
    import kielflow_amd as kf
    n = 256
    axes = kf.cube(n, 16.0)
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.3, a=2.0 * 16.0 / n)
    g = kf.Grid(ctx, axes)
    fft = kf.FFT(g, kf.KF_FP64)
    nd_d, mask_d = ctx.to_device(nd), ctx.to_device(mask)
    runs = {}
    for name, P in (("mixed", kf.KF_MIXED), ("fp64", kf.KF_FP64)):
        ps = kf.Poisson(g, P, tol=1e-10, max_it=60)
        rhs = kf.Rhs(g, ps, M=1.5, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=P)
        phi = g.rfield()
        rhs.set_fields(phi, nd_d, mask_d)
        F = [g.cfield() for _ in range(4)]
        fft.r2c(ctx.to_device(ln_n), F[3])
        runs[name] = dict(ps=ps, rhs=rhs, phi=phi, F=F, rk=kf.Rk4(rhs, 1e-3))
    for block in range(4):
        for r in runs.values():
            r["rk"].step(r["F"], 25)
        a, b = runs["mixed"], runs["fp64"]
        errs = [rel_l2(a["F"][i].get(), b["F"][i].get()) for i in range(4)]
        errs.append(rel_l2(a["phi"].get(), b["phi"].get()))
        print(f"config2 step {25 * (block + 1)}: rel-L2 (ux, uy, uz, ln n, phi) "
              + " ".join(f"{e:.2e}" for e in errs), flush=True)
        assert max(errs) <= MIXED_TOL, (25 * (block + 1), errs)
    for r in runs.values():
        assert r["rhs"].poisson_failures() == 0
        for o in (r["rk"], r["rhs"], r["ps"]):
            o.close()
    fft.close()
    g.close()


def _config3_fixture():
    with open(os.path.join(HERE, "golden", "config3_tangent.json")) as f:
        return json.load(f)


def test_config3_64_hyst_checkpointed_gradients_match_tangent_run(ctx, oracle):
    """64^3 hyperbolic-sinh axes (m = 3), FP64, checkpoints every 5 steps with
    each segment's stages recomputed from its checkpoint: J, dJ/dM, dJ/dQ and
    dJ along a seeded density direction after 20 steps match the oracle's
    FP64 tangent-linear run (tests/golden/config3_tangent.json)."""
    import kielflow_amd as kf
    c = _config3_fixture()
    n, steps, k = c["n"], c["steps"], 5
    print(f"config3 {n}^3 hyst adjoint, {steps} steps, checkpoints every {k}", flush=True)
    axes, oaxes = axes_for(c["kind"], n, c["m"])
    ln_n, nd, mask = kf.setup_host(axes, seed=c["seed"], amp=c["amp"], Q=c["Q"],
                                   a=grain_radius(c["kind"], n, c["m"]))
    og = oracle.Grid(*oaxes)
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=c["poisson_tol"], max_it=100)
    rhs = kf.Rhs(g, ps, M=c["M"], tau=c["tau"], theta=c["theta"], sv_on=c["sv_on"], pen_on=c["pen_on"])
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    y = [g.cfield() for _ in range(4)]
    y[3].upload(og.r2c(ln_n))
    G = nd / c["Q"]
    dV = ctx.to_device(cell_volumes(og))
    adj = kf.Adjoint(rhs, c["dt"], tol=1e-13)
    adj.set_capture_budget(k * adj.step_capture_bytes())
    dFn0 = g.cfield()
    J, dM, dQ = adj.run(y, steps, k, dV, ctx.to_device(G), dFn0)
    assert adj.run_info()[0] == kf.api.KF_ADJ_SEGMENTS
    z = oracle.normals(c["v_seed"], n ** 3).reshape(og.rshape)
    v = og.dealias23(og.r2c(z)) * c["v_scale"]
    d = dFn0.get()
    dv = float(np.sum(v.real * d.real + v.imag * d.imag))
    print(f"config3: J {J:.12e}/{c['J']:.12e} dM {dM:.12e}/{c['dJ_dM']:.12e} "
          f"dQ {dQ:.12e}/{c['dJ_dQ']:.12e} dv {dv:.12e}/{c['dJ_dv']:.12e}", flush=True)
    assert abs(J - c["J"]) <= 1e-9 * abs(c["J"])
    assert abs(dM - c["dJ_dM"]) <= 1e-7 * abs(c["dJ_dM"])
    assert abs(dQ - c["dJ_dQ"]) <= 1e-7 * abs(c["dJ_dQ"])
    assert abs(dv - c["dJ_dv"]) <= 1e-7 * abs(c["dJ_dv"])
    for o in (adj, rhs, ps, g):
        o.close()


def _box(kf, n, Q, seed):
    axes = kf.cube(n, 16.0)
    x0 = kf.api.axis_values(axes[0], n // 2)[0]
    x1 = kf.api.axis_values(axes[0], n // 2 + 1)[0]
    return axes, kf.setup_host(axes, seed=seed, amp=1e-3, Q=Q, a=2.0 * (x1 - x0))


def test_config4_batch_of_uneven_boxes(ctx, oracle):
    """Boxes of 32, 48, 64 and 96 points per side, M in [0.5, 2], Q in
    [-0.5, -0.05], mixed precision: eager steps, then graph-replayed steps
    with eager work between the replays.  The 32^3 and 48^3 boxes match the
    oracle's FP64 RK4 and every box matches an FP64 solo run, at rel-L2 1e-5
    per field."""
    import kielflow_amd as kf
    rng = np.random.default_rng(11)
    sizes = [32, 48, 64, 96, 48, 32]
    Ms = rng.uniform(0.5, 2.0, len(sizes))
    Qs = rng.uniform(-0.5, -0.05, len(sizes))
    setups = [_box(kf, n, Qs[b], 500 + b) for b, n in enumerate(sizes)]
    boxes = [(axes, kf.api.Parameters(Ms[b], 0.1, 100.0, 0.0, 0.0)) for b, (axes, _) in enumerate(setups)]
    print("config4 batch", sizes, flush=True)
    B = kf.Batch(ctx, boxes, precision=kf.KF_MIXED, nstreams=4)
    for b, (axes, (ln_n, nd, mask)) in enumerate(setups):
        B.set_fields(b, ctx.to_device(nd), ctx.to_device(mask))
        F, _ = B.box(b)
        g = kf.Grid(ctx, axes)
        fft = kf.FFT(g, kf.KF_FP64)
        fft.r2c(ctx.to_device(ln_n), F[3])
        fft.close()
        g.close()
    warm, replays = 5, 3
    B.steps(warm)
    B.sync()
    B.capture(6)
    # an FP64 solo run of one box stepped between the replays (eager work on
    # the library's stream while the graphs are live)
    axes0, (ln0, nd0, mask0) = setups[0]
    g0 = kf.Grid(ctx, axes0)
    ps0 = kf.Poisson(g0, kf.KF_FP64, tol=1e-10, max_it=100)
    rhs0 = kf.Rhs(g0, ps0, M=Ms[0], tau=0.1, theta=100.0, sv_on=False, pen_on=True)
    phi0 = g0.rfield()
    rhs0.set_fields(phi0, ctx.to_device(nd0), ctx.to_device(mask0))
    F0 = [g0.cfield() for _ in range(4)]
    fft0 = kf.FFT(g0, kf.KF_FP64)
    fft0.r2c(ctx.to_device(ln0), F0[3])
    rk0 = kf.Rk4(rhs0, 1e-3)
    rk0.step(F0, warm)
    for _ in range(replays):
        B.steps(1)
        rk0.step(F0, 1)
        ctx.sync()
        B.sync()
    assert B.poisson_failures() == 0
    total = warm + replays
    for b, (axes, (ln_n, nd, mask)) in enumerate(setups):
        Fb, _ = B.box(b)
        got = [f.get() for f in Fb]
        if b == 0:
            ref = [f.get() for f in F0]
        else:
            g = kf.Grid(ctx, axes)
            ps = kf.Poisson(g, kf.KF_FP64, tol=1e-10, max_it=100)
            rhs = kf.Rhs(g, ps, M=Ms[b], tau=0.1, theta=100.0, sv_on=False, pen_on=True)
            phi = g.rfield()
            rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
            F = [g.cfield() for _ in range(4)]
            fft = kf.FFT(g, kf.KF_FP64)
            fft.r2c(ctx.to_device(ln_n), F[3])
            rk = kf.Rk4(rhs, 1e-3)
            rk.step(F, total)
            ref = [f.get() for f in F]
            for o in (rk, fft, rhs, ps, g):
                o.close()
        errs = [rel_l2(got[a], ref[a]) for a in range(4) if np.abs(ref[a]).max() > 0]
        print(f"config4 box {b} ({sizes[b]}^3) vs FP64 solo: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
        assert max(errs) <= MIXED_TOL, (b, errs)
        if sizes[b] <= 48 and b in (0, 1):
            og = oracle.Grid.cube(sizes[b], 16.0)
            orhs = oracle.Rhs(og, M=Ms[b], tau=0.1, theta=100.0, sv_on=False, pen_on=True,
                              poisson_mode=oracle.POISSON_FAS, poisson_tol=1e-13, poisson_max=400)
            orhs.set_fields(np.zeros(og.rshape), nd, mask)
            ost = [np.zeros(og.cshape, complex) for _ in range(3)] + [og.r2c(ln_n)]
            t = 0.0
            for _ in range(total):
                t, s = orhs.rk4(1e-3, t, ost)
                assert s == 0
            errs = [rel_l2(got[a], ost[a]) for a in range(4) if np.abs(ost[a]).max() > 0]
            print(f"config4 box {b} ({sizes[b]}^3) vs oracle: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
            assert max(errs) <= MIXED_TOL, (b, errs)
    for o in (rk0, fft0, rhs0, ps0, g0):
        o.close()
    B.close()
