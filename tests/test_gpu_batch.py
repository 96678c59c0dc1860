"""Batched boxes of uneven sizes: the graph-replayed batch step equals the
same boxes stepped alone (eagerly) after every replay, and the mixed batch
stays within 1e-5 of FP64 solo runs."""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _box_setup(kf, n, Q, seed):
    axes = kf.cube(n, 16.0)
    x0 = kf.api.axis_values(axes[0], n // 2)[0]
    x1 = kf.api.axis_values(axes[0], n // 2 + 1)[0]
    return axes, kf.setup_host(axes, seed=seed, amp=1e-3, Q=Q, a=2.0 * (x1 - x0))


def _solo(kf, ctx, axes, fields, M, precision):
    ln_n, nd, mask = fields
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, precision, tol=1e-10, max_it=100)
    rhs = kf.Rhs(g, ps, M=M, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=precision)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    F = [g.cfield() for _ in range(4)]
    fft = kf.FFT(g, kf.KF_FP64)
    fft.r2c(ctx.to_device(ln_n), F[3])
    rk = kf.Rk4(rhs, 1e-3)
    return dict(g=g, ps=ps, rhs=rhs, phi=phi, F=F, fft=fft, rk=rk)


def test_batch_graph_replays_match_solo_runs(ctx):
    import kielflow_amd as kf
    rng = np.random.default_rng(4)
    sizes = [16, 24, 32, 16, 24, 32]
    Ms = rng.uniform(0.5, 2.0, len(sizes))
    Qs = rng.uniform(-0.5, -0.05, len(sizes))
    boxes, setups = [], []
    for b, n in enumerate(sizes):
        axes, fields = _box_setup(kf, n, Qs[b], seed=100 + b)
        boxes.append((axes, kf.api.Parameters(Ms[b], 0.1, 100.0, 0.0, 0.0)))
        setups.append((axes, fields))
    print("batch", sizes, flush=True)
    B = kf.Batch(ctx, boxes, precision=kf.KF_MIXED, nstreams=4)
    solos = []
    fft_keep = []
    for b, (axes, (ln_n, nd, mask)) in enumerate(setups):
        B.set_fields(b, ctx.to_device(nd), ctx.to_device(mask))
        F, phi = B.box(b)
        g = kf.Grid(ctx, axes)
        fft = kf.FFT(g, kf.KF_FP64)
        fft.r2c(ctx.to_device(ln_n), F[3])
        fft_keep.append((fft, g))
        solos.append(_solo(kf, ctx, axes, (ln_n, nd, mask), Ms[b], kf.KF_MIXED))
    warm = 5  # the potential history is full: the graph keeps the steady-state extrapolation
    B.steps(warm)
    for s in solos:
        s["rk"].step(s["F"], warm)
    B.sync()
    B.capture(6)
    # captured solves smooth with two pre and two post sweeps (a fixed
    # iteration count gains nothing from more): the solo runs do the same
    for s in solos:
        s["ps"].set_smoothing(2, 2, 24)
    for rep in range(4):
        B.steps(1)
        B.sync()
        for b, s in enumerate(solos):
            s["rk"].step(s["F"], 1)
            F, phi = B.box(b)
            for a in range(4):
                err = rel_l2(F[a].get(), s["F"][a].get())
                assert err <= 1e-13, (rep, b, a, err)
    assert B.poisson_failures() == 0
    # the mixed batch against FP64 solo runs of the same boxes
    for b, (axes, fields) in enumerate(setups):
        s = _solo(kf, ctx, axes, fields, Ms[b], kf.KF_FP64)
        s["rk"].step(s["F"], warm + 4)
        F, _ = B.box(b)
        for a in range(4):
            ref = s["F"][a].get()
            if np.abs(ref).max() == 0:
                continue
            err = rel_l2(F[a].get(), ref)
            assert err <= 1e-5, (b, a, err)
    B.close()


def test_batch_capture_needs_full_histories(ctx):
    """kf_batch_capture refuses to record a step before the potential
    histories are full (the graph would keep a lower extrapolation order)."""
    import kielflow_amd as kf
    axes = kf.cube(16, 16.0)
    B = kf.Batch(ctx, [(axes, kf.api.Parameters(1.0, 0.1, 100.0, 0.0, 0.0))], precision=kf.KF_MIXED, nstreams=1)
    ln_n, nd, mask = kf.setup_host(axes, seed=3, amp=1e-3, Q=-0.1, a=2.0)
    B.set_fields(0, ctx.to_device(nd), ctx.to_device(mask))
    B.steps(2)
    B.sync()
    with pytest.raises(kf.KfError, match="eager steps"):
        B.capture(6)
    B.steps(3)
    B.sync()
    B.capture(6)
    B.steps(2)
    B.sync()
    assert B.poisson_failures() == 0
    B.close()

