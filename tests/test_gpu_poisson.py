"""GPU nonlinear Poisson solve against the oracle (same discrete equation)."""
import numpy as np
import pytest

from helpers import axes_for, grain_radius, rel_l2

pytestmark = pytest.mark.gpu


def _problem(kind, n, Q=-0.2):
    import kielflow_amd as kf
    axes, oaxes = axes_for(kind, n)
    a = grain_radius(kind, n)
    ln_n, nd, mask = kf.setup_host(axes, seed=3, amp=1e-3, Q=Q, a=a)
    rho = np.exp(ln_n) + nd
    return axes, oaxes, rho


@pytest.mark.parametrize("kind,n", [("eqst", 32), ("eqst", 48), ("sist", 32), ("hyst", 24)])
def test_newton_mg_fp64_matches_oracle(ctx, oracle, kind, n):
    import kielflow_amd as kf
    print("poisson fp64", kind, n, flush=True)
    axes, oaxes, rho = _problem(kind, n)
    og = oracle.Grid(*oaxes)
    ref, it, res = og.poisson_fas(np.zeros(og.rshape), rho, 1e-13, 400)
    assert it >= 0
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=60)
    dphi = g.rfield()
    drho = ctx.to_device(rho)
    iters, resid, code = p.solve(dphi, drho)
    assert resid <= 1e-13
    phi = dphi.get()
    assert og.poisson_residual(phi, rho)[0] < 1e-12
    assert np.abs(phi - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
    # warm start converges in fewer iterations
    it2, res2, _ = p.solve(dphi, drho)
    assert it2 <= 1
    p.close()
    g.close()


def test_newton_mg_mixed_precision(ctx, oracle):
    import kielflow_amd as kf
    print("poisson mixed", flush=True)
    axes, oaxes, rho = _problem("eqst", 64, Q=-0.3)
    g = kf.Grid(ctx, axes)
    p64 = kf.Poisson(g, kf.KF_FP64, tol=1e-12, max_it=60)
    pm = kf.Poisson(g, kf.KF_MIXED, tol=1e-10, max_it=60)
    drho = ctx.to_device(rho)
    a, b = g.rfield(), g.rfield()
    p64.solve(a, drho)
    it, res, code = pm.solve(b, drho)
    assert res <= 1e-10
    assert np.abs(a.get() - b.get()).max() < 1e-8 * np.abs(a.get()).max()


@pytest.mark.parametrize("kind,n,prec", [("eqst", 32, "fp64"), ("eqst", 48, "mixed"), ("sist", 32, "fp64"),
                                         ("hyst", 24, "fp64"), ("eqst", 64, "mixed")])
def test_wavefront_coarsest_smoothing_gives_identical_results(ctx, monkeypatch, kind, n, prec):
    """The coarsest level (4^3 red-black, 3^3 damped Jacobi) smoothed by one
    wavefront gives the solve of the whole workgroup (KF_COARSE_WAVE=0) bit
    for bit, iteration count included."""
    import kielflow_amd as kf
    print("coarsest wavefront", kind, n, prec, flush=True)
    axes, _, rho = _problem(kind, n)
    out = []
    for wave in ("1", "0"):
        monkeypatch.setenv("KF_COARSE_WAVE", wave)
        g = kf.Grid(ctx, axes)
        p = kf.Poisson(g, kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED, tol=1e-12 if prec == "fp64" else 1e-10,
                       max_it=60)
        dphi = g.rfield()
        it, res, code = p.solve(dphi, ctx.to_device(rho))
        out.append((it, res, dphi.get()))
        p.close()
        g.close()
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("kind,n,prec", [("sist", 32, "fp64"), ("hyst", 24, "fp64"), ("sist", 32, "mixed")])
def test_preconditioner_on_the_krylov_vectors_gives_identical_results(ctx, monkeypatch, kind, n, prec):
    """The V-cycle that preconditions BiCGStab on stretched grids reading and
    writing the Krylov vectors directly gives the solve of the version that
    copies them through the finest level's b and x (KF_PRECOND_COPY=1) bit
    for bit."""
    import kielflow_amd as kf
    print("preconditioner in place", kind, n, prec, flush=True)
    axes, _, rho = _problem(kind, n)
    out = []
    for copy in ("0", "1"):
        monkeypatch.setenv("KF_PRECOND_COPY", copy)
        g = kf.Grid(ctx, axes)
        p = kf.Poisson(g, kf.KF_FP64 if prec == "fp64" else kf.KF_MIXED, tol=1e-12 if prec == "fp64" else 1e-10,
                       max_it=60)
        dphi = g.rfield()
        it, res, code = p.solve(dphi, ctx.to_device(rho))
        out.append((it, res, dphi.get()))
        p.close()
        g.close()
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_not_converged_is_an_error(ctx):
    import kielflow_amd as kf
    axes, oaxes, rho = _problem("eqst", 16)
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, tol=1e-15, max_it=1)
    it, res, code = p.solve(g.rfield(), ctx.to_device(rho), raise_on_error=False)
    assert code == kf.api.KF_ERR_NOT_CONVERGED


@pytest.mark.parametrize("iters", [1, 3, 25])
def test_reference_nlj_iterates_match_oracle(ctx, oracle, iters):
    """The reference iteration (Jacobi + damped Newton + Carre SOR), iterate by
    iterate, at 16^3."""
    import kielflow_amd as kf
    print("nlj", iters, flush=True)
    axes, oaxes, rho = _problem("sist", 16)
    og = oracle.Grid(*oaxes)
    ref, it_o, st = og.nlj_solve(np.zeros(og.rshape), rho, max_iter=iters, tol=1e-30)
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, mode=kf.POISSON_NLJ, tol=1e-30, max_it=iters)
    dphi = g.rfield()
    it, res, code = p.solve(dphi, ctx.to_device(rho))
    assert it == it_o == iters
    assert np.abs(dphi.get() - ref).max() < 1e-13


def test_reference_nlj_converged_count(ctx, oracle):
    import kielflow_amd as kf
    axes, oaxes, rho = _problem("eqst", 16)
    og = oracle.Grid(*oaxes)
    ref, it_o, st = og.nlj_solve(np.zeros(og.rshape), rho, max_iter=500, tol=1e-6)
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, mode=kf.POISSON_NLJ, tol=1e-6, max_it=500)
    dphi = g.rfield()
    it, res, code = p.solve(dphi, ctx.to_device(rho))
    assert st.converged == 1 and it == it_o
    assert np.abs(dphi.get() - ref).max() < 1e-12


def test_reference_cascade_matches_oracle(ctx, oracle):
    import kielflow_amd as kf
    print("cascade", flush=True)
    axes, oaxes, rho = _problem("sist", 32)
    og = oracle.Grid(*oaxes)
    phi0 = np.zeros(og.rshape)
    ref, it_o, st = og.mg_cascade(phi0, rho)
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, mode=kf.POISSON_CASCADE)
    dphi = g.rfield()
    it, res, code = p.solve(dphi, ctx.to_device(rho))
    assert np.abs(dphi.get() - ref).max() < 1e-11


@pytest.mark.parametrize("program", [((-2, 1, 1), (6, -1, -1), (1e-3, 1e-4, 1e-7)),
                                     ((-1, 0, -1, 1, 1), (4, 3, 8, -1, -1), (1e-2, 1e-2, 1e-3, 1e-4, 1e-8))])
def test_reference_cascade_level_program_matches_oracle(ctx, oracle, program):
    """solver_poisson_multigrid::set_level_control (cpp:82-147): a caller's
    level steps, iteration counts and tolerances are the ones that run."""
    import kielflow_amd as kf
    from helpers import axes_for
    lvl, steps, tol = program
    print("cascade program", lvl, flush=True)
    axes, oaxes = axes_for("sist", 32)
    og = oracle.Grid(*oaxes)
    rho = 1.0 + og.gauss(-0.1, 1.0)
    ref, it_o, st = og.mg_cascade(np.zeros(og.rshape), rho, lvl=lvl, steps=steps, tol=tol)
    g = kf.Grid(ctx, axes)
    p = kf.Poisson(g, kf.KF_FP64, mode=kf.POISSON_CASCADE)
    p.set_cascade(lvl, steps, tol)
    dphi = g.rfield()
    it, res, code = p.solve(dphi, ctx.to_device(rho))
    assert np.abs(dphi.get() - ref).max() < 1e-11
    # the default program differs: the result is not the default cascade's
    ref_default, _, _ = og.mg_cascade(np.zeros(og.rshape), rho)
    assert np.abs(ref_default - ref).max() > 1e-9
    with pytest.raises(kf.api.KfError):
        p.set_cascade((1, -1), (1, 1), (1e-3, 1e-3))      # above the base level
    with pytest.raises(kf.api.KfError):
        p.set_cascade((-1,), (1,), (1e-3,))               # does not end on the base level
    p.close()
    g.close()
