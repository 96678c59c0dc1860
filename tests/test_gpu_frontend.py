"""The equilibrium start of the reference frontend (frontend.cpp:133-198,
291-310) on the GPU against the oracle: the reference cascade with the
Boltzmann-ion NLJ (solver_poisson_jacobi_nlin.cpp:192-195) iterate for
iterate, and Newton-MG with the exact Jacobian against the oracle's NLJ run
to its fixed point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L_FE = 6.0                                   # frontend.cpp:63-65
M_FE = L_FE * (1.0 - 0.3) / (2 * np.pi)      # frontend.cpp:285-287
THETA = 30.0                                 # frontend.cpp:71
Q_FE = -11498.5 * 8.1720e-06                 # frontend.cpp:74, 137-138
A_FE = 0.15                                  # frontend.cpp:77


def _setup(n):
    import kielflow_amd as kf
    from oracle import oracle as O
    axes = [kf.axis(kf.SIST, n, -0.5 * L_FE, L_FE, M_FE) for _ in range(3)]
    oaxes = [O.axis(O.SIST, n, -0.5 * L_FE, L_FE, M_FE) for _ in range(3)]
    _, nd, _ = kf.setup_host(axes, seed=0, amp=0.0, Q=Q_FE, a=A_FE)
    return axes, O.Grid(*oaxes), nd


def test_frontend_cascade_matches_oracle(ctx):
    """At 64^3.  (The reference's local Newton leaves the ion term out of
    the derivative, so a cascade level that iterates must have D = 2 sum
    1/h^2 well above theta: at 32^3 the cascade's 8^3 level diverges, in the
    oracle as in the reference.)"""
    import kielflow_amd as kf
    n = 64
    print("frontend cascade", n, flush=True)
    axes, og, nd = _setup(n)
    ref_phi, ref_ni, it_o = og.frontend_init(THETA, nd, cascade=True)
    assert it_o > 0
    g = kf.Grid(ctx, axes)
    phi, ni = g.rfield(), g.rfield()
    it = g.frontend_equilibrium(THETA, ctx.to_device(nd), phi, ni, mode=kf.POISSON_CASCADE)
    e_phi = np.abs(phi.get() - ref_phi).max() / np.abs(ref_phi).max()
    e_ni = np.abs(ni.get() - ref_ni).max() / np.abs(ref_ni).max()
    print(f"  iterations {it}/{it_o}, phi {e_phi:.2e}, ni {e_ni:.2e}", flush=True)
    assert e_phi < 1e-11 and e_ni < 1e-11
    # a negative grain: potential well and ion cloud
    assert ref_phi.min() < 0 and ref_ni.max() > 1.0
    g.close()


def test_frontend_newton_mg_matches_the_converged_oracle(ctx):
    import kielflow_amd as kf
    n = 32
    print("frontend newton-mg", n, flush=True)
    axes, og, nd = _setup(n)
    ref_phi, ref_ni, it_o = og.frontend_init(THETA, nd, cascade=False, tol=1e-15, max_it=20000)
    g = kf.Grid(ctx, axes)
    phi, ni = g.rfield(), g.rfield()
    it = g.frontend_equilibrium(THETA, ctx.to_device(nd), phi, ni, mode=kf.POISSON_NEWTON_MG, tol=1e-13,
                                max_iterations=60)
    e_phi = np.abs(phi.get() - ref_phi).max() / np.abs(ref_phi).max()
    e_ni = np.abs(ni.get() - ref_ni).max() / np.abs(ref_ni).max()
    print(f"  Newton iterations {it}, oracle NLJ sweeps {it_o}, phi {e_phi:.2e}, ni {e_ni:.2e}", flush=True)
    assert e_phi < 1e-10 and e_ni < 1e-10
    g.close()


def test_frontend_state_drops_into_the_step(ctx):
    """ln n_i of the equilibrium start is the state of the main loop
    (kielflow.cpp:172-182): one FP64 RK4 step from it matches the oracle."""
    import kielflow_amd as kf
    from helpers import rel_l2
    n = 16
    axes, og, nd = _setup(n)
    g = kf.Grid(ctx, axes)
    phi, ni = g.rfield(), g.rfield()
    g.frontend_equilibrium(THETA, ctx.to_device(nd), phi, ni, mode=kf.POISSON_NEWTON_MG, tol=1e-13)
    ln_n = np.log(ni.get())
    state = [np.zeros(og.cshape, complex) for _ in range(3)] + [og.r2c(ln_n)]
    _, _, mask = kf.setup_host(axes, seed=0, amp=0.0, Q=Q_FE, a=A_FE)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=80)
    rhs = kf.Rhs(g, ps, M=0.6, tau=0.1, theta=THETA, sv_on=True, pen_on=False)
    phi2 = g.rfield()
    phi2.upload(phi.get())
    rhs.set_fields(phi2, ctx.to_device(nd), ctx.to_device(mask))
    dF = [ctx.to_device(s) for s in state]
    kf.Rk4(rhs, 1e-3).step(dF, 1)
    from oracle import oracle as O
    orhs = O.Rhs(og, M=0.6, tau=0.1, theta=THETA, sv_on=True, pen_on=False, poisson_tol=1e-14, poisson_max=400)
    orhs.set_fields(phi.get(), nd, mask)
    ost = [s.copy() for s in state]
    orhs.rk4(1e-3, 0.0, ost)
    for a in range(4):
        if np.abs(ost[a]).max() > 0:
            assert rel_l2(dF[a].get(), ost[a]) < 1e-11, a
