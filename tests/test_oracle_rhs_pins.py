"""Closed-form pins of the oracle's nonlinear right-hand side, penalization
blend, Poisson cascade and drag objective (CPU only).

Every expected value below comes from the PDE that rhs_standard::solve
discretises (rhs_standard.cpp:89-566), evaluated analytically on
band-limited single modes, not from the oracle's code:

  d/dt u_a   = -[ tau u_a + M d_x u_a + d_a(filt(phi) + ln n / theta) + u . grad u_a ]
  d/dt ln n  = -[ SV ln n + M d_x ln n + div u + u . grad ln n ]

All input modes lie inside the 2/3 ball, so dealiasing is the identity on
them, and every product has modes below the Nyquist index, so the products
computed in real space are exact.  A sign, factor, index or operand slip in
any one term changes the result by far more than the tolerances here.
"""
import numpy as np
import pytest

L = 16.0
N = 16


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _coords(O, g):
    x = np.array([O.val_at(g.axes[0], i) for i in range(N)])
    return np.meshgrid(x, x, x, indexing="ij")


def _zero_state(g):
    z = np.zeros(g.cshape, complex)
    return [z.copy() for _ in range(4)]


def _lap_symbol(k, h):
    """-(symbol) of the 7-point Laplacian on one uniform axis:
    L_h cos(k x) = -(2/h^2)(1 - cos(k h)) cos(k x)."""
    return (2.0 / h ** 2) * (1.0 - np.cos(k * h))


def _phi_filter(kh):
    """exp(-2000 k^15) of rhs_standard.cpp:220 at normalised |k|."""
    return np.exp(-2000.0 * kh ** 15)


K1 = 2 * np.pi / L           # wavenumber of mode index 1


def test_advection_closed_form(oracle):
    """u . grad u_a for a divergence-free single-mode velocity (ln n = 0,
    n_d = 0, so phi = 0 and the E and pressure terms vanish)."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, Y, Z = _coords(O, g)
    k = K1
    a, b, c, d = 0.3, -0.2, 0.25, 0.15
    ux = a * np.sin(k * Y) + d * np.cos(k * Z)
    uy = b * np.cos(k * X)
    uz = c * np.sin(k * X)
    grad = {  # analytic partial derivatives d_b u_a
        (0, 0): 0 * X, (0, 1): a * k * np.cos(k * Y), (0, 2): -d * k * np.sin(k * Z),
        (1, 0): -b * k * np.sin(k * X), (1, 1): 0 * X, (1, 2): 0 * X,
        (2, 0): c * k * np.cos(k * X), (2, 1): 0 * X, (2, 2): 0 * X,
    }
    U = (ux, uy, uz)
    tau, M = 0.1, 0.7
    r = O.Rhs(g, M=M, tau=tau, theta=100.0, sv_on=False, pen_on=False, poisson_tol=1e-14)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
    st = _zero_state(g)
    for i in range(3):
        st[i] = g.r2c(U[i])
    out, s = r.eval(0.0, st)
    assert s == 0
    for i in range(3):
        adv = sum(U[j] * grad[(i, j)] for j in range(3))
        expect = -(tau * U[i] + M * grad[(i, 0)] + adv)
        got = g.c2r(out[i])
        assert np.abs(got - expect).max() < 1e-13, i
    # div u = 0 and ln n = 0: the density equation is at rest
    assert np.abs(g.c2r(out[3])).max() < 1e-14


def test_advection_self_product(oracle):
    """u_x = a sin(kx): u_x d_x u_x = (a^2 k / 2) sin(2kx); div u = a k cos(kx)."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, _, _ = _coords(O, g)
    k, a = 2 * K1, 0.4
    r = O.Rhs(g, M=0.0, tau=0.0, theta=100.0, sv_on=False, pen_on=False, poisson_tol=1e-14)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
    st = _zero_state(g)
    st[0] = g.r2c(a * np.sin(k * X))
    out, s = r.eval(0.0, st)
    assert np.abs(g.c2r(out[0]) + 0.5 * a * a * k * np.sin(2 * k * X)).max() < 1e-14
    assert np.abs(g.c2r(out[3]) + a * k * np.cos(k * X)).max() < 1e-14


@pytest.mark.parametrize("sv_on", [False, True])
def test_continuity_closed_form(oracle, sv_on):
    """d/dt ln n = -(SV ln n + M d_x ln n + div u + u . grad ln n)."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, Y, Z = _coords(O, g)
    k = K1
    eps, c, a, M = 2e-3, 0.3, 0.2, 0.9
    ln_n = eps * np.sin(k * X) + eps * np.cos(k * Z)
    ux = np.full(g.rshape, c)
    uz = a * np.sin(k * Z)
    dln = (eps * k * np.cos(k * X), 0 * X, -eps * k * np.sin(k * Z))
    div_u = a * k * np.cos(k * Z)
    r = O.Rhs(g, M=M, tau=0.1, theta=100.0, sv_on=sv_on, pen_on=False, poisson_tol=1e-14)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
    st = _zero_state(g)
    st[0] = g.r2c(ux)
    st[2] = g.r2c(uz)
    st[3] = g.r2c(ln_n)
    out, s = r.eval(0.0, st)
    assert s == 0
    expect = -(M * dln[0] + div_u + ux * dln[0] + uz * dln[2])
    if sv_on:
        # SV = 0.5 (1 - exp(-8 khat^4)) at the index-1 modes of x and z
        # (rhs_standard.cpp:51-77), khat = 1/(N/2)
        sv1 = 0.5 * (1 - np.exp(-8 * (1.0 / (N // 2)) ** 4))
        expect = expect - sv1 * ln_n
    got = g.c2r(out[3])
    assert np.abs(got - expect).max() < 1e-14 + 1e-12 * np.abs(expect).max()


def test_pressure_term_closed_form(oracle):
    """ln n = eps cos(kx) with n_d = 1 - n (so rho = 1 and phi = 0):
    d/dt u_x = -d_x ln n / theta = eps k sin(kx) / theta."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, Y, _ = _coords(O, g)
    k, eps = K1, 1e-3
    for theta in (100.0, 7.0):
        ln_n = eps * np.cos(k * X) + eps * np.sin(2 * k * Y)
        nd = 1.0 - np.exp(ln_n)
        r = O.Rhs(g, M=0.0, tau=0.1, theta=theta, sv_on=False, pen_on=False, poisson_tol=1e-15)
        r.set_fields(np.zeros(g.rshape), nd, np.zeros(g.rshape))
        st = _zero_state(g)
        st[3] = g.r2c(ln_n)
        out, s = r.eval(0.0, st)
        assert s == 0
        assert np.abs(r.phi()).max() < 1e-15
        ex = eps * k * np.sin(k * X) / theta
        ey = -eps * 2 * k * np.cos(2 * k * Y) / theta
        assert np.abs(g.c2r(out[0]) - ex).max() < 1e-16 + 1e-12 * np.abs(ex).max()
        assert np.abs(g.c2r(out[1]) - ey).max() < 1e-16 + 1e-12 * np.abs(ey).max()
        assert np.abs(g.c2r(out[2])).max() < 1e-17
        assert np.abs(g.c2r(out[3])).max() < 1e-17


def test_efield_manufactured_potential(oracle):
    """ln n = 0, n_d = exp(phi*) - 1 - L_h phi*: the Poisson solve returns
    phi*, and d/dt u_a = -d_a filt(phi*) with the filter of
    rhs_standard.cpp:220 and its z normalisation (operations.hpp:23)."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, Y, Z = _coords(O, g)
    h = L / N
    eps = 1e-3
    kx, ky, kz = K1, 2 * K1, 3 * K1
    phis = eps * (np.cos(kx * X) + np.cos(ky * Y) + np.cos(kz * Z))
    Lphi = -eps * (_lap_symbol(kx, h) * np.cos(kx * X) + _lap_symbol(ky, h) * np.cos(ky * Y)
                   + _lap_symbol(kz, h) * np.cos(kz * Z))
    nd = np.exp(phis) - 1.0 - Lphi
    r = O.Rhs(g, M=0.0, tau=0.1, theta=100.0, sv_on=False, pen_on=False, poisson_tol=1e-15,
              poisson_max=400)
    r.set_fields(np.zeros(g.rshape), nd, np.zeros(g.rshape))
    out, s = r.eval(0.0, _zero_state(g))
    assert s == 0
    assert np.abs(r.phi() - phis).max() < 1e-13
    kmax = 2 * np.pi * (N // 2) / L
    kzq = 2 * np.pi * ((N // 2 + 1) // 2) / L     # the kz normalisation quirk
    fx = _phi_filter(kx / kmax)
    fy = _phi_filter(ky / kmax)
    fz = _phi_filter(kz / kzq)
    assert fz < 0.999  # the quirk matters at this mode
    ex = eps * kx * fx * np.sin(kx * X)
    ey = eps * ky * fy * np.sin(ky * Y)
    ez = eps * kz * fz * np.sin(kz * Z)
    for i, e in enumerate((ex, ey, ez)):
        assert np.abs(g.c2r(out[i]) - e).max() < 1e-12 * eps, i
    assert np.abs(g.c2r(out[3])).max() < 1e-15


def test_penalization_blend_uniform_fields(oracle):
    """Uniform mask H = h, uniform u = c and ln n = c_n, n_d = 0: before the
    blend B_a = tau c_a, B_n = 0; after it (rhs_standard.cpp:459-533)
      P_x = 2 h (c_x + M h) + (1 - h) tau c_x + b c_x
      P_y = 2 h c_y + (1 - h) tau c_y + b c_y          (same for z)
      P_n = 2 h (c_n + 10) + b c_n
    with b = cos^2(3 pi x / L) for x < -L/6; the output is -D23 FFT(P)."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, _, _ = _coords(O, g)
    hH, M, tau = 0.3, 1.25, 0.1
    c = (0.05, -0.03, 0.02)
    cn = 0.01
    b = np.where(X < -L / 6, np.cos(3 * np.pi * X / L) ** 2, 0.0)
    r = O.Rhs(g, M=M, tau=tau, theta=100.0, sv_on=False, pen_on=True, poisson_tol=1e-15)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.full(g.rshape, hH))
    st = [g.r2c(np.full(g.rshape, v)) for v in c] + [g.r2c(np.full(g.rshape, cn))]
    out, s = r.eval(0.0, st)
    assert s == 0
    assert np.abs(r.phi() - cn).max() < 1e-15
    P = [2 * hH * (c[0] + M * hH) + (1 - hH) * tau * c[0] + b * c[0],
         2 * hH * c[1] + (1 - hH) * tau * c[1] + b * c[1],
         2 * hH * c[2] + (1 - hH) * tau * c[2] + b * c[2],
         2 * hH * (cn + 10.0) + b * cn]
    # 2/3 rule on this grid: only x-modes are present, kept for |i| < N/3
    idx = np.fft.fftfreq(N, 1.0 / N)
    keep = (np.abs(idx) / (N // 2) < 2.0 / 3.0)[:, None, None]
    for i in range(4):
        expect = -np.fft.rfftn(P[i]) * keep
        assert rel(out[i], expect) < 1e-13, i


def test_drag_objective_manufactured(oracle):
    """phi* = eps sin(kx), n_d = a cos(kx), and ln n chosen so that the
    Poisson solution is phi*: J = -sum n_d d_x phi* dV = -a eps k L^3 / 2."""
    O = oracle
    g = O.Grid.cube(N, L)
    X, _, _ = _coords(O, g)
    h = L / N
    k, eps, a = K1, 1e-3, 1e-3
    phis = eps * np.sin(k * X)
    Lphi = -eps * _lap_symbol(k, h) * np.sin(k * X)
    nd = a * np.cos(k * X) * np.ones(g.rshape)
    ln_n = np.log(np.exp(phis) - Lphi - nd)
    r = O.Rhs(g, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=False, poisson_tol=1e-15,
              poisson_max=400)
    r.set_fields(np.zeros(g.rshape), nd, np.zeros(g.rshape))
    dV = np.full(g.rshape, h ** 3)
    J, _ = r.objective_tl(g.r2c(ln_n), None, dV)
    expect = -a * eps * k * L ** 3 / 2
    assert J == pytest.approx(expect, rel=1e-10)
    # only d_x enters: the same potential along y gives no x drag
    Y = _coords(O, g)[1]
    phis_y = eps * np.sin(k * Y)
    ln_y = np.log(np.exp(phis_y) + eps * _lap_symbol(k, h) * np.sin(k * Y) - nd)
    J2, _ = r.objective_tl(g.r2c(ln_y), None, dV)
    assert abs(J2) < 1e-8 * abs(expect)


def test_cascade_uniform_density_is_exact(oracle):
    """Injection and trilinear prolongation keep a uniform field, and the
    NLJ fixed point of a uniform rho = c is ln c at every level."""
    g = oracle.Grid.cube(32)
    rho = np.full(g.rshape, 1.3)
    phi, it, st = g.mg_cascade(np.zeros(g.rshape), rho, tol=(1e-13, 1e-13, 1e-13, 1e-13),
                               steps=(-1, -1, -1, -1))
    assert it > 0
    assert np.abs(phi - np.log(1.3)).max() < 1e-12


def test_cascade_ends_on_the_finest_discrete_solution(oracle):
    """With the last level run to a tight tolerance, the cascade's end state
    is the fixed point of the finest discrete equation (the FAS solution),
    which the linearised closed form pins separately."""
    g = oracle.Grid.cube(32)
    rho = 1.0 + g.gauss(-0.1, 1.0)
    ref, _, res = g.poisson_fas(np.zeros(g.rshape), rho, 1e-14)
    phi, it, st = g.mg_cascade(np.zeros(g.rshape), rho, steps=(0, -1, -1, -1),
                               tol=(0.01, 5e-4, 1e-4, 1e-13))
    assert st.converged == 1
    assert np.abs(phi - ref).max() < 1e-10 * np.abs(ref).max() + 1e-13


def test_euler_amplification_of_a_linear_mode(oracle):
    """euler_method (euler_method.cpp:74-88) on u_z = eps sin(kx), n = 1:
    du_z/dt = -tau u_z - M d_x u_z, so one step multiplies each mode by 1 + z,
    z = (-tau - i M k) dt."""
    g = oracle.Grid.cube(N, L)
    X, _, _ = _coords(oracle, g)
    tau, M, dt = 0.1, 1.3, 0.01
    r = oracle.Rhs(g, M=M, tau=tau, pen_on=False, sv_on=False, poisson_tol=1e-14)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
    st = _zero_state(g)
    st[2] = g.r2c(1e-3 * np.sin(2 * K1 * X))
    F0 = st[2].copy()
    t, s = r.euler(dt, 0.0, st)
    assert s == 0 and t == dt
    kx = np.array([oracle.k_val_at(g.axes[0], i) for i in range(N)])[:, None, None]
    amp = 1 + (-tau - 1j * M * kx) * dt * np.ones(g.cshape)
    sel = np.abs(F0) > 1e-6
    assert sel.sum() == 2
    assert np.allclose(st[2][sel], amp[sel] * F0[sel], rtol=1e-14, atol=0)
    assert max(np.abs(st[i]).max() for i in (0, 1, 3)) < 1e-15
