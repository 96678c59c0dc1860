"""Pins of the oracle's equilibrium start (the reference frontend,
frontend.cpp:133-198, 291-310, with the __FRONTEND__ NLJ of
solver_poisson_jacobi_nlin.cpp:192-195): the discrete equation
L_h phi = exp(phi) - exp(-theta phi) - n_d in closed form.  CPU only.

The reference's local Newton leaves -theta exp(-theta x) out of the
derivative (cpp:193-195); it converges while the diagonal D = 6/h^2
dominates theta, as on the frontend's grid (h = 6/128).  The boxes here
have h = 1/8 (L = 2, n = 16)."""
import numpy as np
import pytest
from scipy.optimize import brentq

L = 2.0


def test_uniform_charge_gives_the_scalar_root(oracle):
    """n_d = c uniform: phi solves exp(phi) - exp(-theta phi) = c.  (The
    update test compares with sup - inf, which is 0 on a uniform field, so
    the iteration runs to its count.)"""
    g = oracle.Grid.cube(16, L)
    for theta, c in ((10.0, 0.5), (30.0, -0.2), (3.0, 1.7)):
        root = brentq(lambda x: np.exp(x) - np.exp(-theta * x) - c, -5.0, 5.0, xtol=1e-16, rtol=1e-15)
        phi, it, st = g.nlj_ion(np.zeros(g.rshape), np.full(g.rshape, c), theta, max_iter=3000, tol=1e-14)
        assert np.abs(phi - root).max() < 1e-10, (theta, c)


def test_linear_response_of_a_single_mode(oracle):
    """n_d = eps cos(kx): phi = eps cos(kx) / (lambda_h + 1 + theta) to O(eps^2),
    lambda_h = (2/h^2)(1 - cos kh) the 7-point symbol."""
    n = 16
    g = oracle.Grid.cube(n, L)
    h = L / n
    x = np.array([oracle.val_at(g.axes[0], i) for i in range(n)])
    eps, theta = 1e-6, 5.0
    for j in (1, 3):
        k = 2 * np.pi * j / L
        nd = eps * np.cos(k * x)[:, None, None] * np.ones(g.rshape)
        phi, it, st = g.nlj_ion(np.zeros(g.rshape), nd, theta, max_iter=20000, tol=1e-15)
        lam = (2 / h ** 2) * (1 - np.cos(k * h))
        expect = eps * np.cos(k * x)[:, None, None] / (lam + 1 + theta)
        assert np.abs(phi - expect).max() < 1e-11 * 1e3 * eps   # O(eps^2) and the update test


def test_no_charge_gives_the_homogeneous_plasma(oracle):
    g = oracle.Grid.cube(16, L)
    for cascade in (True, False):
        phi, ni, it = g.frontend_init(30.0, np.zeros(g.rshape), cascade=cascade, tol=1e-12)
        assert it >= 0
        assert np.abs(phi).max() == 0.0 and np.abs(ni - 1.0).max() < 1e-15


def test_frontend_output_is_the_filtered_boltzmann_density(oracle):
    """n_i = exp(-theta phi) through exp(-2000 k^15): with the unfiltered
    potential of the same solve, filter(FFT(exp(-theta phi))) is FFT(n_i)."""
    g = oracle.Grid.cube(16, L)
    theta = 10.0
    nd = g.gauss(-1e-3, 0.25)
    phi_raw, it, st = g.nlj_ion(np.zeros(g.rshape), nd, theta, max_iter=3000, tol=1e-12)
    phi, ni, it2 = g.frontend_init(theta, nd, cascade=False, tol=1e-12, max_it=3000)
    assert it2 == it
    a, b = g.r2c(ni), g.filter_phi(g.r2c(np.exp(-theta * phi_raw)))
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
    a, b = g.r2c(phi), g.filter_phi(g.r2c(phi_raw))
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
    # the grain's negative charge pulls the potential down and the ions in
    assert phi_raw.min() < 0 and ni.max() > 1.0
