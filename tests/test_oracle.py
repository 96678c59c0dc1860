"""Pins of the CPU FP64 oracle against golden vectors, closed forms and
invariants (the reference ships no tests or fixtures).  CPU only."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_numpy.npz")
L = 16.0


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


# ----------------------------------------------------------------- FFT
def test_fft_matches_numpy_golden(oracle):
    gold = np.load(GOLD)
    for key in gold.files:
        if not key.startswith("x_"):
            continue
        shape = tuple(int(v) for v in key[2:].split("x"))
        g = oracle.Grid(*[oracle.axis(oracle.EQST, n, -8.0, L) for n in shape])
        X = g.r2c(gold[key])
        assert rel(X, gold["X_" + key[2:]]) < 1e-14
        assert rel(g.c2r(X), gold[key]) < 1e-14


def test_fft_brute_force_dft(oracle):
    rng = np.random.default_rng(1)
    shape = (4, 6, 8)
    x = rng.standard_normal(shape)
    g = oracle.Grid(*[oracle.axis(oracle.EQST, n, -8.0, L) for n in shape])
    nx, ny, nz = shape
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    ey = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    ez = np.exp(-2j * np.pi * np.outer(np.arange(nz // 2 + 1), np.arange(nz)) / nz)
    X = np.einsum("ai,bj,ck,ijk->abc", ex, ey, ez, x)
    assert rel(g.r2c(x), X) < 1e-14


def test_fft1d_mixed_radix(oracle):
    rng = np.random.default_rng(2)
    for n in (2, 3, 4, 6, 8, 12, 24, 48, 96, 5, 7):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert rel(oracle.fft1d(x), np.fft.fft(x)) < 1e-14


def test_c2r_is_transform_of_hermitian_projection(oracle):
    """kz=0 and kz=Nz/2 planes enter C2R through (F + conj F(-k))/2."""
    rng = np.random.default_rng(3)
    shape = (6, 4, 8)
    g = oracle.Grid(*[oracle.axis(oracle.EQST, n, -8.0, L) for n in shape])
    X = rng.standard_normal(g.cshape) + 1j * rng.standard_normal(g.cshape)
    P = X.copy()
    for kz in (0, shape[2] // 2):
        pl = X[:, :, kz]
        flip = np.roll(pl[::-1, ::-1], 1, axis=(0, 1))  # F(-kx, -ky)
        P[:, :, kz] = 0.5 * (pl + np.conj(flip))
    ref = np.fft.irfftn(P, s=shape, axes=(0, 1, 2))
    assert rel(g.c2r(X), ref) < 1e-14
    assert rel(g.c2r(P), ref) < 1e-14


# ----------------------------------------------------------------- axes
def test_axes_metric_matches_map(oracle):
    """dS(i) is the derivative of the index -> x map (checked by centred
    differences on a fine grid) for SiSt and the HySt fix."""
    for t, m in ((oracle.SIST, 16 * (1 - 1 / 3) / (2 * np.pi)), (oracle.HYST, 3.0)):
        n = 2048
        a = oracle.axis(t, n, -8.0, L, m)
        h = L / n
        for i in (3, 100, 700, 1024, 1500, 2040):
            fd = (oracle.val_at(a, i + 1) - oracle.val_at(a, i - 1)) / (2 * h)
            assert abs(fd - oracle.dS(a, i)) < 1e-4 * max(1.0, abs(fd))
    a = oracle.axis(oracle.EQST, 32, -8.0, L)
    assert oracle.k_val_at(a, 16) == pytest.approx(2 * np.pi * 16 / L)   # Nyquist +pi n/L
    assert oracle.k_val_at(a, 17) == pytest.approx(2 * np.pi * (17 - 32) / L)
    assert oracle.val_at(a, 5) == pytest.approx(-8.0 + 5 * 0.5)


def test_sist_periodic_extension(oracle):
    a = oracle.axis(oracle.SIST, 32, -8.0, L, 1.2)
    for i in range(-3, 3):
        assert oracle.val_at(a, i + 32) == pytest.approx(oracle.val_at(a, i) + L, abs=1e-12)


# ----------------------------------------------------------------- spectral operators
def test_derivative_exact_on_resolved_modes(oracle):
    n = 24
    g = oracle.Grid.cube(n)
    x = np.array([oracle.val_at(g.axes[0], i) for i in range(n)])
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    for j in (1, 3, 7):
        f = np.sin(2 * np.pi * j * X / L) * np.cos(2 * np.pi * 2 * Z / L)
        df = (2 * np.pi * j / L) * np.cos(2 * np.pi * j * X / L) * np.cos(2 * np.pi * 2 * Z / L)
        out = g.c2r(g.deriv(0, g.r2c(f)))
        assert np.abs(out - df).max() < 1e-12


def test_derivative_nyquist_convention(oracle):
    """k at index N/2 is +pi N/L (axis_EqSt.cpp:60): i k F at the Nyquist."""
    n = 8
    g = oracle.Grid.cube(n)
    F = np.zeros(g.cshape, complex)
    F[n // 2, 0, 0] = 1.0
    D = g.deriv(0, F)
    assert D[n // 2, 0, 0] == pytest.approx(1j * np.pi * n / L)


@pytest.mark.parametrize("kind", ["sist", "hyst"])
def test_stretched_derivative_chain_rule(oracle, kind):
    """On stretched axes d/dx f = f_xi / S'(xi), converging spectrally."""
    errs = []
    for n in (32, 64):
        if kind == "sist":
            ax = oracle.axis(oracle.SIST, n, -8.0, L, 16 * (1 - 0.5) / (2 * np.pi))
        else:
            ax = oracle.axis(oracle.HYST, n, -8.0, L, 1.0)
        eq = oracle.axis(oracle.EQST, 8, -8.0, L)
        g = oracle.Grid(ax, eq, eq)
        xi = -8.0 + np.arange(n) * L / n
        f = np.sin(2 * np.pi * xi / L)[:, None, None] * np.ones((n, 8, 8))
        dS = np.array([oracle.dS(ax, i) for i in range(n)])
        expect = ((2 * np.pi / L) * np.cos(2 * np.pi * xi / L) / dS)[:, None, None] * np.ones((n, 8, 8))
        out = g.c2r(g.deriv(0, g.r2c(f)))
        errs.append(np.abs(out - expect).max())
    assert errs[1] < errs[0]
    assert errs[1] < (1e-6 if kind == "sist" else 5e-2)


def test_dealias_removes_aliasing_of_products(oracle):
    """Products of two fields inside the 2/3 ball, dealiased, equal the
    dealiased exact product (computed on a 2x padded grid)."""
    n = 12
    g = oracle.Grid.cube(n)
    g2 = oracle.Grid.cube(2 * n)
    rng = np.random.default_rng(4)
    d23 = g.d23_table()

    def band(seed):
        X = rng.standard_normal(g.cshape) + 1j * rng.standard_normal(g.cshape)
        X = X * d23
        return g.c2r(X) * 0 + g.c2r(X)

    f, h = band(0), band(1)
    prod = g.dealias23(g.r2c(f * h))

    def pad(x):
        Xs = np.fft.fftn(x)
        P = np.zeros((2 * n,) * 3, complex)
        idx = np.r_[0:n // 2, -n // 2 + 1:0]
        P[np.ix_(idx, idx, idx)] = Xs[np.ix_(idx, idx, idx)]
        return np.real(np.fft.ifftn(P)) * 8

    fp, hp = pad(f), pad(h)
    exact = np.fft.fftn(fp * hp) / 8
    ex = np.zeros(g.cshape, complex)
    full = np.zeros((n, n, n), complex)
    idx = np.r_[0:n // 2, -n // 2 + 1:0]
    full[np.ix_(idx, idx, idx)] = exact[np.ix_(idx, idx, idx)]
    ex = full[:, :, : n // 2 + 1] * d23
    assert rel(prod, ex) < 1e-12


def test_phi_filter_z_normalisation_quirk(oracle):
    """operations.hpp:23 normalises kz by k_val_at((Nz/2+1)/2)."""
    n = 32
    g = oracle.Grid.cube(n)
    F = np.ones(g.cshape, complex)
    out = g.filter_phi(F)
    kzq = 2 * np.pi * ((n // 2 + 1) // 2) / L
    for kz in (1, 3, 5, 8):
        k = 2 * np.pi * kz / L
        assert out[0, 0, kz].real == pytest.approx(np.exp(-2000 * (k / kzq) ** 15), rel=1e-12, abs=1e-300)
    kx = 2 * np.pi * 5 / L
    assert out[5, 0, 0].real == pytest.approx(np.exp(-2000 * (kx / (2 * np.pi * 16 / L)) ** 15), rel=1e-12)


def test_sv_and_barrier_closed_forms(oracle):
    n = 16
    g = oracle.Grid.cube(n)
    sv = g.sv_table()
    assert sv[0, 0, 0] == 0.0
    assert sv[n // 2, n // 2, n // 2] == pytest.approx(0.5 * (1 - np.exp(-8 * 9)))
    assert sv[n // 2, 0, 0] == pytest.approx(0.5 * (1 - np.exp(-8)))
    b = g.barrier_table()
    x = np.array([oracle.val_at(g.axes[0], i) for i in range(n)])
    exp_b = np.where(x < -L / 6, np.cos(3 * np.pi * x / L) ** 2, 0.0)
    assert np.allclose(b[:, 3, 5], exp_b, atol=1e-15)


# ----------------------------------------------------------------- RNG / set-up
def test_philox_known_answers(oracle):
    """Known-answer vectors of Philox4x32-10 (Random123 distribution)."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for c, k, o in kat:
        assert list(oracle.philox(c, k)) == o


def test_normals_statistics(oracle):
    z = oracle.normals(0, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert np.array_equal(z, oracle.normals(0, 200000))
    assert not np.array_equal(z[:100], oracle.normals(1, 100))


def test_gauss_integral(oracle):
    g = oracle.Grid.cube(32)
    a = 1.0
    nd = g.gauss(-0.1, a)
    assert nd.sum() * (L / 32) ** 3 == pytest.approx(-0.1, rel=1e-10)
    assert g.mask(a).max() == pytest.approx(np.exp(0.0), rel=1e-12) or g.mask(a).max() <= 1.0


# ----------------------------------------------------------------- Poisson
def test_poisson_uniform_density(oracle):
    g = oracle.Grid.cube(16)
    rho = np.full(g.rshape, 1.7)
    phi, it, res = g.poisson_fas(np.zeros(g.rshape), rho, 1e-14)
    assert it >= 0 and np.abs(phi - np.log(1.7)).max() < 1e-13
    phi2, it2, st = g.nlj_solve(np.zeros(g.rshape), rho, max_iter=2000, tol=1e-14)
    assert np.abs(phi2 - np.log(1.7)).max() < 1e-12


def test_poisson_linear_response(oracle):
    """rho = 1 + eps cos(kx): phi = eps cos(kx)/(1 + lambda_h) + O(eps^2),
    lambda_h = (2/h^2)(1 - cos(kh)) the 7-point symbol."""
    n = 16
    g = oracle.Grid.cube(n)
    h = L / n
    x = np.array([oracle.val_at(g.axes[0], i) for i in range(n)])
    eps = 1e-6
    for j in (1, 3):
        k = 2 * np.pi * j / L
        rho = 1 + eps * np.cos(k * x)[:, None, None] * np.ones(g.rshape)
        phi, it, res = g.poisson_fas(np.zeros(g.rshape), rho, 1e-15)
        lam = (2 / h ** 2) * (1 - np.cos(k * h))
        expect = eps * np.cos(k * x)[:, None, None] / (1 + lam)
        assert np.abs(phi - expect).max() < 1e-11


def test_nlj_reference_and_fas_share_the_fixed_point(oracle):
    g = oracle.Grid.cube(16)
    rho = 1.0 + g.gauss(-0.1, 2.0)
    phi, it, res = g.poisson_fas(np.zeros(g.rshape), rho, 1e-13)
    assert res <= 1e-13
    phi2, it2, st = g.nlj_solve(np.zeros(g.rshape), rho, max_iter=5000, tol=1e-13)
    assert st.converged == 1
    assert np.abs(phi2 - phi).max() < 1e-11
    assert g.poisson_residual(phi2, rho)[0] < 1e-11


def test_cascade_reduces_residual(oracle):
    g = oracle.Grid.cube(32)
    rho = 1.0 + g.gauss(-0.1, 1.0)
    r0 = g.poisson_residual(np.zeros(g.rshape), rho)[0]
    phi, it, st = g.mg_cascade(np.zeros(g.rshape), rho)
    assert it > 0
    assert g.poisson_residual(phi, rho)[0] < 0.05 * r0


# ----------------------------------------------------------------- RHS / RK4
def _state(g, Fn=None):
    z = np.zeros(g.cshape, complex)
    return [z.copy(), z.copy(), z.copy(), z.copy() if Fn is None else Fn]


def test_rhs_equilibrium_is_stationary(oracle):
    g = oracle.Grid.cube(16)
    for pen in (False, True):
        r = oracle.Rhs(g, M=1.0, pen_on=pen, sv_on=True)
        r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
        out, st = r.eval(0.0, _state(g))
        assert st == 0
        assert max(np.abs(o).max() for o in out) < 1e-12


@pytest.mark.parametrize("M", [0.0, 1.3])
def test_rk4_amplification_of_a_linear_mode(oracle, M):
    """u = (0, 0, eps sin(kx)), n = 1: du_z/dt = -tau u_z - M d_x u_z exactly,
    so each mode is multiplied by 1 + z + z^2/2 + z^3/6 + z^4/24."""
    n = 16
    g = oracle.Grid.cube(n)
    tau, dt = 0.1, 0.01
    r = oracle.Rhs(g, M=M, tau=tau, pen_on=False, sv_on=False)
    r.set_fields(np.zeros(g.rshape), np.zeros(g.rshape), np.zeros(g.rshape))
    x = np.array([oracle.val_at(g.axes[0], i) for i in range(n)])
    k = 2 * np.pi * 2 / L
    uz = 1e-3 * np.sin(k * x)[:, None, None] * np.ones(g.rshape)
    st = _state(g)
    st[2] = g.r2c(uz)
    F0 = st[2].copy()
    t, s = r.rk4(dt, 0.0, st)
    kx = np.array([oracle.k_val_at(g.axes[0], i) for i in range(n)])[:, None, None]
    z = (-tau - 1j * M * kx) * dt * np.ones(g.cshape)
    amp = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    sel = np.abs(F0) > 1e-6
    assert sel.sum() == 2
    assert np.allclose(st[2][sel], amp[sel] * F0[sel], rtol=1e-13, atol=0)
    assert max(np.abs(st[i]).max() for i in (0, 1, 3)) < 1e-14


def test_cfl_closed_form(oracle):
    n = 16
    g = oracle.Grid.cube(n)
    z = np.zeros(g.cshape, complex)
    assert g.cfl(0.01, z, z, z) == 0.01
    Ux = g.r2c(np.full(g.rshape, 50.0))
    dt_new = g.cfl(0.01, Ux, z, z)
    assert dt_new == pytest.approx(0.4 * (L / n) / 50.0, rel=1e-12)


def test_mg_injection_known_answer(oracle):
    O = oracle
    a = np.arange(8 * 8 * 8, dtype=np.float64).reshape(8, 8, 8)
    assert np.array_equal(O.mg_inject(a, 2), a[::2, ::2, ::2])
    assert np.array_equal(O.mg_inject(a, 8), a[::8, ::8, ::8])


@pytest.mark.parametrize("kind", [1, 2])
def test_mg_prolongation_known_answer(oracle, kind):
    """OP_hto2h: even points copy, odd points interpolate linearly in the
    fine grid's coordinates along z, then y, then x (periodic wrap)."""
    O = oracle
    n = 8
    fine = [O.axis(kind, n, -8.0, 16.0, 3.0 if kind == 2 else 0.0) for _ in range(3)]
    rng = np.random.default_rng(0)
    c = rng.standard_normal((n // 2,) * 3)
    got = O.mg_prolong(c, fine)
    x = [np.array([O.val_at(ax, i) for i in range(-1, n + 2)]) for ax in fine]

    def w(d, i):  # weight of the right neighbour at odd fine index i on axis d
        xs = x[d]
        return (xs[i + 1] - xs[i]) / (xs[i + 2] - xs[i])

    ref = np.zeros((n, n, n))
    ref[::2, ::2, ::2] = c
    for k in range(1, n, 2):
        g = w(2, k)
        ref[::2, ::2, k] = (1 - g) * ref[::2, ::2, k - 1] + g * ref[::2, ::2, (k + 1) % n]
    for j in range(1, n, 2):
        g = w(1, j)
        ref[::2, j, :] = (1 - g) * ref[::2, j - 1, :] + g * ref[::2, (j + 1) % n, :]
    for i in range(1, n, 2):
        g = w(0, i)
        ref[i] = (1 - g) * ref[i - 1] + g * ref[(i + 1) % n]
    assert np.allclose(got, ref, rtol=0, atol=1e-14)
    if kind == 1:  # uniform grid: odd points are plain averages
        assert np.allclose(got[1::2, ::2, ::2], 0.5 * (c + np.roll(c, -1, 0)), atol=1e-14)
