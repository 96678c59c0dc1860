"""The oracle's tangent-linear model against central differences of the
oracle's own forward RHS / RK4 step / objective (CPU, FP64).  The GPU adjoint
is then checked against this tangent run (tests/test_gpu_adjoint.py)."""
import numpy as np
import pytest

from helpers import cell_volumes, initial_state


def _ip(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag)) if np.iscomplexobj(a) else float(np.sum(a * b))


def _rhs(og, nd, mask, M=1.0, sv=False, pen=True):
    from oracle import oracle as O
    r = O.Rhs(og, M=M, tau=0.1, theta=100.0, sv_on=sv, pen_on=pen, poisson_mode=O.POISSON_FAS,
              poisson_tol=1e-14, poisson_max=400)
    r.set_fields(np.zeros(og.rshape), nd, mask)
    return r


def _state(kind, n, seed=2):
    S, nd, mask, og = initial_state(kind, n, seed=seed, Q=-0.2)
    rng = np.random.default_rng(9)
    for a in range(3):
        S[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
    return S, nd, mask, og


@pytest.mark.parametrize("kind,sv,pen", [("eqst", False, True), ("sist", True, True), ("hyst", False, False)])
def test_rhs_tangent_matches_central_differences(kind, sv, pen):
    S, nd, mask, og = _state(kind, 12)
    rng = np.random.default_rng(3)
    V = [og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-3 for _ in range(4)]
    dnd = 1e-2 * rng.standard_normal(og.rshape)
    dM = 0.3
    F, T, st = _rhs(og, nd, mask, sv=sv, pen=pen).eval_tl(S, V, dM, dnd)
    assert st == 0
    F0, st0 = _rhs(og, nd, mask, sv=sv, pen=pen).eval(0.0, S)
    for a in range(4):
        assert np.abs(F[a] - F0[a]).max() <= 1e-11 * np.abs(F0[a]).max()
    h = 1e-3

    def f(s):
        r = _rhs(og, nd + s * h * dnd, mask, M=1.0 + s * h * dM, sv=sv, pen=pen)
        return r.eval(0.0, [x + s * h * v for x, v in zip(S, V)])[0]

    fp, fm = f(1), f(-1)
    for a in range(4):
        fd = (fp[a] - fm[a]) / (2 * h)
        err = np.abs(fd - T[a]).max() / np.abs(T[a]).max()
        assert err <= 1e-7, (a, err)


def test_rk4_and_objective_tangent_match_central_differences():
    S, nd, mask, og = _state("hyst", 12)
    dV = cell_volumes(og)
    rng = np.random.default_rng(4)
    V = [og.dealias23(og.r2c(rng.standard_normal(og.rshape))) * 1e-3 for _ in range(4)]
    dnd = 1e-2 * rng.standard_normal(og.rshape)
    dM, dt, steps = 0.3, 1e-3, 3

    def run(s, tl):
        r = _rhs(og, nd + s * dnd, mask, M=1.0 + s * dM)
        y = [x + s * v for x, v in zip(S, V)]
        ty = [v.copy() for v in V]
        for _ in range(steps):
            if tl:
                assert r.rk4_tl(dt, y, ty, dM, dnd) == 0
            else:
                r.rk4(dt, 0.0, y)
        J, dJ = r.objective_tl(y[3], ty[3] if tl else None, dV, dnd if tl else None)
        return J, dJ, y, ty

    J0, dJ, y0, ty0 = run(0.0, True)
    h = 1e-3
    Jp, _, yp, _ = run(h, False)
    Jm, _, ym, _ = run(-h, False)
    for a in range(4):
        fd = (yp[a] - ym[a]) / (2 * h)
        assert np.abs(fd - ty0[a]).max() <= 1e-7 * np.abs(ty0[a]).max(), a
    fdJ = (Jp - Jm) / (2 * h)
    assert abs(fdJ - dJ) <= 1e-7 * abs(dJ), (fdJ, dJ)
