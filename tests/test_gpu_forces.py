"""Force diagnostics: field_integrate and the backend's force decomposition
(backend.cpp:271-654) on the GPU.

Each term is pinned against the oracle's rhs_standard by a parameter or
state flip that isolates it (the RHS is affine in tau, M, 1/theta and the
spectral viscosity switch, and u -> -u flips exactly the terms linear in u),
so a dropped term, a wrong sign or a swapped component fails one check."""
import numpy as np
import pytest

from helpers import axes_for, cell_volumes, initial_state, rel_l2

pytestmark = pytest.mark.gpu

TERMS = ("DISSIPATION", "SPECTRAL_VISCOSITY", "E_INT", "DIFFUSION", "E_EXT", "ADVECTION")


@pytest.mark.parametrize("kind", ["eqst", "sist", "hyst"])
def test_integrate_closed_forms(ctx, kind):
    """dV = S'(xi) L / N.  EqSt and SiSt: S' is periodic, so the box volume
    is exact and cos^2(2 pi x / L) integrates to L^3 / 2 to spectral
    accuracy.  HySt: S' = m cosh(2 m xi / L) / sinh(m) is not periodic, the
    rule is the trapezoid rule and the volume error falls as 1 / N^2."""
    import kielflow_amd as kf
    print("integrate", kind, flush=True)
    n, L = 32, 16.0
    axes, _ = axes_for(kind, n)
    g = kf.Grid(ctx, axes)
    vol = g.integrate(ctx.to_device(np.ones(g.rshape)))
    if kind == "hyst":
        g2 = kf.Grid(ctx, axes_for(kind, 2 * n)[0])
        vol2 = g2.integrate(ctx.to_device(np.ones(g2.rshape)))
        g2.close()
        ratio = (vol - L ** 3) / (vol2 - L ** 3)
        assert 3.8 < ratio < 4.2, (vol, vol2, ratio)
    else:
        assert abs(vol - L ** 3) <= 1e-12 * L ** 3
        x = np.array([kf.axis_values(axes[0], i)[0] for i in range(n)])
        f = np.broadcast_to(np.cos(2 * np.pi * x / L)[:, None, None] ** 2, g.rshape).copy()
        got = g.integrate(ctx.to_device(f))
        assert abs(got - L ** 3 / 2) <= 1e-9 * L ** 3, got
    # against the host quadrature with the same weights
    w = kf.cell_volumes(axes)
    h = np.random.default_rng(3).standard_normal(g.rshape)
    assert abs(g.integrate(ctx.to_device(h)) - float((h * w).sum())) <= 1e-12 * np.abs(h * w).sum()
    g.close()


@pytest.mark.parametrize("kind", ["eqst", "hyst"])
def test_force_terms_sum_to_the_rhs_and_match_the_oracle(ctx, oracle, kind):
    import kielflow_amd as kf
    print("forces", kind, flush=True)
    n, M, tau, theta = 16, 1.2, 0.1, 100.0
    S, nd, mask, og = initial_state(kind, n, seed=7, Q=-0.2)
    rng = np.random.default_rng(21)
    for a in range(3):
        S[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
    axes, _ = axes_for(kind, n)
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=100)
    rhs = kf.Rhs(g, ps, M=M, tau=tau, theta=theta, sv_on=True, pen_on=False)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    F = [ctx.to_device(x) for x in S]
    rhs.solve(0.0, F)
    R = [f.get() for f in F]
    terms, integ, drag = rhs.force_decomposition([ctx.to_device(x) for x in S], phi)
    assert terms.shape == (6, 3) + tuple(g.rshape)
    # the six terms add up to the velocity RHS (no penalization)
    for a in range(3):
        ref = og.c2r(R[a])
        assert rel_l2(terms[:, a].sum(axis=0), ref) <= 1e-11, (a, rel_l2(terms[:, a].sum(axis=0), ref))
    # integrals are the quadrature of the returned fields
    w = cell_volumes(og)
    for t in range(6):
        for a in range(3):
            ref = float((terms[t, a] * w).sum())
            assert abs(integ[t, a] - ref) <= 1e-12 * max(np.abs(terms[t, a] * w).sum(), 1e-300), (t, a)

    from oracle import oracle as O

    def orhs(state, **kw):
        p = dict(M=M, tau=tau, theta=theta, sv_on=True)
        p.update(kw)
        r = O.Rhs(og, pen_on=False, poisson_mode=O.POISSON_FAS, poisson_tol=1e-14, poisson_max=400, **p)
        r.set_fields(np.zeros(og.rshape), nd, mask)
        out, st = r.eval(0.0, state)
        assert st == 0
        return [og.c2r(x) for x in out[:3]]

    base = orhs(S)
    neg = orhs([-x for x in S[:3]] + [S[3]])
    zero = orhs([np.zeros_like(x) for x in S[:3]] + [S[3]])
    expect = {
        "DISSIPATION": [b - c for b, c in zip(base, orhs(S, tau=0.0))],
        "SPECTRAL_VISCOSITY": [b - c for b, c in zip(base, orhs(S, sv_on=False))],
        "E_EXT": [b - c for b, c in zip(base, orhs(S, M=0.0))],
        "DIFFUSION": [2 * (b - c) for b, c in zip(base, orhs(S, theta=2 * theta))],
        "ADVECTION": [0.5 * (b + c) - z for b, c, z in zip(base, neg, zero)],
    }
    diff_t = expect["DIFFUSION"]
    expect["E_INT"] = [z - d for z, d in zip(zero, diff_t)]
    for t, name in enumerate(TERMS):
        for a in range(3):
            ref = expect[name][a]
            scale = max(np.linalg.norm(ref), 1e-6 * np.linalg.norm(base[a]))
            err = np.linalg.norm(terms[t, a] - ref) / scale
            assert err <= 1e-8, (name, a, err)
    # dissipation is -tau u exactly
    for a in range(3):
        assert rel_l2(terms[0, a], -tau * og.c2r(S[a])) <= 1e-14
    # the drag on the grain against the oracle's adjoint objective
    r = O.Rhs(og, M=M, tau=tau, theta=theta, sv_on=True, pen_on=False, poisson_mode=O.POISSON_FAS,
              poisson_tol=1e-14, poisson_max=400)
    r.set_fields(np.zeros(og.rshape), nd, mask)
    J, _ = r.objective_tl(S[3], None, w)
    assert abs(drag - J) <= 1e-9 * abs(J), (drag, J)
    # only the integrals
    none, integ2, drag2 = rhs.force_decomposition([ctx.to_device(x) for x in S], phi, fields=False)
    assert none is None
    np.testing.assert_allclose(integ2, integ, rtol=1e-12, atol=1e-14 * np.abs(integ).max())
    assert drag2 == pytest.approx(drag, rel=1e-12)
    for o in (rhs, ps, g):
        o.close()


def test_force_decomposition_needs_fp64(ctx):
    import kielflow_amd as kf
    g = kf.Grid(ctx, kf.cube(16))
    ps = kf.Poisson(g, kf.KF_MIXED, tol=1e-8, max_it=50)
    rhs = kf.Rhs(g, ps, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=False, precision=kf.KF_MIXED)
    F = [g.cfield() for _ in range(4)]
    with pytest.raises(kf.KfError):
        rhs.force_decomposition(F, g.rfield())
    for o in (rhs, ps, g):
        o.close()
