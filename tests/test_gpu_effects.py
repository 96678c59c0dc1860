"""The effect_* entry points composed in rhs_standard's order reproduce the
RHS (kf_rhs_solve on the GPU and the oracle's rhs_standard restatement)."""
import numpy as np
import pytest

from helpers import axes_for, initial_state, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,sv", [("eqst", False), ("eqst", True), ("sist", True), ("hyst", False)])
def test_effects_compose_to_the_rhs(ctx, oracle, kind, sv):
    import kielflow_amd as kf
    print("effects", kind, sv, flush=True)
    n, M, tau, theta = 16, 1.2, 0.1, 100.0
    S, nd, mask, og = initial_state(kind, n, seed=5, Q=-0.2)
    rng = np.random.default_rng(12)
    for a in range(3):
        S[a] = og.dealias23(og.r2c(1e-2 * rng.standard_normal(og.rshape)))
    axes, _ = axes_for(kind, n)
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-13, max_it=100)
    rhs = kf.Rhs(g, ps, M=M, tau=tau, theta=theta, sv_on=sv, pen_on=True)
    phi = g.rfield()
    d_mask = ctx.to_device(mask)
    rhs.set_fields(phi, ctx.to_device(nd), d_mask)
    F = [ctx.to_device(x) for x in S]
    rhs.solve(0.0, F)
    ref = [f.get() for f in F]
    # the same terms through the effect interfaces, at the converged potential
    fft = kf.FFT(g, kf.KF_FP64)
    FPhi = g.cfield()
    fft.r2c(phi, FPhi)
    FU = [ctx.to_device(x) for x in S[:3]]
    Fn = ctx.to_device(S[3])
    out = [g.cfield() for _ in range(4)]
    for a in range(3):
        g.effect_force_linear(-tau, FU[a], out[a])
        if sv:
            g.effect_spectral_viscosity(FU[a], out[a])
    if sv:
        g.effect_spectral_viscosity(Fn, out[3])
    g.effect_force_E(FPhi, out[:3])
    g.effect_diffusion_u(theta, Fn, out[:3])
    for f in FU + [Fn]:
        g.dealias23(f)
    for a in range(3):
        g.effect_translation(M, FU[a], out[a])
    g.effect_translation(M, Fn, out[3])
    buf = g.cfield()
    g.effect_continuity(FU, Fn, buf)
    out3 = out[3].get() + buf.get()
    out[3].upload(out3)
    for a in range(3):
        g.effect_advection(FU, FU[a], out[a])
    fix = [-M, 0.0, 0.0, -10.0]
    for a, X in enumerate(FU + [Fn]):
        g.effect_penalization(d_mask, -2.0, fix[a], a == 0, X, out[a])
    got = [o.get() for o in out]
    for a in range(4):
        assert rel_l2(got[a], ref[a]) <= 1e-11, (a, rel_l2(got[a], ref[a]))
    # and the oracle's rhs_standard
    from oracle import oracle as O
    r = O.Rhs(og, M=M, tau=tau, theta=theta, sv_on=sv, pen_on=True, poisson_mode=O.POISSON_FAS,
              poisson_tol=1e-14, poisson_max=400)
    r.set_fields(np.zeros(og.rshape), nd, mask)
    orc, st = r.eval(0.0, S)
    assert st == 0
    for a in range(4):
        assert rel_l2(got[a], orc[a]) <= 1e-10, (a, rel_l2(got[a], orc[a]))
    for o in (fft, rhs, ps, g):
        o.close()


def test_penalization_rejects_positive_damping(ctx):
    import kielflow_amd as kf
    g = kf.Grid(ctx, kf.cube(16))
    X = g.cfield()
    with pytest.raises(kf.KfError):
        g.effect_penalization(g.rfield(), 1.0, 0.0, False, X, g.cfield())
    g.close()
