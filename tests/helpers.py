"""Matching GPU / oracle set-ups built from the same seeded inputs."""
import numpy as np

L = 16.0


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def axes_for(kind, n, m=None):
    """(kf axes, oracle axes) of a cube with n points per side, L = 16."""
    import kielflow_amd as kf
    from oracle import oracle as O
    if kind == "eqst":
        t, mm = kf.EQST, 0.0
    elif kind == "sist":
        t, mm = kf.SIST, (L * (1 - 1 / 3) / (2 * np.pi) if m is None else m)
    else:
        t, mm = kf.HYST, (3.0 if m is None else m)
    return ([kf.axis(t, n, -0.5 * L, L, mm) for _ in range(3)],
            [O.axis(t, n, -0.5 * L, L, mm) for _ in range(3)])


def grain_radius(kind, n, m=None):
    """sigma = 2 cells of the local spacing at the grain (x = 0)."""
    import kielflow_amd as kf
    import ctypes as C
    axes, _ = axes_for(kind, n, m)
    a = axes[0]
    v0, v1 = C.c_double(), C.c_double()
    kf.load().kf_axis_values(C.byref(a), n // 2, C.byref(v0), None, None)
    kf.load().kf_axis_values(C.byref(a), n // 2 + 1, C.byref(v1), None, None)
    return 2.0 * (v1.value - v0.value)


def initial_state(kind, n, seed=0, amp=1e-3, Q=-0.1, m=None):
    """Host arrays: spectral state [FUx, FUy, FUz, Fn] (u = 0), nd, mask."""
    import kielflow_amd as kf
    from oracle import oracle as O
    axes, oaxes = axes_for(kind, n, m)
    a = grain_radius(kind, n, m)
    ln_n, nd, mask = kf.setup_host(axes, seed=seed, amp=amp, Q=Q, a=a)
    og = O.Grid(*oaxes)
    Fn = og.r2c(ln_n)
    z = np.zeros(og.cshape, complex)
    return [z.copy(), z.copy(), z.copy(), Fn], nd, mask, og


def cell_volumes(og):
    """dV = prod_b S'_b(xi) L_b / N_b on the oracle grid's axes."""
    from oracle import oracle as O
    d = [np.array([O.dS(ax, i) * ax.L / ax.N for i in range(ax.N)]) for ax in og.axes]
    return d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :]
