"""ctypes front end of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Arrays follow the reference layout (field_base.hpp:23-24, field_imag.cpp:15-16):
real fields are float64[Nx, Ny, Nz]; complex fields are complex128[Nx, Ny, Nz//2+1]
(interleaved re/im in memory).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EQST, HYST, SIST = 1, 2, 3
POISSON_FAS, POISSON_NLJ, POISSON_CASCADE = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB_PATH


class Axis(C.Structure):
    _fields_ = [("type", C.c_int), ("N", C.c_int), ("l0", C.c_double),
                ("L", C.c_double), ("m", C.c_double)]


class Params(C.Structure):
    _fields_ = [("M", C.c_double), ("tau", C.c_double), ("theta", C.c_double),
                ("mu", C.c_double), ("beta", C.c_double), ("sv_on", C.c_int),
                ("pen_on", C.c_int), ("poisson_mode", C.c_int),
                ("poisson_tol", C.c_double), ("poisson_max", C.c_int)]


class NljState(C.Structure):
    _fields_ = [("norm_max", C.c_double), ("norm_sum", C.c_double),
                ("supremum", C.c_double), ("infinum", C.c_double),
                ("norm_max_old", C.c_double), ("norm_sum_old", C.c_double),
                ("limit_max", C.c_double), ("omega_SOR", C.c_double),
                ("iteration", C.c_int), ("invocation", C.c_int),
                ("max_iterations", C.c_int), ("converged", C.c_int),
                ("ion_theta", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.or_axis_val_at.restype = C.c_double
        L.or_axis_val_at.argtypes = [C.POINTER(Axis), C.c_int]
        L.or_axis_k_val_at.restype = C.c_double
        L.or_axis_k_val_at.argtypes = [C.POINTER(Axis), C.c_int]
        L.or_axis_dS.restype = C.c_double
        L.or_axis_dS.argtypes = [C.POINTER(Axis), C.c_int]
        L.or_grid_create.restype = vp
        L.or_grid_create.argtypes = [C.POINTER(Axis)] * 3
        L.or_grid_destroy.argtypes = [vp]
        for name in ("or_fft_r2c", "or_fft_c2r"):
            getattr(L, name).argtypes = [vp, vp, vp]
        L.or_fft1d.argtypes = [vp, C.c_int, C.c_int]
        for name in ("or_dealias23", "or_filter_phi", "or_sv_table", "or_barrier_table",
                     "or_dealias23_table"):
            getattr(L, name).argtypes = [vp, vp]
        L.or_deriv.argtypes = [vp, C.c_int, vp, vp]
        L.or_nlj_init.argtypes = [C.POINTER(NljState)]
        L.or_nlj_solve.restype = C.c_int
        L.or_nlj_solve.argtypes = [C.POINTER(NljState), vp, vp, vp]
        L.or_mg_cascade_solve.restype = C.c_int
        L.or_mg_cascade_solve.argtypes = [C.POINTER(NljState), vp, vp, vp, C.c_int, vp, vp, vp]
        L.or_poisson_fas.restype = C.c_int
        L.or_poisson_fas.argtypes = [vp, vp, vp, C.c_double, C.c_int, dp]
        L.or_poisson_residual.restype = C.c_double
        L.or_poisson_residual.argtypes = [vp, vp, vp, vp]
        L.or_rhs_create.restype = vp
        L.or_rhs_create.argtypes = [vp, C.POINTER(Params)]
        L.or_rhs_destroy.argtypes = [vp]
        L.or_rhs_set_fields.argtypes = [vp, vp, vp, vp]
        L.or_rhs_get_phi.argtypes = [vp, vp]
        L.or_rhs_eval.restype = C.c_int
        L.or_rhs_eval.argtypes = [vp, C.c_double, vp, vp, vp, vp]
        L.or_rhs_poisson_status.restype = C.c_int
        L.or_rhs_poisson_status.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.or_euler_step.restype = C.c_int
        L.or_euler_step.argtypes = [vp, C.c_double, dp, vp, vp, vp, vp]
        L.or_rk4_step.restype = C.c_int
        L.or_rk4_step.argtypes = [vp, C.c_double, dp, vp, vp, vp, vp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_mg_inject.argtypes = [vp, vp, vp, C.c_int]
        L.or_mg_prolong.argtypes = [vp, vp, vp, C.POINTER(Axis)]
        L.or_rhs_eval_tl.restype = C.c_int
        L.or_rhs_eval_tl.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.c_double, vp]
        L.or_rk4_step_tl.restype = C.c_int
        L.or_rk4_step_tl.argtypes = [vp, C.c_double, C.POINTER(vp), C.POINTER(vp), C.c_double, vp]
        L.or_objective_tl.restype = C.c_double
        L.or_objective_tl.argtypes = [vp, vp, vp, vp, vp, dp]
        L.or_frontend_init.restype = C.c_int
        L.or_frontend_init.argtypes = [vp, C.c_double, vp, C.c_int, C.c_double, C.c_int, vp, vp]
        L.or_cfl.restype = C.c_double
        L.or_cfl.argtypes = [vp, C.c_double, vp, vp, vp]
        L.or_philox4x32_10.argtypes = [vp, vp, vp]
        L.or_normal_fill.argtypes = [C.c_uint64, C.c_size_t, vp]
        L.or_gauss_fill.argtypes = [vp, C.c_double, C.c_double, vp]
        L.or_mask_fill.argtypes = [vp, C.c_double, vp]
        L.or_init_density.argtypes = [vp, C.c_uint64, C.c_double, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def axis(type_, N, l0, L, m=0.0) -> Axis:
    return Axis(type_, N, l0, L, m)


def val_at(ax: Axis, i: int) -> float:
    return lib().or_axis_val_at(C.byref(ax), i)


def k_val_at(ax: Axis, i: int) -> float:
    return lib().or_axis_k_val_at(C.byref(ax), i)


def dS(ax: Axis, i: int) -> float:
    return lib().or_axis_dS(C.byref(ax), i)


class Grid:
    def __init__(self, ax: Axis, ay: Axis, az: Axis):
        self.axes = (ax, ay, az)
        self.n = (ax.N, ay.N, az.N)
        self.nzc = az.N // 2 + 1
        self.h = lib().or_grid_create(C.byref(ax), C.byref(ay), C.byref(az))

    @classmethod
    def cube(cls, n, L=16.0, type_=EQST, m=0.0):
        a = axis(type_, n, -0.5 * L, L, m)
        return cls(a, axis(type_, n, -0.5 * L, L, m), axis(type_, n, -0.5 * L, L, m))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_grid_destroy(self.h)
            self.h = None

    @property
    def rshape(self):
        return self.n

    @property
    def cshape(self):
        return (self.n[0], self.n[1], self.nzc)

    def r2c(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.cshape, np.complex128)
        lib().or_fft_r2c(self.h, _p(x), _p(out))
        return out

    def c2r(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.complex128)
        out = np.empty(self.rshape, np.float64)
        lib().or_fft_c2r(self.h, _p(X), _p(out))
        return out

    def dealias23(self, X):
        X = np.array(X, dtype=np.complex128, order="C")
        lib().or_dealias23(self.h, _p(X))
        return X

    def filter_phi(self, X):
        X = np.array(X, dtype=np.complex128, order="C")
        lib().or_filter_phi(self.h, _p(X))
        return X

    def deriv(self, d: int, X):
        X = np.ascontiguousarray(X, dtype=np.complex128)
        out = np.empty(self.cshape, np.complex128)
        lib().or_deriv(self.h, d, _p(X), _p(out))
        return out

    def sv_table(self):
        out = np.empty(self.cshape, np.float64)
        lib().or_sv_table(self.h, _p(out))
        return out

    def d23_table(self):
        out = np.empty(self.cshape, np.float64)
        lib().or_dealias23_table(self.h, _p(out))
        return out

    def barrier_table(self):
        out = np.empty(self.rshape, np.float64)
        lib().or_barrier_table(self.h, _p(out))
        return out

    def gauss(self, amplitude, a):
        out = np.empty(self.rshape, np.float64)
        lib().or_gauss_fill(self.h, amplitude, a, _p(out))
        return out

    def mask(self, a):
        out = np.empty(self.rshape, np.float64)
        lib().or_mask_fill(self.h, a, _p(out))
        return out

    def init_density(self, seed, amp):
        out = np.empty(self.cshape, np.complex128)
        lib().or_init_density(self.h, seed, amp, _p(out))
        return out

    def poisson_fas(self, phi, rho, tol, max_cycles=200):
        phi = np.array(phi, dtype=np.float64, order="C")
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        res = C.c_double(0)
        it = lib().or_poisson_fas(self.h, _p(phi), _p(rho), tol, max_cycles, C.byref(res))
        return phi, it, res.value

    def poisson_residual(self, phi, rho):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        r = np.empty(self.rshape, np.float64)
        rel = lib().or_poisson_residual(self.h, _p(phi), _p(rho), _p(r))
        return rel, r

    def nlj_solve(self, phi, rho, state: NljState | None = None, max_iter=50, tol=1.45e-6):
        if state is None:
            state = NljState()
            lib().or_nlj_init(C.byref(state))
        state.max_iterations = max_iter
        state.limit_max = tol
        phi = np.array(phi, dtype=np.float64, order="C")
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        it = lib().or_nlj_solve(C.byref(state), self.h, _p(phi), _p(rho))
        return phi, it, state

    def mg_cascade(self, phi, rho, lvl=(-3, 1, 1, 1), steps=(0, -1, -1, -1),
                   tol=(0.01, 0.0005, 0.0001, 0.0001), state: NljState | None = None):
        if state is None:
            state = NljState()
            lib().or_nlj_init(C.byref(state))
        phi = np.array(phi, dtype=np.float64, order="C")
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        lv = np.array(lvl, dtype=np.int32)
        st = np.array(steps, dtype=np.int32)
        tl = np.array(tol, dtype=np.float64)
        it = lib().or_mg_cascade_solve(C.byref(state), self.h, _p(phi), _p(rho), len(lvl),
                                       _p(lv), _p(st), _p(tl))
        return phi, it, state

    def frontend_init(self, theta, nd, cascade=True, tol=1e-12, max_it=5000, phi0=None):
        """(phi, n_i, iterations) of the reference frontend's equilibrium
        start (frontend.cpp:133-198, 291-310)."""
        nd = np.ascontiguousarray(nd, dtype=np.float64)
        phi = np.zeros(self.rshape) if phi0 is None else np.array(phi0, dtype=np.float64, order="C")
        ni = np.empty(self.rshape, np.float64)
        it = lib().or_frontend_init(self.h, theta, _p(nd), int(cascade), tol, max_it, _p(phi), _p(ni))
        return phi, ni, it

    def nlj_ion(self, phi, rho, theta, max_iter=5000, tol=1e-13):
        """The __FRONTEND__ NLJ (solver_poisson_jacobi_nlin.cpp:192-195)."""
        state = NljState()
        lib().or_nlj_init(C.byref(state))
        state.ion_theta = theta
        return self.nlj_solve(phi, rho, state=state, max_iter=max_iter, tol=tol)

    def cfl(self, dt, FUx, FUy, FUz):
        a = [np.ascontiguousarray(f, dtype=np.complex128) for f in (FUx, FUy, FUz)]
        return lib().or_cfl(self.h, dt, _p(a[0]), _p(a[1]), _p(a[2]))


class Rhs:
    """rhs_standard restated (rhs_standard.cpp:7-566)."""

    def __init__(self, grid: Grid, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=True,
                 poisson_mode=POISSON_FAS, poisson_tol=1e-13, poisson_max=200):
        self.grid = grid
        self.params = Params(M, tau, theta, 0.0, 0.0, int(sv_on), int(pen_on), poisson_mode,
                             poisson_tol, poisson_max)
        self.h = lib().or_rhs_create(grid.h, C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_rhs_destroy(self.h)
            self.h = None

    def set_fields(self, phi=None, nd=None, mask=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (phi, nd, mask)]
        lib().or_rhs_set_fields(self.h, *[None if a is None else _p(a) for a in arrs])

    def phi(self):
        out = np.empty(self.grid.rshape, np.float64)
        lib().or_rhs_get_phi(self.h, _p(out))
        return out

    def eval(self, t, state):
        """state: list of 4 complex arrays (FUx, FUy, FUz, Fn); returns new arrays."""
        s = [np.array(f, dtype=np.complex128, order="C") for f in state]
        st = lib().or_rhs_eval(self.h, t, *[_p(a) for a in s])
        return s, st

    def poisson_status(self):
        res = C.c_double(0)
        it = C.c_int(0)
        st = lib().or_rhs_poisson_status(self.h, C.byref(res), C.byref(it))
        return st, res.value, it.value

    def eval_tl(self, state, tangent, dM=0.0, dnd=None):
        """RHS and its tangent along (tangent, dM, dnd): returns (F, T, status)."""
        s = [np.array(f, dtype=np.complex128, order="C") for f in state]
        t = [np.array(f, dtype=np.complex128, order="C") for f in tangent]
        d = None if dnd is None else np.ascontiguousarray(dnd, np.float64)
        st = lib().or_rhs_eval_tl(self.h, (C.c_void_p * 4)(*[_p(a) for a in s]),
                                  (C.c_void_p * 4)(*[_p(a) for a in t]), dM, None if d is None else _p(d))
        return s, t, st

    def rk4_tl(self, dt, y, ty, dM=0.0, dnd=None):
        """One RK4 step of y and of its tangent ty, both in place; returns status."""
        d = None if dnd is None else np.ascontiguousarray(dnd, np.float64)
        return lib().or_rk4_step_tl(self.h, dt, (C.c_void_p * 4)(*[_p(a) for a in y]),
                                    (C.c_void_p * 4)(*[_p(a) for a in ty]), dM, None if d is None else _p(d))

    def objective_tl(self, Fn, Tn, dV, dnd=None):
        """(J, dJ) of the grain drag at spectral ln n Fn along (Tn, dnd)."""
        Fn = np.ascontiguousarray(Fn, np.complex128)
        Tn = None if Tn is None else np.ascontiguousarray(Tn, np.complex128)
        dV = np.ascontiguousarray(dV, np.float64)
        d = None if dnd is None else np.ascontiguousarray(dnd, np.float64)
        dJ = C.c_double(0)
        J = lib().or_objective_tl(self.h, _p(Fn), None if Tn is None else _p(Tn), _p(dV),
                                  None if d is None else _p(d), C.byref(dJ))
        return J, dJ.value

    def euler(self, dt, t, state):
        """euler_method::solve in place; returns (t_new, status)."""
        tt = C.c_double(t)
        st = lib().or_euler_step(self.h, dt, C.byref(tt), *[_p(a) for a in state])
        return tt.value, st

    def rk4(self, dt, t, state):
        """In place on the 4 arrays of `state`; returns (t_new, status)."""
        tt = C.c_double(t)
        st = lib().or_rk4_step(self.h, dt, C.byref(tt), *[_p(a) for a in state])
        return tt.value, st


def set_threads(n: int):
    lib().or_set_threads(n)


def get_threads() -> int:
    return lib().or_get_threads()


def mg_inject(a, f):
    a = np.ascontiguousarray(a, np.float64)
    nin = np.array(a.shape, np.int32)
    out = np.empty(tuple(n // f for n in a.shape), np.float64)
    lib().or_mg_inject(_p(a), _p(nin), _p(out), f)
    return out


def mg_prolong(a, fine_axes):
    a = np.ascontiguousarray(a, np.float64)
    nin = np.array(a.shape, np.int32)
    out = np.empty(tuple(2 * n for n in a.shape), np.float64)
    ax = (Axis * 3)(*fine_axes)
    lib().or_mg_prolong(_p(a), _p(nin), _p(out), ax)
    return out


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.empty(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def normals(seed, n):
    out = np.empty(n, np.float64)
    lib().or_normal_fill(seed, n, _p(out))
    return out


def fft1d(x, sign=-1):
    a = np.array(x, dtype=np.complex128, order="C")
    lib().or_fft1d(_p(a), a.size, sign)
    return a
