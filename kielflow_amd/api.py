"""ctypes binding of libkielflow_amd.so (the C ABI in include/kielflow_amd.h).

Host data is numpy; device data lives in DeviceArray buffers allocated by the
library.  Nothing here computes: every call goes to the HIP library, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkielflow_amd.so")

KF_OK, KF_ERR_ARG, KF_ERR_HIP, KF_ERR_NOT_CONVERGED, KF_ERR_NAN = 0, 1, 2, 3, 4
KF_FP64, KF_MIXED, KF_FP32 = 0, 1, 2
# how an adjoint run got its stage intermediates
KF_ADJ_RECOMPUTE, KF_ADJ_SEGMENTS, KF_ADJ_TRAJECTORY = 0, 1, 2
EQST, HYST, SIST = 1, 2, 3
POISSON_NEWTON_MG, POISSON_NLJ, POISSON_CASCADE = 0, 1, 2


class KfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kielflow_amd error {code}: {msg}")
        self.code = code


class Axis(C.Structure):
    _fields_ = [("type", C.c_int), ("n", C.c_int), ("l0", C.c_double), ("L", C.c_double),
                ("m", C.c_double)]


class Parameters(C.Structure):
    _fields_ = [("M", C.c_double), ("tau", C.c_double), ("theta", C.c_double),
                ("mu", C.c_double), ("beta", C.c_double)]


class Switches(C.Structure):
    _fields_ = [("sv_on", C.c_int), ("pen_on", C.c_int)]


# name -> (restype, argtypes); every int-returning entry point is checked
_vp, _dp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
_AX3 = C.POINTER(Axis)
SIGNATURES = {
    "kf_last_error": (C.c_char_p, []),
    "kf_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "kf_ctx_destroy": (C.c_int, [_vp]),
    "kf_ctx_synchronize": (C.c_int, [_vp]),
    "kf_ctx_stream": (_vp, [_vp]),
    "kf_malloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "kf_free": (C.c_int, [_vp, _vp]),
    "kf_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "kf_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "kf_memcpy_d2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "kf_memset": (C.c_int, [_vp, _vp, C.c_int, C.c_size_t]),
    "kf_grid_create": (C.c_int, [_vp, _AX3, C.POINTER(_vp)]),
    "kf_grid_destroy": (C.c_int, [_vp]),
    "kf_grid_dims": (C.c_int, [_vp, _ip, _ip, _ip, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "kf_axis_values": (C.c_int, [_AX3, C.c_int, _dp, _dp, _dp]),
    "kf_fft_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "kf_fft_destroy": (C.c_int, [_vp]),
    "kf_fft_r2c": (C.c_int, [_vp, _vp, _vp]),
    "kf_fft_c2r": (C.c_int, [_vp, _vp, _vp]),
    "kf_partial_derivative": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "kf_dealias23": (C.c_int, [_vp, _vp]),
    "kf_effect_advection": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "kf_effect_continuity": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "kf_effect_force_E": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "kf_effect_diffusion_u": (C.c_int, [_vp, C.c_double, _vp, _vp, _vp, _vp]),
    "kf_effect_spectral_viscosity": (C.c_int, [_vp, _vp, _vp]),
    "kf_effect_translation": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "kf_effect_force_linear": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "kf_effect_penalization": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_int, _vp, _vp]),
    "kf_filter_phi": (C.c_int, [_vp, _vp]),
    "kf_poisson_create": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_vp)]),
    "kf_poisson_destroy": (C.c_int, [_vp]),
    "kf_poisson_set_tolerance": (C.c_int, [_vp, C.c_double]),
    "kf_poisson_set_max_iterations": (C.c_int, [_vp, C.c_int]),
    "kf_poisson_set_smoothing": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "kf_poisson_set_cascade": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp]),
    "kf_poisson_set_ion_boltzmann": (C.c_int, [_vp, C.c_double]),
    "kf_frontend_equilibrium": (C.c_int, [_vp, C.c_double, _vp, C.c_int, C.c_double, C.c_int, _vp, _vp, _ip]),
    "kf_poisson_solve": (C.c_int, [_vp, _vp, _vp, _ip, _dp]),
    "kf_poisson_converged": (C.c_int, [_vp]),
    "kf_poisson_residual": (C.c_int, [_vp, _vp, _vp, _dp]),
    "kf_rhs_create": (C.c_int, [_vp, C.POINTER(Parameters), C.POINTER(Switches), _vp, C.c_int,
                                C.POINTER(_vp)]),
    "kf_rhs_destroy": (C.c_int, [_vp]),
    "kf_rhs_set_fields": (C.c_int, [_vp, _vp, _vp, _vp]),
    "kf_rhs_set_parameters": (C.c_int, [_vp, C.POINTER(Parameters)]),
    "kf_rhs_solve": (C.c_int, [_vp, C.c_double, _vp, _vp, _vp, _vp]),
    "kf_rhs_poisson_failures": (C.c_int, [_vp, _ip]),
    "kf_rhs_force_general": (C.c_int, [_vp, C.c_int]),
    "kf_rk4_create": (C.c_int, [_vp, C.c_double, C.POINTER(_vp)]),
    "kf_rk4_destroy": (C.c_int, [_vp]),
    "kf_rk4_solve": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "kf_rk4_steps": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "kf_rk4_set_dt": (C.c_int, [_vp, C.c_double]),
    "kf_rk4_time": (C.c_int, [_vp, _dp]),
    "kf_cfl_all": (C.c_int, [_vp, C.c_double, _vp, _vp, _vp, _dp]),
    "kf_field_integrate": (C.c_int, [_vp, _vp, _dp]),
    "kf_poisson_newton_counts": (C.c_int, [_vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "kf_axis_index_at": (C.c_int, [C.POINTER(Axis), C.c_double, _ip]),
    "kf_state_to_real": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "kf_state_plane_z": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "kf_state_from_real": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "kf_force_decomposition": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _dp, _dp]),
    "kf_euler_create": (C.c_int, [_vp, C.c_double, C.POINTER(_vp)]),
    "kf_euler_destroy": (C.c_int, [_vp]),
    "kf_euler_steps": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "kf_euler_set_dt": (C.c_int, [_vp, C.c_double]),
    "kf_cfl_all_noalloc": (C.c_int, [_vp, C.c_double, _vp, _vp, _vp, _dp]),
    "kf_rk4_stage_input": (C.c_int, [_vp, C.c_size_t, C.c_double, _vp, _vp, _vp]),
    "kf_rk4_combine": (C.c_int, [_vp, C.c_size_t, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "kf_adjoint_create": (C.c_int, [_vp, C.c_double, C.POINTER(_vp)]),
    "kf_adjoint_destroy": (C.c_int, [_vp]),
    "kf_adjoint_set_tolerance": (C.c_int, [_vp, C.c_double]),
    "kf_adjoint_reset": (C.c_int, [_vp]),
    "kf_rhs_adjoint": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "kf_rk4_step_adjoint": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "kf_objective": (C.c_int, [_vp, C.POINTER(_vp), _vp, _dp]),
    "kf_objective_adjoint": (C.c_int, [_vp, C.POINTER(_vp), _vp, C.POINTER(_vp)]),
    "kf_adjoint_accumulators": (C.c_int, [_vp, _dp, _vp]),
    "kf_adjoint_run": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_int, _vp, _vp, _dp, _dp, _dp, _vp]),
    "kf_adjoint_spectral_to_real": (C.c_int, [_vp, _vp, _vp]),
    "kf_batch_create": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(_vp)]),
    "kf_batch_destroy": (C.c_int, [_vp]),
    "kf_batch_box": (C.c_int, [_vp, C.c_int] + [C.POINTER(_vp)] * 5),
    "kf_batch_set_fields": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "kf_batch_capture": (C.c_int, [_vp, C.c_int]),
    "kf_batch_steps": (C.c_int, [_vp, C.c_int]),
    "kf_batch_synchronize": (C.c_int, [_vp]),
    "kf_batch_poisson_failures": (C.c_int, [_vp, _ip]),
    "kf_adjoint_set_capture_budget": (C.c_int, [_vp, C.c_size_t]),
    "kf_adjoint_set_keep_potentials": (C.c_int, [_vp, C.c_int]),
    "kf_adjoint_step_capture_bytes": (C.c_int, [_vp, C.POINTER(C.c_size_t)]),
    "kf_adjoint_run_info": (C.c_int, [_vp, _ip, C.POINTER(C.c_long)]),
    "kf_profile_enable": (C.c_int, [C.c_int]),
    "kf_profile_read": (C.c_int, [_ip, _dp]),
    "kf_setup_host": (C.c_int, [_AX3, C.c_ulonglong, C.c_double, C.c_double, C.c_double, _vp, _vp,
                                _vp]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (built by kielflow_amd.build); raise if missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m kielflow_amd.build` "
                              "(the HIP library is the only implementation)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int):
    if code != KF_OK:
        raise KfError(code, load().kf_last_error().decode())


def axis(type_, n, l0, L, m=0.0) -> Axis:
    return Axis(type_, n, l0, L, m)


def axis_values(ax: Axis, i: int):
    """(x, k, S') of point i of an axis (axis::val_at, k_val_at, dS)."""
    v, k, d = C.c_double(), C.c_double(), C.c_double()
    check(load().kf_axis_values(C.byref(ax), i, C.byref(v), C.byref(k), C.byref(d)))
    return v.value, k.value, d.value


def axis_index_at(ax: Axis, val: float) -> int:
    """axis::index_at: the grid index of coordinate val."""
    i = C.c_int(0)
    check(load().kf_axis_index_at(C.byref(ax), float(val), C.byref(i)))
    return i.value


def cell_volumes(axes):
    """dV = prod_b S'_b(xi) L_b / N_b, the quadrature weights of a box."""
    w = [np.array([axis_values(a, i)[2] * a.L / a.n for i in range(a.n)]) for a in axes]
    return w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]


def cube(n, L=16.0, type_=EQST, m=0.0):
    return [axis(type_, n, -0.5 * L, L, m) for _ in range(3)]


def _ax3(axes):
    arr = (Axis * 3)(*axes)
    return arr


def _need(arr, shape, dtype, what):
    """Raise unless `arr` is a device field of exactly this shape and dtype:
    the library takes raw pointers and trusts their extents."""
    if not hasattr(arr, "ptr") or not hasattr(arr, "nbytes"):
        raise TypeError(f"{what}: expected a device field, got {type(arr).__name__}")
    if arr.ptr is None:
        raise ValueError(f"{what}: the device field has been freed")
    dtype = np.dtype(dtype)
    if tuple(arr.shape) != tuple(shape) or np.dtype(arr.dtype) != dtype:
        raise ValueError(f"{what}: expected {dtype.name}{tuple(shape)}, got "
                         f"{np.dtype(arr.dtype).name}{tuple(arr.shape)}")
    return arr.ptr


def _need4(F, grid, what):
    if len(F) != 4:
        raise ValueError(f"{what}: expected the 4 spectral fields (FUx, FUy, FUz, Fn), got {len(F)}")
    return [_need(f, grid.cshape, np.complex128, f"{what}[{i}]") for i, f in enumerate(F)]


class Context:
    def __init__(self, device=0):
        self.lib = load()
        h = _vp()
        check(self.lib.kf_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.kf_ctx_destroy(self.h)
            self.h = None

    def sync(self):
        check(self.lib.kf_ctx_synchronize(self.h))

    @property
    def stream(self):
        return self.lib.kf_ctx_stream(self.h)

    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def zeros(self, shape, dtype) -> "DeviceArray":
        d = DeviceArray(self, shape, dtype)
        check(self.lib.kf_memset(self.h, d.ptr, 0, d.nbytes))
        return d


class DeviceArray:
    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(ctx.lib.kf_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(self.ctx.lib.kf_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(_vp), self.nbytes))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(self.ctx.lib.kf_memcpy_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        check(self.ctx.lib.kf_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.kf_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.h is not None:
                self.free()
        except Exception:
            pass


class Grid:
    def __init__(self, ctx: Context, axes):
        self.ctx = ctx
        self.axes = list(axes)
        h = _vp()
        check(ctx.lib.kf_grid_create(ctx.h, _ax3(self.axes), C.byref(h)))
        self.h = h
        self.n = tuple(a.n for a in self.axes)
        self.nzc = self.n[2] // 2 + 1
        self.rshape = self.n
        self.cshape = (self.n[0], self.n[1], self.nzc)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.kf_grid_destroy(self.h)
            self.h = None

    def cfield(self, dtype=np.complex128):
        return self.ctx.zeros(self.cshape, dtype)

    def rfield(self, dtype=np.float64):
        return self.ctx.zeros(self.rshape, dtype)

    def _c(self, X, what):
        return _need(X, self.cshape, np.complex128, what)

    def _r(self, X, what):
        return _need(X, self.rshape, np.float64, what)

    def deriv(self, d: int, X: DeviceArray, out: DeviceArray):
        check(self.ctx.lib.kf_partial_derivative(self.h, d, self._c(X, "input"), self._c(out, "output")))

    def dealias23(self, X: DeviceArray):
        check(self.ctx.lib.kf_dealias23(self.h, self._c(X, "field")))

    def filter_phi(self, X: DeviceArray):
        check(self.ctx.lib.kf_filter_phi(self.h, self._c(X, "field")))

    # effect_* interfaces (terms of d/dt added to the output buffers)
    def effect_advection(self, FU, FXX, out):
        check(self.ctx.lib.kf_effect_advection(self.h, *[self._c(f, "FU") for f in FU[:3]], self._c(FXX, "FXX"),
                                               self._c(out, "out")))

    def effect_continuity(self, FU, Fni, out):
        check(self.ctx.lib.kf_effect_continuity(self.h, *[self._c(f, "FU") for f in FU[:3]], self._c(Fni, "Fni"),
                                                self._c(out, "out")))

    def effect_force_E(self, FPhi, out):
        check(self.ctx.lib.kf_effect_force_E(self.h, self._c(FPhi, "FPhi"), *[self._c(o, "out") for o in out[:3]]))

    def effect_diffusion_u(self, theta, Fni, out):
        check(self.ctx.lib.kf_effect_diffusion_u(self.h, theta, self._c(Fni, "Fni"),
                                                 *[self._c(o, "out") for o in out[:3]]))

    def effect_spectral_viscosity(self, X, out):
        check(self.ctx.lib.kf_effect_spectral_viscosity(self.h, self._c(X, "input"), self._c(out, "out")))

    def effect_translation(self, v, X, out):
        check(self.ctx.lib.kf_effect_translation(self.h, v, self._c(X, "input"), self._c(out, "out")))

    def effect_force_linear(self, k, X, out):
        check(self.ctx.lib.kf_effect_force_linear(self.h, k, self._c(X, "input"), self._c(out, "out")))

    def effect_penalization(self, mask, eta, fix, fix_scaled_by_mask, X, out):
        check(self.ctx.lib.kf_effect_penalization(self.h, self._r(mask, "mask"), eta, fix, int(fix_scaled_by_mask),
                                                  self._c(X, "input"), self._c(out, "out")))

    def state_to_real(self, F, phi) -> np.ndarray:
        """[Ux, Uy, Uz, ni (linear), Ph] as host double[5][Nx][Ny][Nz]: the
        fields kielflow stores (kielflow.cpp:290-327, IO.cpp:446-481)."""
        ptrs = _need4(F, self, "state")
        out = self.ctx.empty((5,) + tuple(self.rshape), np.float64)
        try:
            check(self.ctx.lib.kf_state_to_real(self.h, *ptrs, self._r(phi, "phi"), out.ptr))
            return out.get()
        finally:
            out.free()

    def state_plane_z(self, F, phi, k: int) -> np.ndarray:
        """The same five fields on z plane k: host double[5][Nx][Ny]."""
        ptrs = _need4(F, self, "state")
        out = self.ctx.empty((5, self.n[0], self.n[1]), np.float64)
        try:
            check(self.ctx.lib.kf_state_plane_z(self.h, *ptrs, self._r(phi, "phi"), int(k), out.ptr))
            return out.get()
        finally:
            out.free()

    def state_from_real(self, fields: np.ndarray, F, phi=None):
        """The inverse: F = FFT(U), F ln n = FFT(log ni), phi = Ph."""
        fields = np.ascontiguousarray(fields, np.float64)
        if fields.shape != (5,) + tuple(self.rshape):
            raise ValueError(f"fields: expected float64{(5,) + tuple(self.rshape)}, got {fields.shape}")
        ptrs = _need4(F, self, "state")
        d = self.ctx.to_device(fields)
        try:
            check(self.ctx.lib.kf_state_from_real(self.h, d.ptr, *ptrs,
                                                  None if phi is None else self._r(phi, "phi")))
        finally:
            d.free()

    def integrate(self, f: DeviceArray):
        """field_integrate: the volume integral of a real field."""
        out = C.c_double(0)
        check(self.ctx.lib.kf_field_integrate(self.h, self._r(f, "field"), C.byref(out)))
        return out.value

    def frontend_equilibrium(self, theta, nd, phi, ni, mode=POISSON_CASCADE, tol=1e-10, max_iterations=100):
        """The equilibrium start of the reference frontend (frontend.cpp:133-198,
        291-310): phi (in/out, the initial guess) and ni = exp(-theta phi),
        both filtered; returns the solver's iteration count."""
        it = C.c_int(0)
        check(self.ctx.lib.kf_frontend_equilibrium(self.h, theta, self._r(nd, "n_d"), mode, tol, max_iterations,
                                                   self._r(phi, "phi"), self._r(ni, "ni"), C.byref(it)))
        return it.value

    def cfl(self, dt, FUx, FUy, FUz):
        out = C.c_double(0)
        check(self.ctx.lib.kf_cfl_all(self.h, dt, self._c(FUx, "FUx"), self._c(FUy, "FUy"), self._c(FUz, "FUz"),
                                      C.byref(out)))
        return out.value


class FFT:
    def __init__(self, grid: Grid, precision=KF_FP64):
        self.grid = grid
        self.precision = precision
        h = _vp()
        check(grid.ctx.lib.kf_fft_create(grid.h, precision, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.grid.ctx.lib.kf_fft_destroy(self.h)
            self.h = None

    def _dtypes(self):
        if self.precision == KF_FP64:
            return np.float64, np.complex128
        return np.float32, np.complex64

    def r2c(self, x: DeviceArray, X: DeviceArray):
        rt, ct = self._dtypes()
        check(self.grid.ctx.lib.kf_fft_r2c(self.h, _need(x, self.grid.rshape, rt, "r2c input"),
                                           _need(X, self.grid.cshape, ct, "r2c output")))

    def c2r(self, X: DeviceArray, x: DeviceArray):
        rt, ct = self._dtypes()
        check(self.grid.ctx.lib.kf_fft_c2r(self.h, _need(X, self.grid.cshape, ct, "c2r input"),
                                           _need(x, self.grid.rshape, rt, "c2r output")))


class Poisson:
    def __init__(self, grid: Grid, precision=KF_FP64, mode=POISSON_NEWTON_MG, tol=1e-10, max_it=100):
        self.grid = grid
        h = _vp()
        check(grid.ctx.lib.kf_poisson_create(grid.h, precision, mode, C.byref(h)))
        self.h = h
        self.set_tolerance(tol)
        self.set_max_iterations(max_it)

    def close(self):
        if getattr(self, "h", None):
            self.grid.ctx.lib.kf_poisson_destroy(self.h)
            self.h = None

    def set_tolerance(self, t):
        check(self.grid.ctx.lib.kf_poisson_set_tolerance(self.h, t))

    def set_max_iterations(self, m):
        check(self.grid.ctx.lib.kf_poisson_set_max_iterations(self.h, m))

    def newton_counts(self):
        """(solves, Newton corrections) finished under the host loop so far."""
        a, b = C.c_long(0), C.c_long(0)
        check(self.grid.ctx.lib.kf_poisson_newton_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_smoothing(self, pre, post, coarse=8):
        check(self.grid.ctx.lib.kf_poisson_set_smoothing(self.h, pre, post, coarse))

    def set_ion_boltzmann(self, theta):
        """Boltzmann ions exp(-theta phi) in the equation (the frontend's solver)."""
        check(self.grid.ctx.lib.kf_poisson_set_ion_boltzmann(self.h, theta))

    def set_cascade(self, level_change, steps, tol):
        """Level program of the reference cascade (KF_POISSON_CASCADE)."""
        n = len(level_change)
        if not len(steps) == len(tol) == n:
            raise ValueError("level_change, steps and tol need the same length")
        check(self.grid.ctx.lib.kf_poisson_set_cascade(self.h, n, (C.c_int * n)(*level_change),
                                                       (C.c_int * n)(*steps), (C.c_double * n)(*tol)))

    def solve(self, phi: DeviceArray, rho: DeviceArray, raise_on_error=True):
        it = C.c_int(0)
        res = C.c_double(0)
        rs = self.grid.rshape
        code = self.grid.ctx.lib.kf_poisson_solve(self.h, _need(phi, rs, np.float64, "phi"),
                                                  _need(rho, rs, np.float64, "rho"), C.byref(it), C.byref(res))
        if raise_on_error:
            check(code)
        return it.value, res.value, code

    def residual(self, phi: DeviceArray, rho: DeviceArray):
        res = C.c_double(0)
        rs = self.grid.rshape
        check(self.grid.ctx.lib.kf_poisson_residual(self.h, _need(phi, rs, np.float64, "phi"),
                                                    _need(rho, rs, np.float64, "rho"), C.byref(res)))
        return res.value


class Rhs:
    def __init__(self, grid: Grid, solver: Poisson, M=1.0, tau=0.1, theta=100.0, sv_on=False,
                 pen_on=True, precision=KF_FP64):
        self.grid = grid
        self.solver = solver
        self.params = Parameters(M, tau, theta, 0.0, 0.0)
        self.switches = Switches(int(sv_on), int(pen_on))
        h = _vp()
        check(grid.ctx.lib.kf_rhs_create(grid.h, C.byref(self.params), C.byref(self.switches),
                                         solver.h, precision, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.grid.ctx.lib.kf_rhs_destroy(self.h)
            self.h = None

    def set_fields(self, phi=None, nd=None, mask=None):
        """n_d and the mask are copied into the library; phi is kept by
        pointer (it is the warm start and the result of every later Poisson
        solve), so this object keeps that field alive while it may be used."""
        rs = self.grid.rshape
        ptrs = [None if a is None else _need(a, rs, np.float64, w)
                for a, w in ((phi, "phi"), (nd, "n_d"), (mask, "mask"))]
        check(self.grid.ctx.lib.kf_rhs_set_fields(self.h, *ptrs))
        if phi is not None:
            self._phi = phi

    def set_parameters(self, M=None, tau=None, theta=None):
        if M is not None:
            self.params.M = M
        if tau is not None:
            self.params.tau = tau
        if theta is not None:
            self.params.theta = theta
        check(self.grid.ctx.lib.kf_rhs_set_parameters(self.h, C.byref(self.params)))

    def solve(self, t, F):
        check(self.grid.ctx.lib.kf_rhs_solve(self.h, t, *_need4(F, self.grid, "rhs state")))

    FORCE_TERMS = ("DISSIPATION", "SPECTRAL_VISCOSITY", "E_INT", "DIFFUSION", "E_EXT", "ADVECTION")

    def force_decomposition(self, F, phi, fields=True):
        """(terms [6, 3, Nx, Ny, Nz] or None, integrals [6, 3], x-drag on the
        grain) of the backend's force decomposition (backend.cpp:271-654)."""
        g = self.grid
        ptrs = _need4(F, g, "state")
        nt = len(self.FORCE_TERMS)
        terms = g.ctx.empty((nt, 3) + tuple(g.rshape), np.float64) if fields else None
        integ = np.zeros(3 * nt, np.float64)
        drag = C.c_double(0)
        check(g.ctx.lib.kf_force_decomposition(self.h, *ptrs, _need(phi, g.rshape, np.float64, "phi"),
                                               terms.ptr if fields else None,
                                               integ.ctypes.data_as(_dp), C.byref(drag)))
        host = None
        if fields:
            host = terms.get()
            terms.free()
        return host, integ.reshape(nt, 3), drag.value

    def force_general(self, on=True):
        check(self.grid.ctx.lib.kf_rhs_force_general(self.h, int(on)))

    def poisson_failures(self):
        c = C.c_int(0)
        check(self.grid.ctx.lib.kf_rhs_poisson_failures(self.h, C.byref(c)))
        return c.value


class Rk4:
    def __init__(self, rhs: Rhs, dt):
        self.rhs = rhs
        h = _vp()
        check(rhs.grid.ctx.lib.kf_rk4_create(rhs.h, dt, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.rhs.grid.ctx.lib.kf_rk4_destroy(self.h)
            self.h = None

    def step(self, F, n=1):
        check(self.rhs.grid.ctx.lib.kf_rk4_steps(self.h, n, *_need4(F, self.rhs.grid, "rk4 state")))

    def set_dt(self, dt):
        check(self.rhs.grid.ctx.lib.kf_rk4_set_dt(self.h, dt))

    @property
    def t(self):
        v = C.c_double(0)
        check(self.rhs.grid.ctx.lib.kf_rk4_time(self.h, C.byref(v)))
        return v.value


class Euler:
    """euler_method (euler_method.cpp:13-91): y = y + dt k per step."""

    def __init__(self, rhs: Rhs, dt):
        self.rhs = rhs
        h = _vp()
        check(rhs.grid.ctx.lib.kf_euler_create(rhs.h, dt, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.rhs.grid.ctx.lib.kf_euler_destroy(self.h)
            self.h = None

    def step(self, F, n=1):
        check(self.rhs.grid.ctx.lib.kf_euler_steps(self.h, n, *_need4(F, self.rhs.grid, "euler state")))

    def set_dt(self, dt):
        check(self.rhs.grid.ctx.lib.kf_euler_set_dt(self.h, dt))


class Adjoint:
    """Discrete adjoint of the FP64 RHS / RK4 step / grain drag objective."""

    def __init__(self, rhs: Rhs, dt: float, tol: float = 1e-12):
        self.rhs = rhs
        h = _vp()
        check(rhs.grid.ctx.lib.kf_adjoint_create(rhs.h, dt, C.byref(h)))
        self.h = h
        check(rhs.grid.ctx.lib.kf_adjoint_set_tolerance(h, tol))

    def close(self):
        if getattr(self, "h", None):
            self.rhs.grid.ctx.lib.kf_adjoint_destroy(self.h)
            self.h = None

    @property
    def lib(self):
        return self.rhs.grid.ctx.lib

    def reset(self):
        check(self.lib.kf_adjoint_reset(self.h))

    def _p4(self, F, what):
        return (_vp * 4)(*_need4(F, self.rhs.grid, what))

    def _r(self, X, what):
        return _need(X, self.rhs.grid.rshape, np.float64, what)

    def rhs_adjoint(self, S, ob, Fb):
        check(self.lib.kf_rhs_adjoint(self.h, self._p4(S, "S"), self._p4(ob, "ob"), self._p4(Fb, "Fb")))

    def step_adjoint(self, y0, yb):
        check(self.lib.kf_rk4_step_adjoint(self.h, self._p4(y0, "y0"), self._p4(yb, "yb")))

    def objective(self, F, dV):
        J = C.c_double(0)
        check(self.lib.kf_objective(self.h, self._p4(F, "F"), self._r(dV, "dV"), C.byref(J)))
        return J.value

    def objective_adjoint(self, F, dV, Fb):
        check(self.lib.kf_objective_adjoint(self.h, self._p4(F, "F"), self._r(dV, "dV"), self._p4(Fb, "Fb")))

    def accumulators(self, nd_bar=None):
        m = C.c_double(0)
        check(self.lib.kf_adjoint_accumulators(self.h, C.byref(m),
                                               None if nd_bar is None else self._r(nd_bar, "nd_bar")))
        return m.value

    def run(self, y, nsteps, k, dV, G, dFn0):
        J, dM, dQ = C.c_double(0), C.c_double(0), C.c_double(0)
        cs = self.rhs.grid.cshape
        check(self.lib.kf_adjoint_run(self.h, self._p4(y, "y"), nsteps, k, self._r(dV, "dV"), self._r(G, "G"),
                                      C.byref(J), C.byref(dM), C.byref(dQ),
                                      _need(dFn0, cs, np.complex128, "dFn0")))
        return J.value, dM.value, dQ.value

    def set_capture_budget(self, nbytes):
        check(self.lib.kf_adjoint_set_capture_budget(self.h, int(nbytes)))

    def set_keep_potentials(self, keep=True):
        check(self.lib.kf_adjoint_set_keep_potentials(self.h, int(bool(keep))))

    def step_capture_bytes(self):
        """Bytes of the kept intermediates of one RK4 step."""
        b = C.c_size_t(0)
        check(self.lib.kf_adjoint_step_capture_bytes(self.h, C.byref(b)))
        return b.value

    def run_info(self):
        """(KF_ADJ_* mode, transposed Poisson iterations) of the last run."""
        k, it = C.c_int(0), C.c_long(0)
        check(self.lib.kf_adjoint_run_info(self.h, C.byref(k), C.byref(it)))
        return k.value, it.value

    def spectral_to_real(self, dF, out):
        """out = R2C^T dF: real-space cotangent of the field whose spectrum is F."""
        check(self.lib.kf_adjoint_spectral_to_real(
            self.h, _need(dF, self.rhs.grid.cshape, np.complex128, "dF"), self._r(out, "out")))


class _View:
    """A device field owned by the library (no free on close)."""

    def __init__(self, ctx, ptr, shape, dtype):
        self.ctx, self.ptr, self.shape, self.dtype = ctx, ptr, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize

    upload = DeviceArray.upload
    get = DeviceArray.get
    copy_from = DeviceArray.copy_from


class Batch:
    """Independent boxes advanced together (kf_batch_*)."""

    def __init__(self, ctx: Context, boxes, sv_on=False, pen_on=True, precision=KF_MIXED, dt=1e-3,
                 poisson_tol=1e-10, nstreams=4):
        """boxes: list of (axes[3], Parameters)."""
        self.ctx = ctx
        n = len(boxes)
        ax = (Axis * (3 * n))(*[a for axes, _ in boxes for a in axes])
        pa = (Parameters * n)(*[p for _, p in boxes])
        sw = Switches(int(sv_on), int(pen_on))
        h = _vp()
        check(ctx.lib.kf_batch_create(ctx.h, n, C.cast(ax, _vp), C.cast(pa, _vp), C.byref(sw), precision, dt,
                                      poisson_tol, nstreams, C.byref(h)))
        self.h = h
        self.shapes = [tuple(a.n for a in axes) for axes, _ in boxes]

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.kf_batch_destroy(self.h)
            self.h = None

    def box(self, b):
        """(state views FUx, FUy, FUz, Fn; potential view) of box b."""
        p = [_vp() for _ in range(5)]
        check(self.ctx.lib.kf_batch_box(self.h, b, *[C.byref(x) for x in p]))
        nx, ny, nz = self.shapes[b]
        cs, rs = (nx, ny, nz // 2 + 1), (nx, ny, nz)
        return [_View(self.ctx, x.value, cs, np.complex128) for x in p[:4]], _View(self.ctx, p[4].value, rs,
                                                                                    np.float64)

    def set_fields(self, b, nd, mask):
        """n_d and the mask are copied into box b's solver."""
        if not 0 <= b < len(self.shapes):
            raise IndexError(f"box {b} of {len(self.shapes)}")
        rs = self.shapes[b]
        check(self.ctx.lib.kf_batch_set_fields(self.h, b, _need(nd, rs, np.float64, "n_d"),
                                               _need(mask, rs, np.float64, "mask")))

    def capture(self, poisson_iterations=4):
        check(self.ctx.lib.kf_batch_capture(self.h, poisson_iterations))

    def steps(self, n):
        check(self.ctx.lib.kf_batch_steps(self.h, n))

    def sync(self):
        check(self.ctx.lib.kf_batch_synchronize(self.h))

    def poisson_failures(self):
        c = C.c_int(0)
        check(self.ctx.lib.kf_batch_poisson_failures(self.h, C.byref(c)))
        return c.value


PROF_FFT_X, PROF_FFT_Y, PROF_FFT_ZR2C, PROF_FFT_ZC2R, PROF_POISSON, PROF_RB0 = 0, 1, 2, 3, 4, 5


def profile_enable(family: int):
    check(load().kf_profile_enable(family))


def profile_read():
    n = C.c_int(0)
    ms = C.c_double(0)
    check(load().kf_profile_read(C.byref(n), C.byref(ms)))
    return n.value, ms.value


def setup_host(axes, seed=0, amp=1e-3, Q=-0.1, a=1.0):
    """Initial ln n, grain charge n_d and solid mask on the host (kielflow.cpp:172-229)."""
    n = [ax.n for ax in axes]
    ln_n = np.empty(n, np.float64)
    nd = np.empty(n, np.float64)
    mask = np.empty(n, np.float64)
    check(load().kf_setup_host(_ax3(axes), seed, amp, Q, a, ln_n.ctypes.data_as(_vp),
                               nd.ctypes.data_as(_vp), mask.ctypes.data_as(_vp)))
    return ln_n, nd, mask
