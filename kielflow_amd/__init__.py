"""kielflow_amd -- MI355X-native pseudo-spectral RHS, RK4 step and nonlinear
Poisson solve of kielflow, behind a C ABI (include/kielflow_amd.h).

The compute path is libkielflow_amd.so (hand-written HIP for gfx950);
`kielflow_amd.api` binds it with ctypes.  There is no CPU fallback.
"""
from .api import (EQST, HYST, KF_FP32, KF_FP64, KF_MIXED, POISSON_CASCADE, POISSON_NEWTON_MG,
                  POISSON_NLJ, SIST, FFT, Adjoint, Batch, Context, DeviceArray, Euler, Grid, KfError, Poisson, Rhs,
                  Rk4, axis, axis_index_at, axis_values, cell_volumes, cube, load, setup_host)

__all__ = ["EQST", "HYST", "SIST", "KF_FP32", "KF_FP64", "KF_MIXED", "POISSON_NEWTON_MG",
           "POISSON_NLJ", "POISSON_CASCADE", "FFT", "Adjoint", "Batch", "Context", "DeviceArray", "Grid", "KfError",
           "Poisson", "Rhs", "Rk4", "Euler", "axis", "axis_index_at", "axis_values", "cell_volumes", "cube", "load",
           "setup_host"]
