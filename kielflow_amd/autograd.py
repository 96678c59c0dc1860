"""torch.autograd.Function over the C ABI's checkpointed adjoint, for gradcheck.

`DragProblem` fixes a box, the grain shape G (unit charge), the quadrature
cell volumes dV, the step count and the checkpoint interval.  `GrainDrag`
maps (M, Q, ln n0) -> J, the x-drag on the grain after `steps` RK4 steps
from the state u = 0, Fn = OP_FFT(ln n0) (kielflow.cpp:172-193), with
n_d = Q G.  Its backward returns the gradients that one forward+reverse sweep
of kf_adjoint_run produced: dJ/dM, dJ/dQ and dJ/d(ln n0) (real field).

torch is used for host tensors only; every field operation runs in the HIP
library.  The library is loaded before torch is imported so that the HIP
runtime it was built against is the one that serves the device.
"""
import numpy as np

from . import api as _api

_api.load()
import torch  # noqa: E402  (after the HIP library, see the module docstring)


class DragProblem:
    def __init__(self, ctx, axes, G, dV, steps, k=4, dt=1e-3, tau=0.1, theta=100.0, sv_on=False, pen_on=True,
                 mask=None, poisson_tol=1e-13, adjoint_tol=1e-12):
        self.ctx = ctx
        self.grid = _api.Grid(ctx, axes)
        self.solver = _api.Poisson(self.grid, _api.KF_FP64, tol=poisson_tol, max_it=200)
        self.rhs = _api.Rhs(self.grid, self.solver, M=1.0, tau=tau, theta=theta, sv_on=sv_on, pen_on=pen_on)
        self.phi = self.grid.rfield()
        self.G = np.ascontiguousarray(G, np.float64)
        self.dG = ctx.to_device(self.G)
        self.dV = ctx.to_device(np.ascontiguousarray(dV, np.float64))
        self.mask = ctx.to_device(np.zeros(self.grid.rshape) if mask is None else mask)
        self.rhs.set_fields(self.phi, ctx.to_device(-0.1 * self.G), self.mask)
        self.adj = _api.Adjoint(self.rhs, dt, tol=adjoint_tol)
        self.fft = _api.FFT(self.grid, _api.KF_FP64)
        self.steps, self.k, self.dt = steps, k, dt

    def _setup(self, M, Q, ln_n0):
        self.rhs.set_parameters(M=float(M))
        self.rhs.set_fields(None, self.ctx.to_device(float(Q) * self.G), None)
        self.phi.upload(np.zeros(self.grid.rshape))
        y = [self.grid.cfield() for _ in range(4)]
        self.fft.r2c(self.ctx.to_device(np.ascontiguousarray(ln_n0, np.float64)), y[3])
        return y

    def value(self, M, Q, ln_n0):
        """J only (forward steps and the objective)."""
        y = self._setup(M, Q, ln_n0)
        rk = _api.Rk4(self.rhs, self.dt)
        rk.step(y, self.steps)
        rk.close()
        return self.adj.objective(y, self.dV)

    def value_and_grad(self, M, Q, ln_n0):
        """J, dJ/dM, dJ/dQ, dJ/d(ln n0) (real, host array)."""
        y = self._setup(M, Q, ln_n0)
        dF = self.grid.cfield()
        J, dM, dQ = self.adj.run(y, self.steps, self.k, self.dV, self.dG, dF)
        dl = self.grid.rfield()
        self.adj.spectral_to_real(dF, dl)
        return J, dM, dQ, dl.get()


class GrainDrag(torch.autograd.Function):
    """J(M, Q, ln n0); M and Q are 0-d float64 tensors, ln n0 a real
    [Nx, Ny, Nz] float64 tensor on the host."""

    @staticmethod
    def forward(ctx, M, Q, ln_n0, problem):
        args = (float(M), float(Q), ln_n0.detach().cpu().numpy())
        if any(ctx.needs_input_grad[:3]):
            J, dM, dQ, dl = problem.value_and_grad(*args)
            ctx.save_for_backward(torch.tensor(dM, dtype=torch.float64), torch.tensor(dQ, dtype=torch.float64),
                                  torch.from_numpy(dl))
        else:
            J = problem.value(*args)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        dM, dQ, dl = ctx.saved_tensors
        return gJ * dM, gJ * dQ, gJ * dl, None


def grain_drag(M, Q, ln_n0, problem):
    return GrainDrag.apply(M, Q, ln_n0, problem)
