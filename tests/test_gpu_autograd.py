"""gradcheck of the torch.autograd.Function over the checkpointed adjoint:
J(M, Q, ln n0) on a 16^3 box, the field input parameterised by three smooth
seeded modes so that torch's finite differences stay few."""
import numpy as np
import pytest

from helpers import axes_for, cell_volumes, initial_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["eqst", "hyst"])
def test_gradcheck_grain_drag(ctx, kind):
    from kielflow_amd import autograd as KA
    import torch
    print("gradcheck", kind, flush=True)
    n, Q0 = 16, -0.2
    state, nd, mask, og = initial_state(kind, n, seed=3, Q=Q0)
    axes, _ = axes_for(kind, n)
    prob = KA.DragProblem(ctx, axes, nd / Q0, cell_volumes(og), steps=5, k=2, mask=mask)
    base = og.c2r(state[3])
    rng = np.random.default_rng(11)
    modes = np.stack([og.c2r(og.dealias23(og.r2c(rng.standard_normal(og.rshape)))) for _ in range(3)])
    modes *= 1e-3 / np.abs(modes).max()
    tb = torch.from_numpy(base)
    tm = torch.from_numpy(modes)

    def J(M, Q, c):
        ln_n0 = tb + torch.tensordot(c, tm, dims=1)
        return KA.grain_drag(M, Q, ln_n0, prob)

    M = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    Q = torch.tensor(Q0, dtype=torch.float64, requires_grad=True)
    c = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    out = J(M, Q, c)
    out.backward()
    print(f"J={out.item():.6e} dJ/dM={M.grad.item():.6e} dJ/dQ={Q.grad.item():.6e} dJ/dc={c.grad.tolist()}",
          flush=True)
    assert abs(M.grad.item()) > 0 and abs(Q.grad.item()) > 0 and c.grad.abs().max() > 0
    scale = max(abs(M.grad.item()), abs(Q.grad.item()), c.grad.abs().max().item())
    assert torch.autograd.gradcheck(J, (M, Q, c), eps=1e-4, atol=1e-6 * scale, rtol=1e-4)
