"""torch.autograd wrapper of the C-ABI adjoint (test harness only).

CulvertDischarge.apply(alpha_window, u_inlet) returns the time-averaged
discharge J of a culvert run on the GPU; backward() hands torch the gradients
the hand-written adjoint (flb_adjoint) computed in the same call."""
import numpy as np
import torch

from flowlb_amd import core


class CulvertDischarge(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha_window, u_inlet, case, window, steps, t0, t1, ckpt):
        flags, rho0, u0 = case.build()
        N = case.N
        u0 = u0.copy()
        u0[:N][(flags & core.INLET) != 0] = np.float32(float(u_inlet))
        x0, x1, y0, y1 = window
        alpha = np.zeros((case.ny, case.nx), np.float32)
        alpha[y0:y1 + 1, x0:x1 + 1] = alpha_window.detach().cpu().numpy()
        with core.Solver(flags, case.nx, case.ny, case.options(), rho0=rho0, u0=u0) as s:
            s.set_brinkman(alpha.ravel())
            J, g, duin, _ = s.adjoint(steps, case.nx - 2, t0, t1, ckpt)
        gw = torch.from_numpy(g.reshape(case.ny, case.nx)[y0:y1 + 1, x0:x1 + 1].astype(np.float64))
        ctx.save_for_backward(gw, torch.tensor(duin, dtype=torch.float64))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_j):
        gw, gu = ctx.saved_tensors
        return grad_j * gw, grad_j * gu, None, None, None, None, None, None
