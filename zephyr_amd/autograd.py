"""torch.autograd binding of the discrete adjoint (used for checks).

FinalPressure.apply(rho0, ctx, n_steps, every, window, p0, gamma) returns
J = integral over `window` of p^2 at the final time for the initial state
rho = rho0, u = v = 0, p = p0, e = e(rho, p); its backward is the device
adjoint (zfv_adjoint_forward / zfv_adjoint_backward) with the chain rule
through e(rho, p) = e0 + (p + gamma P0) / ((gamma - 1) rho).
"""
import numpy as np
import torch

from . import problems


class FinalPressure(torch.autograd.Function):
    @staticmethod
    def forward(fctx, rho0, zctx, n_steps, every, window, p0=1.0, gamma=1.4):
        rho = rho0.detach().cpu().numpy().astype(np.float64)
        p = np.full_like(rho, p0)
        zero = np.zeros_like(rho)
        zctx.set_state(rho, zero, zero.copy(), p, problems.energy_rP(rho, p, gamma))
        J = zctx.adjoint_forward(n_steps, every, window)
        fctx.zctx = zctx
        fctx.save_for_backward(rho0)
        fctx.p0, fctx.gamma = p0, gamma
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(fctx, gJ):
        (rho0,) = fctx.saved_tensors
        lr, lu, lv, lp, le = fctx.zctx.adjoint_backward()
        rho = rho0.detach().cpu().numpy().astype(np.float64)
        # e = (p + gamma P0) / ((gamma - 1) rho) with P0 = e0 = 0: de/drho = -e / rho
        e = problems.energy_rP(rho, np.full_like(rho, fctx.p0), fctx.gamma)
        g = lr - le * e / rho
        return gJ * torch.from_numpy(g).to(rho0.dtype), None, None, None, None, None, None
