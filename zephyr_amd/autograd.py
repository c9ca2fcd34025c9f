"""torch.autograd binding of the discrete adjoint (used for checks).

FinalPressure.apply(rho0, ctx, n_steps, every, window, p0, gamma, u0, v0)
returns J = integral over `window` of p^2 at the final time for the initial
state rho = rho0, u = u0, v = v0 (default 0), p = p0 (default 1), e = e(rho, p),
where p0, u0, v0 are numbers or per-cell arrays held fixed; its backward is the
device
adjoint (zfv_adjoint_forward / zfv_adjoint_backward) with the chain rule
through e(rho, p) = e0 + (p + gamma P0) / ((gamma - 1) rho).
"""
import numpy as np
import torch

from . import problems


class FinalPressure(torch.autograd.Function):
    @staticmethod
    def forward(fctx, rho0, zctx, n_steps, every, window, p0=1.0, gamma=1.4, u0=0.0, v0=0.0):
        rho = rho0.detach().cpu().numpy().astype(np.float64)

        def field(x):
            return np.broadcast_to(np.asarray(x, dtype=np.float64), rho.shape).copy()

        p = field(p0)
        zctx.set_state(rho, field(u0), field(v0), p, problems.energy_rP(rho, p, gamma))
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
        p = np.broadcast_to(np.asarray(fctx.p0, dtype=np.float64), rho.shape)
        e = problems.energy_rP(rho, p, fctx.gamma)
        g = lr - le * e / rho
        return gJ * torch.from_numpy(g).to(rho0.dtype), None, None, None, None, None, None, None, None
