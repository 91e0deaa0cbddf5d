"""The engine's discrete adjoint as a torch.autograd.Function (validation
harness only: the engine itself never imports torch).

FinalPositions.apply(E, eps, sigma, w, nsteps, ckpt, skin) builds the
workload w with per-element Young's moduli E, pair eps and a common sigma,
runs nsteps steps through fm_adjoint_forward and returns the final
positions, caller order, as an (n, 2) float64 tensor; its backward passes
the incoming gradient to fm_adjoint_backward as the terminal cotangent and
returns dL/dE, dL/deps, dL/dsigma."""
import numpy as np
import torch

from femmd_old_amd import configs


class FinalPositions(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, eps, sigma, w, nsteps, ckpt, skin=None):
        v = configs.with_params(w, E=E.detach().cpu().numpy(), eps=float(eps), sigma=float(sigma))
        s = configs.build_system(v, skin=skin)
        s.adjoint_forward(int(nsteps), int(ckpt))
        x, y, _, _ = s.get_state()
        ctx.system = s
        return torch.from_numpy(np.stack([x, y], 1).astype(np.float64))

    @staticmethod
    def backward(ctx, g):
        g = g.detach().cpu().numpy()
        r = ctx.system.adjoint_backward(g[:, 0], g[:, 1])
        ctx.system.close()
        return (torch.from_numpy(r["dE"]), torch.tensor(r["deps"], dtype=torch.float64),
                torch.tensor(r["dsigma"], dtype=torch.float64), None, None, None, None)


def mse_min_image(pos, ids, target, L):
    """(1/K) sum_k |x_{ids[k]} - t_k|^2 with the minimum image (the rounding
    carries no gradient)."""
    d = pos[ids] - target
    d = d - L * torch.round(d / L)
    return (d * d).sum(1).mean()
