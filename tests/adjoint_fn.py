"""torch.autograd.Function over the library's checkpointed adjoint (test harness).

``WallShear.apply(H, scale, sim, reset, nsteps, window, K)`` runs the coupled
steps from the state ``reset()`` restores, with the per-node body-force field
H ([3, n] padded) and the per-cell force scales, and returns the wall-shear
objective J; its backward returns dJ/dH and dJ/dscale from
``lbmibm_adjoint_wall_shear`` (one forward sweep plus the reverse sweep with
recomputation every K steps).
"""
from __future__ import annotations

import numpy as np
import torch


class WallShear(torch.autograd.Function):
    @staticmethod
    def forward(ctx, H, scale, sim, reset, nsteps, window, K):
        reset()
        h = H.detach().cpu().numpy().astype(np.float64)
        sim.set_body_force_field(h[0], h[1], h[2])
        sim.set_cell_force_scale(scale.detach().cpu().numpy().astype(np.float64))
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            J, gH, gs = sim.adjoint_wall_shear(0, nsteps, window, K)
            ctx.save_for_backward(torch.from_numpy(np.stack(gH)), torch.from_numpy(np.array(gs)))
        else:
            J = sim.wall_shear(0, nsteps, window)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gH, gs = ctx.saved_tensors
        return gJ * gH, gJ * gs, None, None, None, None, None
