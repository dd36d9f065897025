"""torch.autograd.Function around the differentiable rollout.

A gradient harness only: the rollout and its adjoint run in libbrasspan.so
(bp_forward_checkpointed / bp_backward); torch carries the parameters and
gradients on the host.

    theta = torch.tensor(..., requires_grad=True)   # (scenes, 6): wind xyz, drag, lift, xsph_c
    v0 = torch.zeros(n, 3, requires_grad=True)      # initial velocities, original order
    xT = Rollout.apply(theta, v0, scenes, params, 32, 1 / 120, 8)
    loss = ((xT - target) ** 2).sum(dim=1).mean()
    loss.backward()                                  # theta.grad, v0.grad
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import brasspan as bp


def _params_with(p: bp.bp_params, th) -> bp.bp_params:
    q = bp.bp_params.from_buffer_copy(p)
    for a in range(3):
        q.wind[a] = float(th[a])
    q.drag_coeff = float(th[3])
    q.lift_coeff = float(th[4])
    q.xsph_c = float(th[5])
    return q


class Rollout(torch.autograd.Function):
    """x_T (N x 3, original order) of `substeps` substeps of `subdt` from the
    scenes' initial positions, with per-scene parameters theta and initial
    velocities v0; one solver handle (batched when there are several scenes)."""

    @staticmethod
    def forward(ctx, theta, v0, scene_list, params_list, substeps, subdt, every):
        th = theta.detach().cpu().numpy().astype(np.float32).reshape(len(scene_list), 6)
        prms = [_params_with(p, th[k]) for k, p in enumerate(params_list)]
        if len(scene_list) == 1:
            s = bp.Solver(scene_list[0], prms[0])
        else:
            s = bp.BatchSolver(scene_list, prms)
        pos, vel, mp = s.get_state()
        v = v0.detach().cpu().numpy().astype(np.float32).reshape(-1, 3)
        s.set_state(pos, v[mp], mp)
        s.forward_checkpointed(int(substeps), float(subdt), int(every))
        xT, _ = s.state_original_order()
        ctx.solver = s
        return torch.from_numpy(xT.copy())

    @staticmethod
    def backward(ctx, grad_xT):
        gx0, gv0, gth = ctx.solver.backward(grad_xT.detach().cpu().numpy())
        ctx.solver.close()
        return torch.from_numpy(gth), torch.from_numpy(gv0), None, None, None, None, None
