"""torch.autograd wrapper of the GPU rollout and its discrete adjoint.

``Rollout(mesh, params, nsteps)`` is a callable ``(u0, kappa_e) -> u_T``
whose backward runs the library's adjoint (transposed solves with
checkpoint-and-recompute, fem_adjoint_* in include/femhip.h).  The forward is
the backward-Euler heat / advection-diffusion step of the reference
(CPUSolver.cpp:72-102) with the stiffness assembled per element as
K = sum_e kappa_e K_e.  Tensors cross the C ABI as host float64 arrays; the
arithmetic runs in the HIP library.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .solver import AdvectionDiffusionParameters, BoundaryCondition, FEMContext, SolverOptions, Adjoint

__all__ = ["Rollout"]


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, kappa_e, model):
        uT = model._run_forward(u0.detach().cpu().double().numpy(), kappa_e.detach().cpu().double().numpy())
        ctx.model = model
        ctx.fid = model._fid
        ctx.save_for_backward(u0, kappa_e)
        return torch.from_numpy(uT).to(device=u0.device, dtype=u0.dtype)

    @staticmethod
    def backward(ctx, grad_uT):
        model = ctx.model
        if model._fid != ctx.fid:  # another forward ran since: recompute this one
            u0, kappa_e = ctx.saved_tensors
            model._run_forward(u0.detach().cpu().double().numpy(), kappa_e.detach().cpu().double().numpy())
        gu, gk = model.adjoint.backward(grad_uT.detach().cpu().double().numpy())
        u0, kappa_e = ctx.saved_tensors
        return (torch.from_numpy(gu).to(device=u0.device, dtype=u0.dtype),
                torch.from_numpy(gk).to(device=kappa_e.device, dtype=kappa_e.dtype), None)


class Rollout:
    """u_T = rollout(u0, kappa_e) over ``nsteps`` steps, differentiable in
    both arguments (u0 in unknown space, kappa_e one per triangle)."""

    def __init__(self, mesh: dict, params, nsteps: int, checkpoint_every: int = 10,
                 boundary_condition: int = BoundaryCondition.Dirichlet, velocity=(1.0, 0.0, 0.0),
                 velocity_e=None, normalize_velocity: bool = True, options: Optional[SolverOptions] = None):
        self.nsteps = int(nsteps)
        self.every = int(checkpoint_every)
        self.velocity = velocity
        self.velocity_e = velocity_e
        self.normalize = normalize_velocity
        self.ctx = FEMContext(mesh, boundary_condition, velocity=velocity, velocity_e=velocity_e,
                              normalize_velocity=normalize_velocity)
        self.adjoint = Adjoint(self.ctx, params, options)
        self._kappa = None
        self._fid = 0

    @property
    def num_unknowns(self) -> int:
        return self.ctx.num_unknowns()

    def _run_forward(self, u0: np.ndarray, kappa_e: np.ndarray) -> np.ndarray:
        if self._kappa is None or not np.array_equal(self._kappa, kappa_e):
            self.ctx.assemble_matrices(self.velocity, self.velocity_e, self.normalize, kappa_e)
            self._kappa = kappa_e.copy()
        uT = self.adjoint.forward(u0, self.nsteps, self.every)
        self._fid += 1
        return uT

    def __call__(self, u0: torch.Tensor, kappa_e: torch.Tensor) -> torch.Tensor:
        return _RolloutFn.apply(u0, kappa_e, self)

    def close(self):
        self.adjoint.close()
        self.ctx.close()
