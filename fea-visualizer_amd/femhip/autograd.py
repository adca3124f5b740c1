"""torch.autograd wrapper of the GPU rollout and its discrete adjoint.

``Rollout(mesh, params, nsteps)`` is a callable ``(u0, kappa_e) -> u_T``
whose backward runs the library's adjoint (transposed solves with
checkpoint-and-recompute, fem_adjoint_* in include/femhip.h).  The forward is
the backward-Euler heat / advection-diffusion step of the reference
(CPUSolver.cpp:72-102) with the stiffness assembled per element as
K = sum_e kappa_e K_e.  GPU tensors stay on the device: their pointers cross
the C ABI (fem_adjoint_forward_device / _backward_device); only a changed
kappa_e comes to the host, once, for the re-assembly.  CPU tensors cross as
host float64 arrays.  The arithmetic runs in the HIP library either way.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .solver import BoundaryCondition, FEMContext, SolverOptions, Adjoint

__all__ = ["Rollout"]


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, kappa_e, model):
        uT = model._run_forward(u0, kappa_e)
        ctx.model = model
        ctx.fid = model._fid
        ctx.save_for_backward(u0, kappa_e)
        return uT

    @staticmethod
    def backward(ctx, grad_uT):
        model = ctx.model
        u0, kappa_e = ctx.saved_tensors
        if model._fid != ctx.fid:  # another forward ran since: recompute this one
            model._run_forward(u0, kappa_e)
        gu, gk = model._run_backward(grad_uT)
        return gu.to(device=u0.device, dtype=u0.dtype), gk.to(device=kappa_e.device, dtype=kappa_e.dtype), None


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
        self.ntri = int(np.asarray(mesh["tri"]).shape[0])
        self.adjoint = Adjoint(self.ctx, params, options)
        self._kappa = None  # kappa_e of the current assembly (a tensor on its device)
        self._fid = 0
        self._device = None

    @property
    def num_unknowns(self) -> int:
        return self.ctx.num_unknowns()

    def _assemble(self, kappa_e: torch.Tensor):
        k = kappa_e.detach()
        if self._kappa is None or self._kappa.device != k.device or not torch.equal(self._kappa, k):
            self.ctx.assemble_matrices(self.velocity, self.velocity_e, self.normalize,
                                       k.to(device="cpu", dtype=torch.float64).numpy())
            self._kappa = k.clone()

    @staticmethod
    def _check_one_runtime():
        if len(L.hip_runtimes()) > 1:
            raise RuntimeError("device tensors need one HIP runtime in the process: import torch before the femhip "
                               "library is first used (it was loaded first, and torch then mapped its own runtime)")

    def _run_forward(self, u0: torch.Tensor, kappa_e: torch.Tensor) -> torch.Tensor:
        self._assemble(kappa_e)
        if u0.is_cuda:
            self._check_one_runtime()
            u = u0.detach().to(torch.float64).contiguous()
            uT = torch.empty_like(u)
            torch.cuda.current_stream(u.device).synchronize()  # the library reads on its own stream
            self.adjoint.forward_device(u.data_ptr(), uT.data_ptr(), self.nsteps, self.every)
            out = uT.to(u0.dtype)
        else:
            out = torch.from_numpy(self.adjoint.forward(u0.detach().double().numpy(), self.nsteps,
                                                        self.every)).to(u0.dtype)
        self._fid += 1
        return out

    def _run_backward(self, grad_uT: torch.Tensor):
        if grad_uT.is_cuda:
            self._check_one_runtime()
            g = grad_uT.detach().to(torch.float64).contiguous()
            gu = torch.empty_like(g)
            gk = torch.empty(self.ntri, dtype=torch.float64, device=g.device)
            torch.cuda.current_stream(g.device).synchronize()
            self.adjoint.backward_device(g.data_ptr(), gu.data_ptr(), gk.data_ptr())
            return gu, gk
        gu, gk = self.adjoint.backward(grad_uT.detach().double().numpy())
        return torch.from_numpy(gu), torch.from_numpy(gk)

    def __call__(self, u0: torch.Tensor, kappa_e: torch.Tensor) -> torch.Tensor:
        return _RolloutFn.apply(u0, kappa_e, self)

    def close(self):
        self.adjoint.close()
        self.ctx.close()
