"""Gradients of the chemical fields: the field replay and its adjoint as a
torch.autograd.Function on PyTorch-ROCm.

A run records its owner trajectory (`HipSimulation.record`); `FieldFit`
replays the field steps over it on the GPU, holding the labels fixed (the
Metropolis sweep is discrete), as a function of the rates (`FieldParams`:
diffusion and decay per species, secretion and uptake per cell type and
species) and of the initial fields. `replay(fit, theta, u0)` is
differentiable in both: its backward pass is the hand-written adjoint of
glazier_amd/csrc/adjoint.hip, with checkpoints every `checkpoint_every`
steps. torch only carries device memory and streams here; the arithmetic
is the library's.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np

from .engine import HipSimulation, _check, library
from .model import CFieldParams, FieldParams

_bound = False


def _bind():
    global _bound
    if _bound:
        return library()
    lib = library()
    P, v = ctypes.POINTER, ctypes.c_void_p
    lib.ghip_record.argtypes = [v, ctypes.c_uint64]
    lib.ghip_recorded_steps.argtypes = [v, P(ctypes.c_uint64)]
    lib.ghip_recorded_types.argtypes = [v, ctypes.c_uint64, P(ctypes.c_uint8), ctypes.c_size_t]
    lib.ghip_fit_create.argtypes = [v, ctypes.c_uint32, P(v)]
    lib.ghip_fit_destroy.argtypes = [v]
    lib.ghip_fit_model_params.argtypes = [v, P(CFieldParams)]
    lib.ghip_fit_forward.argtypes = [v, P(CFieldParams), v, v, v]
    lib.ghip_fit_backward.argtypes = [v, v, P(CFieldParams), v, v]
    lib.ghip_fit_backward_noalloc.argtypes = [v, v, P(CFieldParams), v, v]
    _bound = True
    return lib


def record(sim: HipSimulation, max_steps: int) -> None:
    """Start recording the owner map of each step (0 stops)."""
    _check(_bind().ghip_record(sim._h, int(max_steps)))


def recorded_steps(sim: HipSimulation) -> int:
    out = ctypes.c_uint64()
    _check(_bind().ghip_recorded_steps(sim._h, ctypes.byref(out)))
    return int(out.value)


def recorded_types(sim: HipSimulation, step: int) -> np.ndarray:
    out = np.empty(sim.sites, dtype=np.uint8)
    _check(_bind().ghip_recorded_types(sim._h, int(step), out.ctypes.data_as(
        ctypes.POINTER(ctypes.c_uint8)), out.size))
    return out


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class FieldFit:
    """The recorded steps of `sim`, replayable as a function of the rates."""

    def __init__(self, sim: HipSimulation, checkpoint_every: int = 20):
        lib = _bind()
        self.model = sim.model
        self.sites = sim.sites
        self.n_species = len(sim.model.species)
        self.n_types = sim.model.n_types
        self.steps = recorded_steps(sim)
        self.checkpoint_every = int(checkpoint_every)
        h = ctypes.c_void_p()
        _check(lib.ghip_fit_create(sim._h, self.checkpoint_every, ctypes.byref(h)))
        self._h = h
        self.generation = 0

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(_bind().ghip_fit_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def model_params(self) -> FieldParams:
        c = CFieldParams()
        _check(_bind().ghip_fit_model_params(self._h, ctypes.byref(c)))
        return FieldParams.from_c(c, self.n_types, self.n_species)

    def forward(self, params: FieldParams, u0):
        """Fields after every recorded step, from `u0` (a float32 device
        tensor of n_species x n_sites)."""
        import torch

        u0 = u0.detach().to(device="cuda", dtype=torch.float32).contiguous()
        assert u0.numel() == self.n_species * self.sites
        out = torch.empty_like(u0)
        c = params.to_c()
        _check(_bind().ghip_fit_forward(self._h, ctypes.byref(c), ctypes.c_void_p(u0.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr()), _stream()))
        self.generation += 1
        return out

    def backward(self, grad_final):
        """(dL/dparams, dL/du0) of the last forward pass."""
        import torch

        g = grad_final.detach().to(device="cuda", dtype=torch.float32).contiguous()
        gu0 = torch.empty_like(g)
        gc = CFieldParams()
        _check(_bind().ghip_fit_backward(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.byref(gc),
                                         ctypes.c_void_p(gu0.data_ptr()), _stream()))
        return FieldParams.from_c(gc, self.n_types, self.n_species), gu0


def _function():
    import torch

    class FieldReplay(torch.autograd.Function):
        """u_final = replay(theta, u0) over a FieldFit's trajectory.

        theta is FieldParams.vector() as an f64 tensor: diffusion, decay,
        then the cell types' secretion and uptake rows. The backward pass
        reruns the forward first if another forward pass ran in between."""

        @staticmethod
        def forward(ctx, theta, u0, fit):
            params = FieldParams.from_vector(theta.detach().cpu().numpy(), fit.n_types, fit.n_species)
            out = fit.forward(params, u0)
            ctx.fit = fit
            ctx.params = params
            ctx.generation = fit.generation
            ctx.save_for_backward(u0)
            ctx.theta_device = theta.device
            return out

        @staticmethod
        def backward(ctx, grad_out):
            fit = ctx.fit
            (u0,) = ctx.saved_tensors
            if fit.generation != ctx.generation:
                fit.forward(ctx.params, u0)
            grad, gu0 = fit.backward(grad_out)
            gtheta = torch.as_tensor(grad.vector(), dtype=torch.float64, device=ctx.theta_device)
            return gtheta, gu0.view_as(u0), None

    return FieldReplay


_FieldReplay = None


def replay(fit: FieldFit, theta, u0):
    """Differentiable replay: fields after the recorded steps."""
    global _FieldReplay
    if _FieldReplay is None:
        _FieldReplay = _function()
    return _FieldReplay.apply(theta, u0, fit)


def fit_rates(fit: FieldFit, target, start: FieldParams, u0, steps: int = 50, lr: float = 0.05,
              free: Optional[np.ndarray] = None) -> List[dict]:
    """Adam on the log of the free rates (the nonzero ones of `start` unless
    `free` says otherwise) against the L2 loss to `target`. Returns one
    record per step: the loss and the rate vector before that step."""
    import torch

    v0 = start.vector()
    free = (v0 > 0.0) if free is None else np.asarray(free, dtype=bool)
    fixed = torch.as_tensor(np.where(free, 0.0, v0), dtype=torch.float64)
    log_rates = torch.tensor(np.log(np.where(free, v0, 1.0)), dtype=torch.float64, requires_grad=True)
    mask = torch.as_tensor(free)
    opt = torch.optim.Adam([log_rates], lr=lr)
    history = []
    for _ in range(steps):
        opt.zero_grad()
        theta = torch.where(mask, torch.exp(log_rates), fixed)
        out = replay(fit, theta, u0)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        history.append({"loss": float(loss.detach()), "theta": theta.detach().numpy().copy()})
        opt.step()
    theta = torch.where(mask, torch.exp(log_rates), fixed).detach()
    out = replay(fit, theta, u0)
    history.append({"loss": float(((out - target) ** 2).sum().detach()), "theta": theta.numpy().copy()})
    return history
