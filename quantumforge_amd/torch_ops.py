"""PyTorch integration: VQE / QAOA energies as a differentiable torch op.

`VQEEnergy.apply(params, ...)` returns the energies <psi(theta_b)|H|psi(theta_b)>
of a batch of parameter sets; its backward uses the gradients the HIP
engine's two-state adjoint sweep computed in the same call (one forward
sweep, one energy kernel, one reverse sweep for all sets), so
`energies.sum().backward()` costs no extra circuit runs.
"""
from __future__ import annotations

import numpy as np
import torch

from .core import ParameterizedCircuit, PauliSum
from .gradients import GradientStrategy, batch_gradient


class VQEEnergy(torch.autograd.Function):
    """energies[b] = <H> at params[b]; params is a float64 tensor [sets, parameters]."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, backend, circuit: ParameterizedCircuit, qubits: int, hamiltonian: PauliSum,
                precision: str = "complex128", strategy: GradientStrategy = GradientStrategy.Adjoint):
        if params.dtype != torch.float64:
            raise TypeError("parameters must be float64 (the engine binds angles in FP64)")
        p = params.detach().to("cpu").contiguous().numpy()
        squeeze = p.ndim == 1
        p2 = np.atleast_2d(p)
        energies, grads = batch_gradient(backend, circuit, qubits, p2, hamiltonian, strategy, precision)
        g = torch.from_numpy(grads).to(params.device)
        e = torch.from_numpy(energies).to(params.device)
        if squeeze:
            g, e = g[0], e[0]
        ctx.save_for_backward(g)
        return e

    @staticmethod
    def backward(ctx, grad_energy: torch.Tensor):
        (g,) = ctx.saved_tensors
        if g.dim() == 1:
            return grad_energy * g, None, None, None, None, None, None
        return grad_energy[:, None] * g, None, None, None, None, None, None


def vqe_energy(params: torch.Tensor, backend, circuit: ParameterizedCircuit, qubits: int, hamiltonian: PauliSum,
               precision: str = "complex128") -> torch.Tensor:
    """Differentiable batched energies (see VQEEnergy)."""
    return VQEEnergy.apply(params, backend, circuit, qubits, hamiltonian, precision)
