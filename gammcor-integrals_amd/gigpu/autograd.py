"""torch.autograd interface of the differentiation form.

`coulomb_energy(coords, context, basis, density)` returns
E_J = 1/2 sum_{pqrs} D_pq (pq|rs) D_rs as a differentiable function of the
atom coordinates (natoms x 3, bohr).  The forward pass is the energy-only
launch of gi_coulomb_energy_grad; the backward pass is the hand-written
gradient kernels behind the same call (derivative integrals evaluated on the
fly, Hermite tables recomputed rather than saved from the forward pass).
Both run on the GPU of `context`; torch only carries the tensors.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import lib


def _at(basis, coords: torch.Tensor):
    r = coords.detach().to("cpu", torch.float64).numpy().reshape(-1, 3)
    if r.shape[0] != basis.natoms:
        raise ValueError("coords must be natoms x 3 (%d atoms)" % basis.natoms)
    return dataclasses.replace(basis, atom_coords=np.array(r))


def _density(density) -> np.ndarray:
    if isinstance(density, torch.Tensor):
        return density.detach().to("cpu", torch.float64).numpy()
    return np.asarray(density, dtype=np.float64)


class CoulombEnergy(torch.autograd.Function):
    """E_J(R) for a fixed AO density; gradient with respect to R only."""

    @staticmethod
    def forward(ctx, coords, context, basis, density, kappa):
        D = _density(density)
        context.set_basis(_at(basis, coords))
        e = context.coulomb_energy(D, kappa, grad=False)
        ctx.context, ctx.basis, ctx.density, ctx.kappa = context, basis, D, kappa
        ctx.save_for_backward(coords)
        return torch.tensor(e, dtype=coords.dtype, device=coords.device)

    @staticmethod
    def backward(ctx, grad_energy):
        (coords,) = ctx.saved_tensors
        ctx.context.set_basis(_at(ctx.basis, coords))
        _, g = ctx.context.coulomb_energy(ctx.density, ctx.kappa, grad=True)
        g = torch.from_numpy(g).to(device=coords.device, dtype=coords.dtype).reshape(coords.shape)
        return grad_energy * g, None, None, None, None


def coulomb_energy(coords: torch.Tensor, context: "lib.Context", basis, density, kappa: float = 0.0):
    """Differentiable E_J(coords); `basis` supplies everything but the
    coordinates, `context` the device."""
    return CoulombEnergy.apply(coords, context, basis, density, kappa)
