"""Differentiable halo operations (torch.autograd) over the HIP kernels.

exchange_halo is a copy from every owner into the ghost slots that hold its
images (plus a constant periodic shift for positions); reverse_halo is the
adjoint scatter-add of the ghost values into their owners.  So:

- ``HaloPositions``: owned positions [N, 3] -> owned + ghost positions
  [N+G, 3] along the plan of the last rebuild (refreshed in place, the
  reference's reused-plan refresh).  Backward = the reverse_halo kernel on the
  three gradient components (ordered, deterministic).
- ``HaloCopy``: an owned field [N, C] -> [N+G, C] (forward_exchange).
  Backward = reverse_exchange(Sum).
- ``ReverseSum``: [N+G, C] -> owned [N, C] with every ghost added to its owner
  (reverse_exchange(Sum) / reverse_halo).  Backward = forward_exchange.

Each Function keeps the LoopbackHalo it ran with (its device plan holds the
ghost -> owner maps built by the forward pass) and a plan generation number;
a backward after the plan was rebuilt raises instead of using stale maps.
"""
from __future__ import annotations

import torch

from . import _abi
from .halo import DeviceParticles, LoopbackHalo


def _generation(halo: LoopbackHalo) -> int:
    return getattr(halo, "_generation", 0)


def bump_generation(halo: LoopbackHalo) -> None:
    halo._generation = _generation(halo) + 1


def _check_same_plan(ctx) -> None:
    if _generation(ctx.halo) != ctx.generation:
        raise RuntimeError("dsdecompx_amd: the halo plan was rebuilt after the forward pass")


def _reverse_sum(halo: LoopbackHalo, values: torch.Tensor, owned: int) -> torch.Tensor:
    """[N+G, C] float64 -> [N, C]: owner value plus its ghosts, ascending send slot."""
    if values.dim() == 2 and values.shape[1] == 3 and values.dtype == torch.float64:
        # one transposing copy into SoA rows, the reverse_halo kernel on the
        # rows, and a transposed view of the owned part as the result
        rows = torch.empty((3, values.shape[0]), dtype=torch.float64, device=values.device)
        rows.copy_(values.t())   # always a copy: the caller's tensor is not modified
        _abi.check(_abi.lib().dsdx_reverse_halo(halo.plan, rows[0].data_ptr(), rows[1].data_ptr(),
                                                rows[2].data_ptr(), owned,
                                                torch.cuda.current_stream().cuda_stream),
                   "reverse_halo")
        return rows[:, :owned].t()
    values = values.contiguous().clone()
    if values.dim() == 1:
        values = values[:, None]
    halo.reverse_exchange([(values, _abi.SUM)], owned)
    return values[:owned]


def _forward_copy(halo: LoopbackHalo, owned_values: torch.Tensor, ghosts: int) -> torch.Tensor:
    """[N, C] -> [N+G, C]: ghosts receive their owner's value."""
    n = owned_values.shape[0]
    out = torch.empty((n + ghosts,) + tuple(owned_values.shape[1:]), dtype=owned_values.dtype,
                      device=owned_values.device)
    out[:n] = owned_values
    halo.forward_exchange([out], n)
    return out


class HaloPositions(torch.autograd.Function):
    """Owned positions -> owned + ghost positions along the saved plan."""

    @staticmethod
    def forward(ctx, pos: torch.Tensor, halo: LoopbackHalo, particles: DeviceParticles):
        n = halo.owned
        g = halo.ghosts
        if pos.shape != (n, 3) or pos.dtype != torch.float64:
            raise ValueError("positions must be float64 [owned, 3]")
        particles.x[:n] = pos[:, 0]
        particles.y[:n] = pos[:, 1]
        particles.z[:n] = pos[:, 2]
        halo.refresh(particles, n, positions_only=True)
        ctx.halo = halo
        ctx.generation = _generation(halo)
        ctx.owned = n
        return torch.stack([particles.x[:n + g], particles.y[:n + g], particles.z[:n + g]], 1)

    @staticmethod
    def backward(ctx, grad):
        _check_same_plan(ctx)
        return _reverse_sum(ctx.halo, grad.to(torch.float64), ctx.owned), None, None


class HaloCopy(torch.autograd.Function):
    """forward_exchange as a function of the owned values."""

    @staticmethod
    def forward(ctx, owned_values: torch.Tensor, halo: LoopbackHalo):
        ctx.halo = halo
        ctx.generation = _generation(halo)
        ctx.owned = owned_values.shape[0]
        return _forward_copy(halo, owned_values.contiguous(), halo.ghosts)

    @staticmethod
    def backward(ctx, grad):
        _check_same_plan(ctx)
        out = _reverse_sum(ctx.halo, grad, ctx.owned)
        return out.reshape((ctx.owned,) + tuple(grad.shape[1:])), None


class ReverseSum(torch.autograd.Function):
    """reverse_exchange(Sum) / reverse_halo as a function of all values."""

    @staticmethod
    def forward(ctx, values: torch.Tensor, halo: LoopbackHalo):
        ctx.halo = halo
        ctx.generation = _generation(halo)
        ctx.owned = halo.owned
        out = _reverse_sum(halo, values, halo.owned)
        return out.reshape((halo.owned,) + tuple(values.shape[1:]))

    @staticmethod
    def backward(ctx, grad):
        _check_same_plan(ctx)
        return _forward_copy(ctx.halo, grad.contiguous(), ctx.halo.ghosts), None


def rebuild(halo: LoopbackHalo, particles: DeviceParticles, owned: int, fields=()) -> int:
    """exchange_halo(rebuild) that also invalidates earlier autograd graphs."""
    g = halo.exchange_halo(particles, owned, fields=fields)
    bump_generation(halo)
    return g


def halo_positions(pos, halo, particles):
    return HaloPositions.apply(pos, halo, particles)


def halo_copy(values, halo):
    return HaloCopy.apply(values, halo)


def reverse_sum(values, halo):
    return ReverseSum.apply(values, halo)
