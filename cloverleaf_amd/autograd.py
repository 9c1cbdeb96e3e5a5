"""torch.autograd binding of the adjoint form.

    import torch, cloverleaf_amd as clv
    from cloverleaf_amd.autograd import kinetic_energy
    adj = clv.Adjoint(started_chunk, dt, steps, checkpoint_every)
    rho = torch.tensor(rho0, device="cuda", requires_grad=True)   # (ny, nx) float64
    e = torch.tensor(e0, device="cuda", requires_grad=True)
    ke = kinetic_energy(adj, rho, e)     # forward: fixed-dt run, final kinetic energy
    ke.backward()                         # rho.grad, e.grad: the checkpointed adjoint

The forward and backward passes run the library's HIP kernels on the library's
stream; torch's device work is synchronized around each call.  Tensors on the
CPU are copied through host memory.  One Adjoint holds the checkpoints of one
forward run: when another forward has used it since, backward first re-runs the
forward from the saved inputs.
"""
from __future__ import annotations

import torch

from . import Adjoint


def _sync(t: torch.Tensor):
    if t.is_cuda:
        torch.cuda.synchronize(t.device)


def _run_forward(adj: Adjoint, d: torch.Tensor, e: torch.Tensor) -> list:
    if d.is_cuda:
        _sync(d)
        adj.set_initial(d.data_ptr(), e.data_ptr(), pitch=adj.nx)
    else:
        adj.set_initial(d.numpy(), e.numpy())
    return adj.forward()


class KineticEnergy(torch.autograd.Function):
    """ke(density0, energy0) of a fixed-dt run; backward = clv_adjoint_backward."""

    @staticmethod
    def forward(ctx, adj: Adjoint, density0: torch.Tensor, energy0: torch.Tensor) -> torch.Tensor:
        if density0.dtype != torch.float64 or energy0.dtype != torch.float64:
            raise TypeError("kinetic_energy: density0 and energy0 must be float64")
        if tuple(density0.shape) != (adj.ny, adj.nx) or tuple(energy0.shape) != (adj.ny, adj.nx):
            raise ValueError(f"kinetic_energy: fields must be ({adj.ny}, {adj.nx})")
        d = density0.detach().contiguous()
        e = energy0.detach().contiguous()
        s = _run_forward(adj, d, e)
        ctx.adj = adj
        ctx.generation = adj.generation
        ctx.device = d.device
        ctx.save_for_backward(d, e)
        return torch.tensor(s[3], dtype=torch.float64, device=d.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        adj: Adjoint = ctx.adj
        ke_bar = float(grad_out.detach().cpu())
        if adj.generation != ctx.generation:
            d, e = ctx.saved_tensors
            _run_forward(adj, d, e)
        if ctx.device.type == "cuda":
            gd = torch.empty((adj.ny, adj.nx), dtype=torch.float64, device=ctx.device)
            ge = torch.empty_like(gd)
            _sync(gd)
            adj.backward(ke_bar, gd.data_ptr(), ge.data_ptr(), pitch=adj.nx)
        else:
            gdn, gen = adj.backward(ke_bar)
            gd, ge = torch.from_numpy(gdn), torch.from_numpy(gen)
        return None, gd, ge


def kinetic_energy(adj: Adjoint, density0: torch.Tensor, energy0: torch.Tensor) -> torch.Tensor:
    return KineticEnergy.apply(adj, density0, energy0)
