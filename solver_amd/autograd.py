"""torch.autograd bindings of the discrete adjoint (include/of_hip.h, of_adjoint_*).

`rollout` advances a periodic 3D Taylor-Green state by `steps` device steps and returns the
final velocity; its backward is the checkpointed discrete adjoint, not a trace of the forward.
`rollout_kinetic_energy` returns the final kinetic energy
(solver/taylor_green.cpp:269-288) with the same backward seeded by the energy's gradient.

Tensors are float64 device tensors in the reference storage layout (sz, sy, sx). Gradients are
with respect to the active entries of the inputs; inactive (ghost) entries get zero gradient,
because the forward reads only active entries.

Example::

    import solver_amd as sa
    from solver_amd.autograd import rollout_kinetic_energy

    ctx = sa.Context(0)
    adj = sa.Adjoint(ctx, 128, 128, 128, 1 / 400, dt, checkpoint_every=20)
    u0.requires_grad_()
    ke = rollout_kinetic_energy(adj, 200, u0, v0, w0)
    ke.backward()          # u0.grad = d KE(T) / d u0
"""
from __future__ import annotations

import torch

from . import Adjoint


def _check_run(ctx):
    if ctx.adj._generation != ctx.generation:
        raise RuntimeError("the Adjoint object ran another forward pass before this backward; use one Adjoint per "
                           "live graph")


def _forward_common(ctx, adj: Adjoint, steps, u0, v0, w0, p0, fu, fv, fw):
    forced = fu is not None
    if forced and (fv is None or fw is None):
        raise ValueError("forcing needs all three components")
    adj.set_forcing((fu.detach().contiguous(), fv.detach().contiguous(), fw.detach().contiguous())
                    if forced else None)
    vel = [t.detach().contiguous() for t in (u0, v0, w0)]
    p = p0.detach().contiguous() if p0 is not None else None
    ke, _ = adj.forward(vel, p, steps)
    adj._generation = getattr(adj, "_generation", 0) + 1
    ctx.adj = adj
    ctx.generation = adj._generation
    ctx.has_p = p0 is not None
    ctx.forced = forced
    return ke


def _grads(ctx, seed):
    _check_run(ctx)
    gu, gp, gf = ctx.adj.backward(seed=seed, want_pressure=ctx.has_p, want_forcing=ctx.forced)
    gf = gf if ctx.forced else (None, None, None)
    return gu[0], gu[1], gu[2], gp, gf[0], gf[1], gf[2]


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, adj, steps, u0, v0, w0, p0, fu, fv, fw):
        _forward_common(ctx, adj, steps, u0, v0, w0, p0, fu, fv, fw)
        return tuple(adj.final_velocity())

    @staticmethod
    def backward(ctx, gu, gv, gw):
        seed = [gu, gv, gw]
        if any(g is None for g in seed):
            like = ctx.adj.final_velocity()
            seed = [g if g is not None else torch.zeros_like(t) for g, t in zip(seed, like)]
        return (None, None) + _grads(ctx, seed)


class _RolloutKE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, adj, steps, u0, v0, w0, p0, fu, fv, fw):
        ke = _forward_common(ctx, adj, steps, u0, v0, w0, p0, fu, fv, fw)
        return torch.tensor(ke, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, g):
        grads = _grads(ctx, None)
        scale = g.to(torch.float64)
        return (None, None) + tuple(x * scale if x is not None else None for x in grads)


def rollout(adj: Adjoint, steps: int, u0, v0, w0, p0=None, forcing=None):
    """Final velocity (u, v, w) after `steps` steps; differentiable w.r.t. u0, v0, w0, p0 and the
    forcing field (a tuple of three face tensors, or None)."""
    fu, fv, fw = forcing if forcing is not None else (None, None, None)
    return _Rollout.apply(adj, int(steps), u0, v0, w0, p0, fu, fv, fw)


def rollout_kinetic_energy(adj: Adjoint, steps: int, u0, v0, w0, p0=None, forcing=None):
    """Final kinetic energy after `steps` steps (a 0-d tensor), differentiable like `rollout`."""
    fu, fv, fw = forcing if forcing is not None else (None, None, None)
    return _RolloutKE.apply(adj, int(steps), u0, v0, w0, p0, fu, fv, fw)
