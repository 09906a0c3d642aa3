"""torch.autograd front end of the projection.

`project(solver, u, v, w, dt)` runs `mgpoisl` on CUDA float64 tensors in the
reference layout and returns the projected velocities and the pressure
increment.  Gradients with respect to u, v, w flow through the library's
hand-written adjoint (`mgp_adjoint_device`): the transposed divergence and
gradient stencils around an adjoint Poisson solve with the transposed SIP
factors.

The differentiated map is the projection of the interior face velocities:
before the solve the input's ghost layers are refilled from the interior
(the 2-layer exchange with corners the caller performs before mgpoisl), and
after it the library's own exchange fills the output ghosts, whose
gradients are routed back to the interior cells they copy.  Faces that are
boundary conditions (block faces without a connected neighbour, and faces
next to a blocked cell under the ghostcell operator) are treated as
constants.  The pressure increment is not differentiable: with periodic or
Neumann boundaries its mean is fixed by the solver, not by the operator.

`rollout` applies a step function repeatedly, optionally in checkpointed
segments whose forward pass is recomputed during the backward pass.
"""
from __future__ import annotations

import torch


def _exchange_pairs(solver):
    """(destination, source) positions of the 2-layer exchange with corners:
    every ghost filled by the exchange and the interior cell it copies."""
    pairs = getattr(solver, "_xpairs", None)
    if pairs is None:
        g = solver.grids
        inner = torch.from_numpy(g.interior_mask()).cuda()
        ids = torch.where(inner, torch.arange(g.n3d, dtype=torch.float64, device="cuda"),
                          torch.full((g.n3d,), -1.0, dtype=torch.float64, device="cuda"))
        solver.connect([ids], layers=2, corners=True)
        dst = torch.nonzero((ids >= 0) & ~inner).squeeze(1)
        src = ids[dst].to(torch.int64)
        pairs = (dst, src)
        solver._xpairs = pairs
    return pairs


def exchange_transpose(solver, fields):
    """Transpose of the 2-layer exchange, in place: each ghost's value is
    added to the interior cell it copies and the ghost is cleared."""
    dst, src = _exchange_pairs(solver)
    for t in fields:
        vals = t[dst]
        t[dst] = 0.0
        t.index_add_(0, src, vals)


class Projection(torch.autograd.Function):
    """(u, v, w) -> (u', v', w', dp) of one mgpoisl call."""

    @staticmethod
    def forward(ctx, u, v, w, solver, dt):
        out = [t.detach().clone().contiguous() for t in (u, v, w)]
        solver.connect(out, layers=2, corners=True)
        p = torch.zeros_like(out[0])
        solver.last_stats = solver.mgpoisl(out[0], out[1], out[2], p, dt)
        ctx.solver = solver
        ctx.dt = dt
        ctx.mark_non_differentiable(p)
        return out[0], out[1], out[2], p

    @staticmethod
    def backward(ctx, gu, gv, gw, gp):
        s = ctx.solver
        n = s.grids.n3d
        g = [torch.zeros(n, dtype=torch.float64, device="cuda") if x is None
             else x.detach().clone().contiguous() for x in (gu, gv, gw)]
        exchange_transpose(s, g)
        s.last_adjoint_stats = s.adjoint(g[0], g[1], g[2], ctx.dt)
        return g[0], g[1], g[2], None, None


def project(solver, u, v, w, dt):
    """Differentiable projection; returns (u', v', w', dp)."""
    return Projection.apply(u, v, w, solver, dt)


def rollout(step, state, nsteps, checkpoint_every=0):
    """Apply `step` (tuple of tensors -> tuple of tensors) `nsteps` times.
    With checkpoint_every = m > 0 the steps run in segments of m under
    torch.utils.checkpoint: only the segment inputs are kept, and the
    backward pass recomputes each segment before differentiating it."""
    from torch.utils.checkpoint import checkpoint

    def segment(m):
        def run(*s):
            for _ in range(m):
                s = step(*s)
            return tuple(s)
        return run

    state = tuple(state)
    done = 0
    while done < nsteps:
        m = min(checkpoint_every if checkpoint_every > 0 else nsteps, nsteps - done)
        if checkpoint_every > 0:
            state = checkpoint(segment(m), *state, use_reentrant=False)
        else:
            state = segment(m)(*state)
        done += m
    return state
