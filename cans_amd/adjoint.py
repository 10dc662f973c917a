"""Gradients through RK3 steps on the GPU (FP64; each direction periodic or bounded by
walls, a periodic z on a uniform grid, walls in z on any grid; bulk forcing of u and v
along periodic directions; explicit diffusion, or 3-D implicit diffusion on tri-periodic
grids).  Under CANS_MIXED the projections (forward and adjoint) run with FP32 pressure
storage and transforms, the velocity and everything else in FP64.  Scalars (explicit
diffusion, bulk-forced or not, walled where the box is; the first one buoyant when the case says so)
are advanced by the forward steps, and their adjoints carried by the backward ones
(cans_rk3_substep_adjoint_scal); the gradient is taken with respect to the velocity.

``kinetic_energy_after_steps(sim, u0, v0, w0, dt, nsteps, every)`` runs nsteps fused
RK3 steps (cans_rk3_substep) from the interior velocity (u0, v0, w0) with zero
pressure and returns the volume-averaged kinetic energy of the final velocity as a
torch scalar whose backward pass runs the adjoint substeps
(cans_rk3_substep_adjoint).  The average weighs u and v by their cells' dzf(k) and
w by dzc(k); the normal velocities' wall faces (u at i = n1, v at j = n2, w at k = n3),
which the boundary conditions fix, are left out, and their gradient is zero.  The forward pass keeps the velocity and pressure every
`every` steps; the backward pass restarts from each checkpoint in turn (last first),
recomputes its steps while keeping the velocity at the start of every substep, and
runs their adjoints in reverse.  Recomputed steps repeat the forward ones bit for bit
(the fused substep is deterministic), so the gradient is that of the computed map.

Tensors are the fields' interiors in torch's (k, j, i) order: shape (n3, n2, n1),
float64, on the context's device.
"""
import torch

from cans_amd import _lib as LB


def _interior(f):
    return f.view[1:-1, 1:-1, 1:-1]


def _weights(sim, like):
    """Per-k volume weights of u, v, w divided by the box volume (torch (n3, 1, 1) columns)."""
    g, n = sim.grid, sim.case.n
    n3 = n[2]
    vol = float(g.dzf[1:n3 + 1].sum()) * n[0] * n[1]
    wf = torch.tensor(g.dzf[1:n3 + 1] / vol, dtype=like.dtype, device=like.device).view(n3, 1, 1)
    wc = torch.tensor(g.dzc[1:n3 + 1] / vol, dtype=like.dtype, device=like.device).view(n3, 1, 1)
    return wf, wf, wc


def _clear_wall_faces(sim, fields):
    """Zero the normal velocities' wall faces of (u, v, w) interiors in (k, j, i) order."""
    cb = sim.case.cbcpre
    if cb[0:2] != "PP":
        fields[0][:, :, -1] = 0.
    if cb[2:4] != "PP":
        fields[1][:, -1, :] = 0.
    if cb[4:6] != "PP":
        fields[2][-1, :, :] = 0.


def _addr(f, buf):
    """Address of element (0,0,0) of a copy of field f's buffer."""
    return buf.data_ptr() + buf.element_size() * f.off


class _KineticEnergyAfterSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, v0, w0, sim, dt, nsteps, every, s0):
        if sim.ctx.precision not in (LB.CANS_FP64, LB.CANS_MIXED):
            raise LB.CansError("the adjoint needs a CANS_FP64 or CANS_MIXED simulation")
        for f, a in zip((sim.u, sim.v, sim.w), (u0, v0, w0)):
            f.buf.zero_()
            _interior(f).copy_(a)
        for f in (sim.p, sim.pp):
            f.buf.zero_()
        for i, f in enumerate(sim.s):
            f.buf.zero_()
            if s0 is not None:
                _interior(f).copy_(s0[i])
        sim.init_bounds()
        fields = (sim.u, sim.v, sim.w, sim.p) + tuple(sim.s)
        ckpt = {}
        for s in range(nsteps):
            if s % every == 0:
                ckpt[s] = [f.buf.clone() for f in fields]
            sim.rk3_step(dt)
        fin = [_interior(f).clone() for f in (sim.u, sim.v, sim.w)]
        _clear_wall_faces(sim, fin)
        wts = _weights(sim, fin[0])
        E = 0.5 * sum((m * x * x).sum() for m, x in zip(wts, fin))
        ctx.sim, ctx.dt, ctx.nsteps, ctx.every, ctx.ckpt, ctx.fin = sim, dt, nsteps, every, ckpt, fin
        return E

    @staticmethod
    def backward(ctx, gE):
        sim, dt, nsteps, every = ctx.sim, ctx.dt, ctx.nsteps, ctx.every
        wts = _weights(sim, ctx.fin[0])
        # adjoint velocity and right-hand side, each with a second set the substeps write into
        ub, ubo, fb, fbo = ([sim.ctx.field() for _ in range(3)] for _ in range(4))
        pp = sim.ctx.field(kind=LB.CANS_FIELD_PRESSURE)
        for f, m, x in zip(ub, wts, ctx.fin):  # dE/du = m u (zero on the wall faces: fin is)
            _interior(f).copy_(m * x * gE)
        ns = len(sim.s)
        # the scalars' adjoints (zero at the end: the loss reads the velocity only), FP64 fields
        sb, sbo, fsb, fsbo = ([sim.ctx.field() for _ in range(ns)] for _ in range(4))
        for f in sb + fsb:
            f.buf.zero_()
        fields = (sim.u, sim.v, sim.w, sim.p) + tuple(sim.s)
        for s0 in sorted(ctx.ckpt, reverse=True):
            for f, b in zip(fields, ctx.ckpt[s0]):
                f.buf.copy_(b)
            steps = []
            for _ in range(s0, min(s0 + every, nsteps)):
                starts = []
                for irk in (1, 2, 3):
                    starts.append([f.buf.clone() for f in (sim.u, sim.v, sim.w) + tuple(sim.s)])
                    sim.rk3_substep(irk, dt)
                steps.append(starts)
            for starts in reversed(steps):
                for irk in (3, 2, 1):
                    st = starts[irk - 1]
                    q = [_addr(f, b) for f, b in zip((sim.u, sim.v, sim.w), st[:3])]
                    scal = None
                    if ns:
                        scal = ([_addr(f, b) for f, b in zip(sim.s, st[3:])], sb, fsb, sbo, fsbo)
                    sim.rk3_substep_adjoint(irk, dt, q, ub, fb, ubo, fbo, pp, scal)
                    ub, ubo, fb, fbo = ubo, ub, fbo, fb
                    sb, sbo, fsb, fsbo = sbo, sb, fsbo, fsb
            del steps
        grads = [_interior(f).clone() for f in ub]
        _clear_wall_faces(sim, grads)
        return grads[0], grads[1], grads[2], None, None, None, None, None


def kinetic_energy_after_steps(sim, u0, v0, w0, dt, nsteps, every=4, s0=None):
    """Volume-averaged kinetic energy after nsteps RK3 steps from (u0, v0, w0) (and the scalars s0, a
    sequence of tensors, zero if None); differentiable in the velocity."""
    return _KineticEnergyAfterSteps.apply(u0, v0, w0, sim, dt, nsteps, every, s0)
