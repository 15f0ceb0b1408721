"""Gradients through the KDK leapfrog with the engine's force adjoint.

`trajectory(x0, v0, m, dt, steps, ...)` returns the final positions and
velocities as torch tensors; gradients of any loss on them flow back to the
initial positions, velocities and masses through `LeapfrogAdjoint`, a
torch.autograd.Function whose backward pass

  * walks the trajectory backwards with the time-reversed leapfrog
    (gc_step_reverse), restarting from stored checkpoints every
    `checkpoint_every` steps so the reversal error stays bounded, and
  * applies the discrete adjoint of each KDK step
        v_half = v_k + h a(x_k),  x_{k+1} = x_k + dt v_half,
        v_{k+1} = v_half + h a(x_{k+1}),                  h = dt / 2
    with one force vector-Jacobian product per step (gc_force_vjp, a
    hand-written all-pairs kernel): the two half kicks that read a(x_k)
    share one VJP with cotangent dt * vbar_half (h at the two ends).

Tensors are (n, 3) float64 on the engine's device; the engine runs on
torch's current stream.  Everything between torch and the engine is
device-to-device.
"""
from __future__ import annotations

import torch

from .engine import Engine


def _soa(t: torch.Tensor) -> torch.Tensor:
    """(n, 3) -> contiguous (3, n) float64."""
    return t.detach().to(torch.float64).t().contiguous()


def _aos(t: torch.Tensor) -> torch.Tensor:
    return t.t().contiguous()


class LeapfrogAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, engine: Engine, dt: float, steps: int, checkpoint_every: int):
        n = m.shape[0]
        stream = torch.cuda.current_stream(x0.device).cuda_stream
        engine.set_stream(stream)
        xs, vs, ms = _soa(x0), _soa(v0), m.detach().to(torch.float64).contiguous()
        engine.set_state_device(n, xs.data_ptr(), vs.data_ptr(), ms.data_ptr())
        engine.compute_accelerations()
        checkpoints = {0: (xs.clone(), vs.clone())}
        done = 0
        while done < steps:
            k = min(checkpoint_every, steps - done)
            engine.step(dt, k)
            done += k
            xk, vk = torch.empty_like(xs), torch.empty_like(vs)
            engine.get_state_device(xk.data_ptr(), vk.data_ptr())
            checkpoints[done] = (xk, vk)
        ctx.engine, ctx.dt, ctx.steps, ctx.n = engine, dt, steps, n
        ctx.checkpoints = checkpoints
        ctx.save_for_backward(ms)
        xT, vT = checkpoints[steps]
        return _aos(xT), _aos(vT)

    @staticmethod
    def backward(ctx, gx, gv):
        (ms,) = ctx.saved_tensors
        eng, dt, steps, n = ctx.engine, ctx.dt, ctx.steps, ctx.n
        h = 0.5 * dt
        dev = ms.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        eng.set_stream(stream)
        xbar = _soa(gx) if gx is not None else torch.zeros(3, n, dtype=torch.float64, device=dev)
        vbar = _soa(gv) if gv is not None else torch.zeros(3, n, dtype=torch.float64, device=dev)
        mbar = torch.zeros(n, dtype=torch.float64, device=dev)

        def restore(k):
            xk, vk = ctx.checkpoints[k]
            eng.set_state_device(n, xk.data_ptr(), vk.data_ptr(), ms.data_ptr())
            eng.compute_accelerations()

        def vjp(cot):
            lam = cot.contiguous()
            eng.force_vjp(lam.data_ptr(), xbar.data_ptr(), mbar.data_ptr(), accumulate=True)

        if steps == 0:
            return _aos(xbar), _aos(vbar), mbar, None, None, None, None
        restore(steps)
        vjp(h * vbar)                        # a(x_T) read by the last half kick
        for k in range(steps - 1, -1, -1):
            vhalf = vbar + dt * xbar         # x_{k+1} = x_k + dt v_half
            if k in ctx.checkpoints:
                restore(k)
            else:
                eng.step_reverse(dt, 1)      # (x_{k+1}, v_{k+1}) -> (x_k, v_k)
            vjp((dt if k > 0 else h) * vhalf)
            vbar = vhalf
        return _aos(xbar), _aos(vbar), mbar, None, None, None, None


def trajectory(x0, v0, m, dt, steps, engine: Engine | None = None, checkpoint_every=20,
               precision="fp64", softening=0.05, G=1.0):
    """Final (x, v) after `steps` KDK steps; differentiable in x0, v0, m."""
    if engine is None:  # kept alive by the autograd graph for the backward pass
        engine = Engine("hip-direct", precision, softening=softening, G=G)
    return LeapfrogAdjoint.apply(x0, v0, m, engine, float(dt), int(steps), int(checkpoint_every))


def reference_trajectory(x0, v0, m, dt, steps, softening=0.05, G=1.0):
    """Pure-torch FP64 KDK with all-pairs forces (autograd reference)."""
    def accel(x):
        d = x[None, :, :] - x[:, None, :]              # d_ij = x_j - x_i
        s = (d * d).sum(-1) + softening * softening
        f = s.pow(-1.5) * (1 - torch.eye(x.shape[0], dtype=x.dtype, device=x.device))
        return G * (f[:, :, None] * d * m[None, :, None]).sum(1)

    x, v = x0, v0
    a = accel(x)
    for _ in range(steps):
        v = v + (0.5 * dt) * a
        x = x + dt * v
        a = accel(x)
        v = v + (0.5 * dt) * a
    return x, v
