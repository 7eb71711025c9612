"""torch.autograd harness of the differentiable rollout.

    xT, vT = rollout(engine, x0, v0, gravity, dt, steps, ckpt_every)
    rho = density(engine, xT)

`rollout` runs `steps` PBF steps on the engine's HIP kernels (mtlp_rollout,
keeping (x, v) every ckpt_every steps); its backward is mtlp_adjoint_rollout,
which recomputes each segment from its checkpoint with a tape and walks it
backwards, giving dL/dx0, dL/dv0 and dL/dgravity.  `density` is rho_i =
W(0) + sum_j W(r_ij) at given positions (fresh spatial hash) with
mtlp_density_vjp as its backward.

Tensors are float32 CUDA (HIP) tensors of shape (n, 4), original particle
order; the w of gradients is 0.  Work runs on torch's current stream.  The
engine holds the checkpoints of one rollout: call backward on a rollout
before starting the next one on the same engine.
"""
import numpy as np
import torch


def _stream():
    return torch.cuda.current_stream()


def _check_state(t, n, name):
    if t.dtype != torch.float32 or t.dim() != 2 or t.shape != (n, 4) or not t.is_cuda:
        raise ValueError(f"{name} must be a float32 CUDA tensor of shape ({n}, 4)")
    return t.contiguous()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, gravity, engine, dt, steps, ckpt_every):
        n = engine.particle_count
        x0 = _check_state(x0, n, "x0")
        v0 = _check_state(v0, n, "v0")
        g = [float(c) for c in gravity.detach().to("cpu", torch.float64).reshape(3)]
        p = engine.params
        if [float(c) for c in p.gravity] != [float(np.float32(c)) for c in g]:
            for k in range(3):
                p.gravity[k] = g[k]
            engine.params = p
        s = _stream()
        engine.set_state_device(x0, v0, s)
        engine.rollout(dt, steps, ckpt_every, s)
        xT = torch.empty_like(x0)
        vT = torch.empty_like(v0)
        engine.get_state_device(xT, vT, s)
        engine._rollout_serial = getattr(engine, "_rollout_serial", 0) + 1
        ctx.engine = engine
        ctx.serial = engine._rollout_serial
        ctx.args = (dt, steps, ckpt_every)
        ctx.gshape, ctx.gdtype, ctx.gdevice = gravity.shape, gravity.dtype, gravity.device
        return xT, vT

    @staticmethod
    def backward(ctx, gx, gv):
        engine = ctx.engine
        if getattr(engine, "_rollout_serial", None) != ctx.serial:
            raise RuntimeError("the engine ran another rollout since this one; its checkpoints are gone")
        dt, steps, every = ctx.args
        n = engine.particle_count
        s = _stream()
        dev = gx.device if gx is not None else gv.device
        xb0 = torch.empty((n, 4), dtype=torch.float32, device=dev)
        vb0 = torch.empty((n, 4), dtype=torch.float32, device=dev)
        gb = torch.empty(3, dtype=torch.float32, device=dev)
        engine.adjoint_rollout(dt, steps, every,
                               None if gx is None else gx.contiguous(),
                               None if gv is None else gv.contiguous(),
                               xb0, vb0, gb, s)
        gbar = gb.to(ctx.gdevice, ctx.gdtype).reshape(ctx.gshape)
        return xb0, vb0, gbar, None, None, None, None


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, engine):
        n = engine.particle_count
        x = _check_state(x, n, "x")
        rho = torch.empty(n, dtype=torch.float32, device=x.device)
        engine.density(x, rho, _stream())
        ctx.engine = engine
        ctx.save_for_backward(x)
        return rho

    @staticmethod
    def backward(ctx, grho):
        (x,) = ctx.saved_tensors
        xb = torch.empty_like(x)
        ctx.engine.density_vjp(x, grho.contiguous().to(torch.float32), xb, _stream())
        return xb, None


def rollout(engine, x0, v0, gravity, dt, steps, ckpt_every):
    """(x_T, v_T) after `steps` steps from (x0, v0) under `gravity` (3)."""
    return _Rollout.apply(x0, v0, gravity, engine, float(dt), int(steps), int(ckpt_every))


def density(engine, x):
    """rho at positions x (n, 4), original order."""
    return _Density.apply(x, engine)


def centroid_density_loss(engine, xT, target, w_density=0.1):
    """|centroid(x_T) - target|^2 + w mean((rho(x_T) / rho0 - 1)^2)."""
    c = xT[:, :3].mean(dim=0) - target
    rho = density(engine, xT)
    dev = rho / engine.params.rest_density - 1.0
    return (c * c).sum() + w_density * (dev * dev).mean()


__all__ = ["rollout", "density", "centroid_density_loss"]
