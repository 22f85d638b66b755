"""PyTorch autograd wrappers around libncg's FP64 gravity and RK4 adjoints.

    acc = GravityField.apply(x, m, radius, central, G, softening, ctx)
    xT, vT = RK4Trajectory.apply(x0, v0, m, radius, central0, G, softening, dt,
                                 nsteps, checkpoint_every, ctx)

x, v: (n, 3) float64 CUDA tensors; m, radius: (n,); central: (8,) =
{x, y, z, vx, vy, vz, M, present}.  Forward passes run the FP64-fast form on
the GPU (RK4Trajectory keeps a checkpoint every checkpoint_every steps,
ncg_rk4_forward_ckpt_dev); backward passes call the hand-written VJP kernels
(ncg_vjp_gravity_dev) and the discrete RK4 adjoint from those checkpoints
(ncg_rk4_adjoint_ckpt_dev).  Gradients flow to x/x0, v0, m and the central body;
radii (softening lengths) are not differentiated.  Collision-free dynamics
only: contact events are not differentiable.
"""
import torch

from ._abi import Context

_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def pack_bodies(x, v, m, radius):
    n = x.shape[0]
    b = torch.empty((n, 8), dtype=torch.float64, device=x.device)
    b[:, 0:3] = x
    b[:, 3:6] = v if v is not None else 0.0
    b[:, 6] = m
    b[:, 7] = radius
    return b.contiguous()


def _sync(ctx):
    torch.cuda.current_stream().synchronize()
    ctx.synchronize()


class GravityField(torch.autograd.Function):
    """Accelerations a_i (mutual + central) and the central body's CM_acc."""

    @staticmethod
    def forward(fctx, x, m, radius, central, G=1.0, softening=0.0, ctx=None):
        ctx = ctx or default_context()
        b = pack_bodies(x.detach(), None, m.detach(), radius.detach())
        cen = central.detach().contiguous()
        n = x.shape[0]
        acc = torch.zeros((n, 3), dtype=torch.float64, device=x.device)
        cm = torch.zeros(3, dtype=torch.float64, device=x.device)
        present = bool(cen[7].item() != 0.0)
        c4 = tuple(float(t) for t in (cen[0], cen[1], cen[2], cen[6])) if present else None
        _sync(ctx)
        ctx.gravity_allpairs_dev(b, acc, None, c4, G, softening, "fast64", cm if present else None)
        ctx.synchronize()
        fctx.save_for_backward(b, cen)
        fctx.G, fctx.soft, fctx.ncg = G, softening, ctx
        return acc, cm

    @staticmethod
    def backward(fctx, g_acc, g_cm):
        b, cen = fctx.saved_tensors
        ctx = fctx.ncg
        n = b.shape[0]
        dev = b.device
        lam = (g_acc if g_acc is not None else torch.zeros((n, 3), dtype=torch.float64, device=dev)).contiguous()
        lamc = (g_cm if g_cm is not None else torch.zeros(3, dtype=torch.float64, device=dev)).contiguous()
        gx = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        gm = torch.zeros(n, dtype=torch.float64, device=dev)
        gc = torch.zeros(4, dtype=torch.float64, device=dev)
        _sync(ctx)
        ctx.vjp_gravity_dev(b, cen, lam, gx, gm, gc, lamc, fctx.G, fctx.soft)
        ctx.synchronize()
        gcen = torch.zeros(8, dtype=torch.float64, device=dev)
        gcen[0:3] = gc[0:3]
        gcen[6] = gc[3]
        return gx, gm, None, gcen, None, None, None


class RK4Trajectory(torch.autograd.Function):
    """Final (x, v) after nsteps FP64 RK4 steps; central body moves too."""

    @staticmethod
    def forward(fctx, x0, v0, m, radius, central0, G=1.0, softening=0.0, dt=0.01, nsteps=1,
                checkpoint_every=20, ctx=None):
        ctx = ctx or default_context()
        b = pack_bodies(x0.detach(), v0.detach(), m.detach(), radius.detach())
        c = central0.detach().contiguous().clone()
        _sync(ctx)
        # The forward run keeps a checkpoint every `checkpoint_every` steps.
        ck, ckc = ctx.rk4_forward_ckpt_dev(b, c, nsteps, dt, checkpoint_every, G, softening)
        ctx.synchronize()
        fctx.save_for_backward(ck, ckc)
        fctx.args = (G, softening, dt, nsteps, checkpoint_every, ctx)
        return b[:, 0:3].clone(), b[:, 3:6].clone(), c.clone()

    @staticmethod
    def backward(fctx, g_x, g_v, g_c):
        ck, ckc = fctx.saved_tensors
        G, soft, dt, nsteps, K, ctx = fctx.args
        n = ck.shape[1]
        dev = ck.device
        gfin = torch.zeros((n, 6), dtype=torch.float64, device=dev)
        if g_x is not None:
            gfin[:, 0:3] = g_x
        if g_v is not None:
            gfin[:, 3:6] = g_v
        gcf = torch.zeros(6, dtype=torch.float64, device=dev)
        if g_c is not None:
            gcf[:] = g_c[0:6]
        g0 = torch.zeros((n, 6), dtype=torch.float64, device=dev)
        gm = torch.zeros(n, dtype=torch.float64, device=dev)
        gc0 = torch.zeros(7, dtype=torch.float64, device=dev)
        _sync(ctx)
        ctx.rk4_adjoint_ckpt_dev(ck, ckc, gfin.contiguous(), g0, gm, gc0, nsteps, dt, K, G, soft,
                                 gcf.contiguous())
        ctx.synchronize()
        gcen = torch.zeros(8, dtype=torch.float64, device=dev)
        gcen[0:6] = gc0[0:6]
        gcen[6] = gc0[6]
        return g0[:, 0:3], g0[:, 3:6], gm, None, gcen, None, None, None, None, None, None
