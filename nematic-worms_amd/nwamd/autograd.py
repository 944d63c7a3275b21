"""torch.autograd binding of the adjoint of the fused MD step.

    theta = torch.tensor([f_a, kappa, epsilon, gamma], dtype=torch.float64, requires_grad=True)
    run = WormRun(params, ext, num_filaments, dt, kT_noise, seed, offset, steps, checkpoint_every)
    pos = run(theta, pos0, vel0)            # final positions, (N, 4) float32 on the GPU
    loss = bond_order(pos, run)             # or any torch expression of pos
    loss.backward()                         # theta.grad = d loss / d theta

The forward pass runs nwAdjointForward (the fused steps, checkpointed); the
backward pass hands the adjoint of the final positions to nwAdjointBackward.
torch provides the device memory and the stream; the library runs on that
stream while it works.  One run is differentiable at a time: the library
keeps the checkpoints of the last forward pass only.
"""
import copy
import ctypes

import torch

from . import abi

_last_run = [0]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class _Stream:
    """Puts the library on torch's current stream for the duration."""

    def __init__(self, lib):
        self.lib = lib

    def __enter__(self):
        self.prev = self.lib.nwGetStream()
        self.lib.nwSetStream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def __exit__(self, *exc):
        self.lib.nwSetStream(self.prev)


class WormRun:
    """`steps` fused MD steps of dry worms as a differentiable function of
    theta = (f_a, kappa, epsilon, gamma) (SimParams / NWExtParams from
    `params` / `ext`, the parameters replaced by theta)."""

    def __init__(self, params, ext, num_filaments, dt, kT_noise, seed, offset, steps, checkpoint_every=100,
                 langevin=True):
        self.params, self.ext = copy.copy(params), copy.copy(ext)
        self.W, self.dt, self.kT = num_filaments, dt, kT_noise
        self.seed, self.offset = seed, offset
        self.steps, self.K, self.langevin = steps, checkpoint_every, langevin

    def with_theta(self, theta):
        p, e = copy.copy(self.params), copy.copy(self.ext)
        fa, kappa, eps, gamma = (float(x) for x in theta)
        e.driveForce, p.bondBendingK, e.wcaEpsilon, p.gamma = fa, kappa, eps, gamma
        return p, e

    def __call__(self, theta, pos0, vel0):
        return _WormRunFn.apply(theta, pos0, vel0, self)


class _WormRunFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, pos0, vel0, run):
        lib = abi.load()
        p, e = run.with_theta(theta.detach().cpu())
        pos = pos0.detach().clone().contiguous()
        vel = vel0.detach().clone().contiguous()
        zeros = [torch.zeros_like(pos) for _ in range(3)]
        lib.setParameters(ctypes.byref(p))
        lib.nwSetExtParameters(ctypes.byref(e))
        lib.nwRngSetSeed(run.seed, run.offset)
        with _Stream(lib):
            lib.nwAdjointForward(_ptr(pos), _ptr(vel), _ptr(zeros[0]), _ptr(zeros[1]), _ptr(zeros[2]), run.W, run.dt,
                                 p.gamma, run.kT, 1 if run.langevin else 0, run.steps, run.K)
        _last_run[0] += 1
        ctx.run_id = _last_run[0]
        ctx.vel_final = vel
        return pos

    @staticmethod
    def backward(ctx, grad_pos):
        if ctx.run_id != _last_run[0]:
            raise RuntimeError("WormRun: only the last forward run can be differentiated")
        lib = abi.load()
        gp = grad_pos.detach().to(torch.float32).contiguous()
        g = (ctypes.c_double * 4)()
        with _Stream(lib):
            lib.nwAdjointBackward(_ptr(gp), None, g)
        return torch.tensor(list(g), dtype=torch.float64), None, None, None


class _BondOrderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, run):
        lib = abi.load()
        p, _ = run.with_theta([run.ext.driveForce, run.params.bondBendingK, run.ext.wcaEpsilon, run.params.gamma])
        lib.setParameters(ctypes.byref(p))
        pos = pos.detach().contiguous()
        grad = torch.empty_like(pos)
        with _Stream(lib):
            S = lib.nwBondOrderGrad(_ptr(pos), run.W, _ptr(grad))
        ctx.save_for_backward(grad)
        return torch.tensor(S, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (g.to(grad.device, torch.float32) * grad), None


def bond_order(pos, run):
    """Bond-tangent nematic order S of the worms at `pos` (differentiable)."""
    return _BondOrderFn.apply(pos, run)
