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


# ---- batched replicas with SRD solvent -------------------------------------
class ReplicaSrdRun:
    """`steps` batched steps of R replicas with SRD solvent as a
    differentiable function of every replica's drive f_a:

        run = ReplicaSrdRun(reps, solvent, dt, gap, shifts, steps, checkpoint_every)
        fa = torch.tensor([...], dtype=torch.float64, requires_grad=True)   # (R,)
        pos, vel = run(fa, pos0, vel0)       # all beads, then every replica's solvent
        P = polar_speed(pos, vel, run)       # (R,) polar speed of every replica
        ((P - 0.5) ** 2).sum().backward()    # fa.grad = d loss / d f_a

    reps: NWReplicaParams[R] (no Langevin thermostat), solvent:
    NWReplicaSolvent[R], shifts: (collisions, R, 3) grid shifts.  The forward
    pass runs nwBatchAdjointForward, the backward pass nwBatchAdjointBackward
    (collisions' cells and signs held fixed); the loss may depend on the
    beads' final positions and velocities, not on the solvent's.  The other
    parameters' gradients of the last backward pass: `grad_theta` (R x 4:
    f_a, kappa, epsilon, gamma)."""

    def __init__(self, reps, solvent, dt, gap, shifts, steps, checkpoint_every=50):
        self.reps, self.solvent = reps, solvent
        self.R = len(reps)
        self.dt, self.gap, self.steps, self.K = dt, gap, steps, checkpoint_every
        self.shifts = torch.as_tensor(shifts, dtype=torch.float32).reshape(-1).contiguous()
        self.bead_off = None
        self.grad_theta = None

    def set_batch(self, fa):
        lib = abi.load()
        R = self.R
        reps = (abi.NWReplicaParams * R)()
        for r in range(R):
            reps[r] = self.reps[r]
            reps[r].ext.driveForce = float(fa[r])
        lib.nwBatchSetReplicas(reps, R)
        sol = (abi.NWReplicaSolvent * R)(*self.solvent)
        lib.nwBatchSetSolvent(sol, R)
        off = (ctypes.c_uint * (R + 1))()
        lib.nwBatchLayout(off, None)
        self.bead_off = list(off)
        return lib

    def __call__(self, fa, pos0, vel0):
        return _ReplicaSrdFn.apply(fa, pos0, vel0, self)


class _ReplicaSrdFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fa, pos0, vel0, run):
        lib = run.set_batch(fa.detach().cpu())
        N = run.bead_off[-1]
        pos = pos0.detach().clone().contiguous()
        vel = vel0.detach().clone().contiguous()
        zeros = [torch.zeros((N, 4), dtype=torch.float32, device=pos.device) for _ in range(3)]
        sh = run.shifts
        with _Stream(lib):
            lib.nwBatchAdjointForward(_ptr(pos), _ptr(vel), _ptr(zeros[0]), _ptr(zeros[1]), _ptr(zeros[2]), run.dt,
                                      run.gap, ctypes.cast(sh.data_ptr(), ctypes.POINTER(ctypes.c_float)),
                                      run.steps, run.K)
        _last_run[0] += 1
        ctx.run_id = _last_run[0]
        ctx.run = run
        return pos, vel

    @staticmethod
    def backward(ctx, grad_pos, grad_vel):
        if ctx.run_id != _last_run[0]:
            raise RuntimeError("ReplicaSrdRun: only the last forward run can be differentiated")
        run = ctx.run
        N = run.bead_off[-1]
        for g in (grad_pos, grad_vel):
            if g is not None and bool(torch.any(g[N:] != 0)):
                raise RuntimeError("ReplicaSrdRun: the loss may not depend on the solvent's final state")
        lib = abi.load()
        gp = grad_pos[:N].detach().to(torch.float32).contiguous() if grad_pos is not None else None
        gv = grad_vel[:N].detach().to(torch.float32).contiguous() if grad_vel is not None else None
        if gp is None:
            gp = torch.zeros((N, 4), dtype=torch.float32, device=grad_vel.device)
        g = (ctypes.c_double * (4 * run.R))()
        with _Stream(lib):
            lib.nwBatchAdjointBackward(_ptr(gp), _ptr(gv) if gv is not None else None, g)
        run.grad_theta = torch.tensor(list(g), dtype=torch.float64).reshape(run.R, 4)
        return run.grad_theta[:, 0].clone(), None, None, None


class _PolarSpeedFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, vel, run):
        lib = abi.load()
        N = run.bead_off[-1]
        pos, vel = pos.detach().contiguous(), vel.detach().contiguous()
        gP = torch.zeros_like(pos)
        gV = torch.zeros_like(vel)
        P = (ctypes.c_double * run.R)()
        with _Stream(lib):
            lib.nwBatchPolarSpeedGrad(_ptr(pos), _ptr(vel), _ptr(gP), _ptr(gV), P)
        counts = torch.tensor([run.bead_off[r + 1] - run.bead_off[r] for r in range(run.R)], device=pos.device)
        ctx.save_for_backward(gP[:N], gV[:N], counts)
        ctx.NT = pos.shape[0]
        return torch.tensor(list(P), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gP, gV, counts = ctx.saved_tensors
        scale = torch.repeat_interleave(g.to(gP.device, torch.float32), counts)[:, None]
        N = gP.shape[0]
        pad = torch.zeros((ctx.NT - N, 4), dtype=torch.float32, device=gP.device)
        return torch.cat([scale * gP, pad]), torch.cat([scale * gV, pad]), None


def polar_speed(pos, vel, run):
    """Polar speed <v . t> of every replica of `run` (differentiable in the
    beads' positions and velocities)."""
    return _PolarSpeedFn.apply(pos, vel, run)
