"""FP64 reference of the fused MD step and of its gradient -- TEST
INFRASTRUCTURE ONLY (the checker of nwAdjointForward / nwAdjointBackward).

The step is the one ParticleSystem::update runs for dry worms
(simulation/particleSystem.cpp:257-330, particles_kernel_impl.cuh):
velocity Verlet (filament_integrator :71-142), Langevin (langevinKernel
:859-874) with the Philox normals of the library's generator, bending
(filamentKernel :751-818), harmonic chain bonds and the drive along the bond
(the archived AVS/NWWormsKernel_InternalForce.h:37-59,
NWWormsKernel_DriveForce.h:10-40) and WCA between beads of different worms
(AVS/NWDeviceFunctions.h:73-81), here over all bead pairs (small systems).

Derivatives are not written out: the force Jacobian is assembled column by
column by complex-step differentiation (F(r + i h e_c).imag / h, exact to
rounding for these analytic expressions), and so are the parameter
derivatives; the reverse sweep then applies the chain rule of the step,
    v' = v + (F + G) dt / 2, r' = r + v' dt + F dt^2 / 2, G' = F, F' = Phi(r', v').
Pinned by central finite differences of the forward map (tests)."""
import numpy as np

import orc

H_CS = 1e-30  # complex step


class Model:
    """W worms of n beads in a periodic box; theta = (f_a, kappa, epsilon, gamma)."""

    def __init__(self, p, e, W, dt, kT_noise, langevin, seed, offset):
        self.n = p.filamentSize
        self.W = W
        self.N = W * self.n
        self.L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
        self.k, self.r0 = float(p.bondSpringK), float(p.bondSpringL)
        self.sigma, self.rc = float(e.wcaSigma), float(e.wcaCutoff)
        self.dt, self.kT = float(dt), float(kT_noise)
        self.langevin = bool(langevin)
        self.seed, self.offset = seed, offset
        self.worm = np.repeat(np.arange(W), self.n)

    def mi(self, d):
        return d - self.L * np.round(d.real / self.L)

    def forces(self, r, v, z, theta):
        fa, kappa, eps, gamma = theta
        W, n = self.W, self.n
        cplx = np.iscomplexobj(r) or np.iscomplexobj(v) or any(np.iscomplexobj(t) for t in theta)
        F = np.zeros((self.N, 3), np.complex128 if cplx else np.float64)
        if self.langevin:
            F += -gamma * v + np.sqrt(2.0 * gamma * self.kT) * z
        R = r.reshape(W, n, 3)
        Fw = F.reshape(W, n, 3)
        if n >= 2:
            d = self.mi(R[:, 1:] - R[:, :-1])
            ln = np.sqrt((d * d).sum(-1))[..., None]
            fb = self.k * (ln - self.r0) / ln * d
            Fw[:, :-1] += fb + fa * d / ln
            Fw[:, 1:] -= fb
        if n >= 3:
            rij, rjk = d[:, :-1], d[:, 1:]
            dij = np.sqrt((rij * rij).sum(-1))[..., None]
            djk = np.sqrt((rjk * rjk).sum(-1))[..., None]
            dd = (rij * rjk).sum(-1)[..., None]
            A = kappa / (dij * djk)
            B = rij * dd / (rij * rij).sum(-1)[..., None]
            C = rjk * dd / (rjk * rjk).sum(-1)[..., None]
            fi = -A * (rjk - B)
            fk = -A * (C - rij)
            Fw[:, :-2] += fi
            Fw[:, 1:-1] += -fi - fk
            Fw[:, 2:] += fk
        # WCA over all pairs of beads of different worms
        i, j = np.triu_indices(self.N, 1)
        keep = self.worm[i] != self.worm[j]
        i, j = i[keep], j[keep]
        d = self.mi(r[j] - r[i])
        rr = (d * d).sum(-1)
        close = rr.real < self.rc * self.rc
        i, j, d, rr = i[close], j[close], d[close], rr[close]
        u = 1.0 / rr
        s6 = (self.sigma * self.sigma * u) ** 3
        f = (24.0 * eps * u * s6 * (2.0 * s6 - 1.0))[:, None]
        np.add.at(F, i, -f * d)
        np.add.at(F, j, f * d)
        return F

    def noise(self, off):
        return orc.random_gaussian(self.seed, off, 4 * self.N).reshape(self.N, 4)[:, :3]

    def run(self, r, v, F, G, theta, steps):
        """Forward; returns the final state and the per-step (r', v', z, offset)."""
        r, v, F, G = (np.array(x, np.float64)[:, :3] for x in (r, v, F, G))
        off = self.offset
        traj = []
        for _ in range(steps):
            v = v + 0.5 * (F + G) * self.dt
            r = r + v * self.dt + 0.5 * F * self.dt * self.dt
            r = r - self.L * np.floor((r + self.L / 2) / self.L)  # periodic wrap (unit derivative)
            z = self.noise(off) if self.langevin else np.zeros_like(r)
            G, F = F, self.forces(r, v, z, theta)
            traj.append((r, v, z))
            if self.langevin:
                off += self.N
        return r, v, F, G, traj

    def jacobian_r(self, r, v, z, theta):
        N3 = 3 * self.N
        J = np.empty((N3, N3))
        for c in range(N3):
            dr = np.zeros(N3, np.complex128)
            dr[c] = 1j * H_CS
            J[:, c] = (self.forces(r + dr.reshape(-1, 3), v, z, theta).imag / H_CS).reshape(-1)
        return J

    def dforce_dtheta(self, r, v, z, theta):
        out = []
        for m in range(4):
            t = [complex(x) for x in theta]
            t[m] += 1j * H_CS
            out.append((self.forces(r, v, z, t).imag / H_CS).reshape(-1))
        return out

    def gradient(self, r, v, F, G, theta, steps, dLdr, dLdv=None):
        """d loss / d theta for loss adjoints dLdr (and dLdv) of the final state."""
        *_, traj = self.run(r, v, F, G, theta, steps)
        lr = np.array(dLdr, np.float64)[:, :3].reshape(-1)
        lv = np.zeros_like(lr) if dLdv is None else np.array(dLdv, np.float64)[:, :3].reshape(-1)
        lF = np.zeros_like(lr)
        lG = np.zeros_like(lr)
        gamma = theta[3]
        grad = np.zeros(4)
        dt = self.dt
        for (rk, vk, zk) in reversed(traj):
            J = self.jacobian_r(rk, vk, zk, theta)
            for m, dF in enumerate(self.dforce_dtheta(rk, vk, zk, theta)):
                grad[m] += lF @ dF
            Rr = lr + J.T @ lF
            dFdv = -gamma if self.langevin else 0.0
            V = lv + dFdv * lF + dt * Rr
            lr, lv, lF, lG = Rr, V, lG + 0.5 * dt * dt * Rr + 0.5 * dt * V, 0.5 * dt * V
        return grad

    def loss_linear(self, r, v, F, G, theta, steps, w):
        rf, *_ = self.run(r, v, F, G, theta, steps)
        return float((np.asarray(w, np.float64)[:, :3] * rf).sum())
