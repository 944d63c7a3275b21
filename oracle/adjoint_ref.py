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

With an Srd run, an SRD collision follows the MD part of steps 0, gap,
2 gap, ... (particleSystem.cpp:249-339, 341-413): the particles' cells from
the FP32 SRD hash of the oracle (calcHashD :312-342 with the collision's grid
shift), the rotation signs from the library's uniform draw, the cell momentum
(cellCenterOfMomentumSRD :559-594), the rotation about it (collideSolventKernel
:651-691, rotate2D :640-648) and the solvent's streaming (solvent_integrator
:145-205).  The backward holds the cells and signs fixed, so a collision is a
linear map of the velocities; its Jacobian is the map applied to the identity.

Derivatives are not written out: the force Jacobian is assembled column by
column by complex-step differentiation (F(r + i h e_c).imag / h, exact to
rounding for these analytic expressions), and so are the parameter
derivatives; the reverse sweep then applies the chain rule of the step,
    v' = v + (F + G) dt / 2, r' = r + v' dt + F dt^2 / 2, G' = F, F' = Phi(r', v'),
    v'' = C v' (a collision step).
Pinned by central finite differences of the forward map (tests)."""
import ctypes

import numpy as np

import orc

H_CS = 1e-30  # complex step


class Srd:
    """SRD collisions of a run: Ns solvent particles after the beads, a
    collision after steps 0, gap, 2 gap, ... with grid shift shifts[c] (3
    floats) for collision c; s the SolventParams (grid, cell, alpha), the beads
    rotated when rot_beads (else they only join the cell momentum)."""

    def __init__(self, s, Ns, gap, shifts, rot_beads):
        self.s, self.Ns, self.gap = s, Ns, gap
        self.shifts = np.asarray(shifts, np.float32).reshape(-1, 3)
        self.rot_beads = bool(rot_beads)
        self.Cs = int(s.numCells)
        self.alpha = float(np.float32(s.alpha))


class Collision:
    """One collision with its cells h (per particle), signs (per cell) and
    masses m fixed: the linear map of the (NT, 3) velocities."""

    def __init__(self, h, sign, m, rotated, alpha, Cs):
        self.h, self.sign, self.m, self.rotated = h, sign, m, rotated
        self.c, self.s = np.cos(alpha), np.sin(alpha)
        self.M = np.bincount(h, weights=m, minlength=Cs)
        self.Cs = Cs

    def apply(self, V):
        """V: (NT, 3, B) -> the collided velocities."""
        h = self.h
        P = np.zeros((self.Cs,) + V.shape[1:])
        np.add.at(P, h, self.m[:, None, None] * V)
        u = P[h] / self.M[h][:, None, None]
        d = V - u
        sn = (self.sign[h] * self.s)[:, None]
        rot = np.stack([d[:, 0] * self.c - d[:, 1] * sn, d[:, 0] * sn + d[:, 1] * self.c, d[:, 2]], axis=1)
        return np.where(self.rotated[:, None, None], u + rot, V)

    def jacobian(self):
        NT = self.h.shape[0]
        E = np.eye(3 * NT).reshape(NT, 3, 3 * NT)
        return self.apply(E).reshape(3 * NT, 3 * NT)


class Model:
    """W worms of n beads in a periodic box; theta = (f_a, kappa, epsilon, gamma)."""

    def __init__(self, p, e, W, dt, kT_noise, langevin, seed, offset, srd=None):
        self.p = p
        self.srd = srd
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
        # WCA between beads of different worms: all pairs under a complex
        # step, else the pairs within the cutoff from a periodic k-d tree
        # (the same pairs and terms, summed in another order)
        i, j = self.pairs(r) if not cplx else np.triu_indices(self.N, 1)
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

    def pairs(self, r):
        """index pairs i < j closer than the cutoff under the minimum image,
        with some margin (the caller applies the exact cutoff)"""
        from scipy.spatial import cKDTree
        x = np.mod(np.asarray(r, np.float64) + self.L / 2, self.L)
        x = np.minimum(x, np.nextafter(self.L, 0))  # [0, L) for the tree
        ij = cKDTree(x, boxsize=self.L).query_pairs(self.rc * (1 + 1e-9), output_type="ndarray")
        if len(ij) == 0:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        a, b = ij[:, 0].astype(np.int64), ij[:, 1].astype(np.int64)
        return np.minimum(a, b), np.maximum(a, b)

    def noise(self, off):
        return orc.random_gaussian(self.seed, off, 4 * self.N).reshape(self.N, 4)[:, :3]

    def wrap(self, r):
        return r - self.L * np.floor((r + self.L / 2) / self.L)  # periodic wrap (unit derivative)

    def collide(self, k, r, v, rs, vs, w, off):
        """Collision c = k / gap of the full state; returns the new (v, vs),
        the Collision and the generator offset after its draw."""
        sr = self.srd
        s = type(sr.s).from_buffer_copy(sr.s)
        s.randOffset.x, s.randOffset.y, s.randOffset.z = (float(x) for x in sr.shifts[k // sr.gap])
        pos = np.zeros((self.N + sr.Ns, 4), np.float32)
        pos[:, :3] = np.concatenate([r, rs])
        h, _ = orc.calc_hash(self.p, s, pos, srd=True)
        uni = orc.random_uniform(self.seed, off, sr.Cs)
        sign = np.where(uni > 0.5, 1.0, -1.0)
        rotated = np.ones(self.N + sr.Ns, bool)
        rotated[:self.N] = sr.rot_beads
        col = Collision(h.astype(np.int64), sign, w, rotated, sr.alpha, sr.Cs)
        V = col.apply(np.concatenate([v, vs])[:, :, None])[:, :, 0]
        return V[:self.N], V[self.N:], col, off + (sr.Cs + 3) // 4

    def run(self, r, v, F, G, theta, steps):
        """Forward; returns the final state and the per-step (r', v', z,
        collision).  With an Srd run, r and v hold the beads then the solvent
        (w of v: the masses), and the returned r, v do too."""
        r, v = np.array(r, np.float64), np.array(v, np.float64)
        w = v[:, 3].copy()
        N = self.N
        rs, vs = r[N:, :3], v[N:, :3]
        r, v = r[:N, :3], v[:N, :3]
        F, G = (np.array(x, np.float64)[:N, :3] for x in (F, G))
        off = self.offset
        traj = []
        for k in range(steps):
            v = v + 0.5 * (F + G) * self.dt
            r = self.wrap(r + v * self.dt + 0.5 * F * self.dt * self.dt)
            z = self.noise(off) if self.langevin else np.zeros_like(r)
            G, F = F, self.forces(r, v, z, theta)
            col = None
            if self.langevin:
                off += self.N
            if self.srd is not None and k % self.srd.gap == 0:
                vpre = v
                v, vs, col, off = self.collide(k, r, v, rs, vs, w, off)
                rs = self.wrap(rs + vs * (self.dt * self.srd.gap))
                traj.append((r, vpre, z, col))
            else:
                traj.append((r, v, z, col))
        if self.srd is not None:
            return np.concatenate([r, rs]), np.concatenate([v, vs]), F, G, traj
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
        N3 = 3 * self.N
        lr = np.array(dLdr, np.float64)[:self.N, :3].reshape(-1)
        lv = np.zeros_like(lr) if dLdv is None else np.array(dLdv, np.float64)[:self.N, :3].reshape(-1)
        lvs = np.zeros(3 * self.srd.Ns) if self.srd is not None else None  # the solvent's velocity adjoints
        lF = np.zeros_like(lr)
        lG = np.zeros_like(lr)
        gamma = theta[3]
        grad = np.zeros(4)
        dt = self.dt
        for (rk, vk, zk, col) in reversed(traj):
            if col is not None:
                lall = col.jacobian().T @ np.concatenate([lv, lvs])
                lv, lvs = lall[:N3], lall[N3:]
            J = self.jacobian_r(rk, vk, zk, theta)
            for m, dF in enumerate(self.dforce_dtheta(rk, vk, zk, theta)):
                grad[m] += lF @ dF
            Rr = lr + J.T @ lF
            dFdv = -gamma if self.langevin else 0.0
            V = lv + dFdv * lF + dt * Rr
            lr, lv, lF, lG = Rr, V, lG + 0.5 * dt * dt * Rr + 0.5 * dt * V, 0.5 * dt * V
        return grad

    def loss_linear(self, r, v, F, G, theta, steps, w, wv=None):
        """sum of w . r (and wv . v) over the beads of the final state"""
        rf, vf, *_ = self.run(r, v, F, G, theta, steps)
        out = float((np.asarray(w, np.float64)[:self.N, :3] * rf[:self.N]).sum())
        if wv is not None:
            out += float((np.asarray(wv, np.float64)[:self.N, :3] * vf[:self.N]).sum())
        return out
