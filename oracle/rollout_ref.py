"""FP64 CPU reference of the implicit rollout over a ground plane and of its
adjoint, built on the oracle's C restatement (SNK1, ground barrier, mass,
convert, parameter adjoint) with dense numpy solves in place of MAS-PCG.

TEST INFRASTRUCTURE ONLY: the checker for the library's differentiable
rollout on small meshes.  Same time discretisation as the library's driver:
x~ = x_t + dt v_t + dt^2 g on free vertices (pinned: x_t), Newton on the
incremental potential with the ground step bound (slack 0.8) and energy
backtracking, v_{t+1} = (x_{t+1} - x_t) / dt; adjoint by the implicit
function theorem with the exact (unprojected) Hessian.
"""
import numpy as np

import oracle as O


def _dense(kr, kc, kv, n):
    A = np.zeros((3 * n, 3 * n))
    for k in range(len(kr)):
        M = kv[k].reshape(3, 3, order="F")
        i, j = kr[k], kc[k]
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += M
        if i != j:
            A[3 * j:3 * j + 3, 3 * i:3 * i + 3] += M.T
    return A


class RolloutRef:
    def __init__(self, sc, slack=0.8, newton_tol=1e-11, max_newton=60):
        m = sc.meta
        self.sc = sc
        self.n, self.T = sc.n, sc.T
        self.dt = sc.dt
        self.surf = np.ascontiguousarray(m["surf"], np.uint32)
        self.nrm = np.asarray(m["ground_normal"], np.float64)
        self.off = float(m["ground_offset"])
        self.dhat, self.kappa = sc.dhat, sc.kappa
        mu1, lam1 = _lame(m["nu"])
        self.c_mu, self.c_lam = mu1, lam1
        self.g = np.array([0.0, m["gravity"], 0.0])
        self.free = np.repeat(sc.btype == 0, 3).astype(np.float64)
        self.slack, self.newton_tol, self.max_newton = slack, newton_tol, max_newton
        self.btype = sc.btype.astype(np.int32)

    def set_log_E(self, log_E):
        E = np.exp(np.asarray(log_E, np.float64))
        self.mu, self.lam = self.c_mu * E, self.c_lam * E

    def x_tilde(self, x, v):
        xt = x + self.dt * v + self.dt ** 2 * np.tile(self.g, self.n)
        return self.free * xt + (1 - self.free) * x

    def energy(self, x, xt):
        sc = self.sc
        X = x.reshape(-1, 3)
        ek = 0.5 * np.sum(sc.mass[:, None] * (X - xt.reshape(-1, 3)) ** 2)
        ef = sum(O.snk1_energy(X[sc.tets[t].astype(int)].reshape(-1), sc.DmInv[t], sc.vol[t], self.mu[t], self.lam[t])
                 for t in range(self.T)) * self.dt ** 2
        eg = O.ground_energy(self.surf, x, self.nrm, self.off, self.dhat, self.kappa)
        return ek + ef + eg

    def system(self, x, xt, project=True):
        """(dense Hessian, gradient b) at x: kinetic + ground + SNK1 + mass, pinned rows masked."""
        sc = self.sc
        X = x.reshape(-1, 3)
        shape = O.kinetic_gradient(X, xt.reshape(-1, 3), sc.mass)
        gp = O.ground_detect(self.surf, x, self.nrm, self.off, self.dhat)
        fb, gv, gr, gc = O.ground_assemble(gp, x, self.nrm, self.off, self.dhat, self.kappa, btype=self.btype)
        with O.fem_projection(project):
            grad, fv, fr, fc = O.fem_assemble(X, sc.tets, sc.DmInv, sc.vol, self.mu, self.lam, self.dt, self.btype,
                                              0, shape.copy())
        mv, mr, mc = O.mass_triplets(sc.mass)
        R = np.concatenate([gr, fr, mr]); Cc = np.concatenate([gc, fc, mc]); V = np.concatenate([gv, fv, mv])
        b = O.fem_rhs(fb.reshape(-1, 3), grad, self.btype).reshape(-1)
        return _dense(*O.convert(R, Cc, V), self.n), b

    def step(self, x_t, v_t):
        xt = self.x_tilde(x_t, v_t)
        x = x_t.copy()
        prev = np.inf
        for _ in range(self.max_newton):
            A, b = self.system(x, xt)
            dx = np.linalg.solve(A, b) * self.free
            dxn = np.abs(dx).max()
            alpha = min(1.0, O.ground_step_bound(self.surf, x, dx, self.nrm, self.off, self.slack))
            if dxn > 1e-6:
                e0 = self.energy(x, xt)
                while True:
                    xn = x - alpha * dx
                    if self.energy(xn, xt) <= e0 or alpha < 1e-10:
                        break
                    alpha *= 0.5
                x = xn
            else:
                x = x - alpha * dx
            if alpha == 1.0 and (dxn <= self.newton_tol or (dxn < 1e-9 and dxn > 0.5 * prev)):
                break
            prev = dxn
        else:
            raise RuntimeError("reference Newton did not converge")
        return x, (x - x_t) / self.dt

    def rollout(self, log_E, v0, steps):
        self.set_log_E(log_E)
        x = self.sc.x_rest.reshape(-1).copy()
        v = np.asarray(v0, np.float64).reshape(-1) * self.free
        states = [(x, v)]
        for _ in range(steps):
            states.append(self.step(*states[-1]))
        return states

    def gradient(self, log_E, v0, target, steps):
        """(L, dL/d ln E, dL/d v0) for L = sum |x_T - target|^2."""
        states = self.rollout(log_E, v0, steps)
        xT = states[-1][0]
        L = float(np.sum((xT - target.reshape(-1)) ** 2))
        gx = 2.0 * (xT - target.reshape(-1)) * self.free
        gv = np.zeros_like(gx)
        gmu = np.zeros(self.T); glam = np.zeros(self.T)
        sc = self.sc
        m3 = np.repeat(sc.mass, 3)
        for t in range(steps - 1, -1, -1):
            x_t, v_t = states[t]
            x_n = states[t + 1][0]
            A, _ = self.system(x_n, self.x_tilde(x_t, v_t), project=False)
            xbar = (gx + gv / self.dt) * self.free
            lam = np.linalg.solve(A, xbar) * self.free
            dmu, dlam = O.fem_param_adjoint(x_n.reshape(-1, 3), sc.tets, sc.DmInv, sc.vol, self.dt, lam, self.btype)
            gmu -= dmu; glam -= dlam
            ml = m3 * lam * self.free
            gx, gv = ml - (gv / self.dt) * self.free, self.dt * ml
        return L, gmu * self.mu + glam * self.lam, gv.reshape(-1, 3)


def _lame(nu):
    lmu = 1.0 / (2.0 * (1.0 + nu))
    llam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return 4.0 * lmu / 3.0, llam + 5.0 * lmu / 6.0
