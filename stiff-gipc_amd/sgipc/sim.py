"""Implicit time stepping of one tet body over a ground plane, and its
adjoint, on the library's Newton-step path.

Each time step minimises the incremental potential
    E(x) = 1/2 |x - x~|_M^2 + dt^2 Psi_SNK1(x; E_t) + kappa B_ground(x),
    x~ = x_t + dt v_t + dt^2 g (pinned vertices: x~ = x_t),
with Newton's method: ground detection, assembly (kinetic, ground barrier,
SNK1, mass), BCOO convert, MAS build and MAS-PCG through the C ABI, then a
step bounded by the ground step size and a backtracking line search on the
FP64 energies.  This is the caller loop of the reference's
GIPC::IPC_Solver / solve_subIP (StiffGIPC/GIPC.cu:10700-10964) reduced to
the ground barrier; v_{t+1} = (x_{t+1} - x_t) / dt.

The adjoint differentiates the converged minimiser by the implicit function
theorem: with g(x; x~, theta) = 0 at x_{t+1} and H = dg/dx (the exact
Hessian, SNK1 projection off), lambda = H^-1 xbar solves with the same
MAS-PCG, then x~bar = M lambda and theta_bar -= (dg/dtheta)^T lambda (the
per-tet parameter-Jacobian kernel).  States are checkpointed every `ckpt`
steps and recomputed segment by segment in the backward pass.
RolloutFunction exposes rollout + adjoint as a torch.autograd.Function over
(per-tet ln E, initial velocities).  The reference has no adjoint.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import capi
from .scene import lame_coefficients
from .system import NewtonSystem, _p


class GroundSim:
    def __init__(self, scene, device="cuda:0", newton_tol=1e-11, max_newton=60, pcg_tol=1e-10, adj_tol=1e-11,
                 slack=0.8, max_pcg=20000, stall_tol=1e-9):
        m = scene.meta
        self.sc = scene
        self.surf_np = np.ascontiguousarray(m["surf"], np.uint32)
        self.sys = NewtonSystem(scene, device, precond=capi.PRECOND_MAS, extra_triplets=len(self.surf_np))
        s = self.sys
        # bitwise-repeatable solves: finite differences of a rollout then see
        # no run-to-run round-off
        s.ctx.call("sgipc_set_deterministic", 1)
        self.dev = s.dev
        self.n, self.T = s.n, s.T
        self.dt = float(scene.dt)
        self.normal = np.ascontiguousarray(m["ground_normal"], np.float64)
        self.offset = float(m["ground_offset"])
        self.dhat, self.kappa = float(scene.dhat), float(scene.kappa)
        self.c_mu, self.c_lam = lame_coefficients(m["nu"])
        self.grav = torch.tensor([0.0, m["gravity"], 0.0], dtype=torch.float64, device=self.dev)
        self.newton_tol, self.max_newton, self.stall_tol = newton_tol, max_newton, stall_tol
        self.pcg_tol, self.adj_tol, self.slack, self.max_pcg = pcg_tol, adj_tol, slack, max_pcg
        i32 = dict(dtype=torch.int32, device=self.dev)
        self.surf = torch.tensor(self.surf_np.astype(np.int32), **i32)     # uint32 ids, int32 storage
        self.close = torch.zeros(max(len(self.surf_np), 1), **i32)
        self.n_close = 0
        self.free = torch.tensor(np.repeat(scene.btype == 0, 3), dtype=torch.float64, device=self.dev)
        self.mass3 = s.mass.repeat_interleave(3)
        self.dmu = torch.zeros(self.T, dtype=torch.float64, device=self.dev)
        self.dlam = torch.zeros(self.T, dtype=torch.float64, device=self.dev)
        self.stats = dict(newton=0, pcg=0, line_search=0, steps=0)

    # ---- parameters -----------------------------------------------------------
    def set_log_E(self, log_E):
        E = torch.exp(log_E.detach().to(self.dev, torch.float64))
        self.sys.mu.copy_(self.c_mu * E)
        self.sys.lam.copy_(self.c_lam * E)

    # ---- library calls --------------------------------------------------------
    def _call(self, name, *a):
        return self.sys.ctx.call(name, *a)

    def _detect(self):
        k = C.c_int(0)
        self._call("sgipc_ground_detect", len(self.surf_np), _p(self.surf), _p(self.sys.x), self.normal.ctypes.data,
                   self.offset, self.dhat, _p(self.close), C.byref(k))
        self.n_close = k.value

    def _assemble(self):
        s = self.sys
        self._call("sgipc_kinetic_gradient", self.n, _p(s.x), _p(s.x_tilde), _p(s.mass), _p(s.shape))
        s.fb.zero_()
        off = 0
        if self.n_close:
            self._call("sgipc_assemble_ground", self.n_close, _p(self.close), _p(s.x), self.normal.ctypes.data,
                       self.offset, self.dhat, self.kappa, _p(s.btype), 0, off, _p(s.fb))
            off += self.n_close
        self._call("sgipc_assemble_fem", self.T, _p(s.x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam),
                   self.dt, _p(s.btype), 0, off, _p(s.shape))
        off += 10 * self.T
        self._call("sgipc_mass_triplets", self.n, _p(s.mass), 0, off)
        s.nt = off + self.n
        self._call("sgipc_fem_rhs", self.n, _p(s.fb), _p(s.shape), _p(s.btype), _p(s.b))

    def energy(self):
        s = self.sys
        ek, ef, eg = C.c_double(), C.c_double(), C.c_double()
        self._call("sgipc_kinetic_energy", self.n, _p(s.x), _p(s.x_tilde), _p(s.mass), C.byref(ek))
        self._call("sgipc_fem_energy", self.T, _p(s.x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam),
                   self.dt, C.byref(ef))
        self._call("sgipc_ground_energy", len(self.surf_np), _p(self.surf), _p(s.x), self.normal.ctypes.data,
                   self.offset, self.dhat, self.kappa, C.byref(eg))
        return ek.value + ef.value + eg.value

    def _step_bound(self, d):
        a = C.c_double(1.0)
        self._call("sgipc_ground_step_bound", len(self.surf_np), _p(self.surf), _p(self.sys.x), _p(d),
                   self.normal.ctypes.data, self.offset, self.slack, C.byref(a))
        return a.value

    def _solve(self, tol):
        s = self.sys
        s.convert()
        s.build_precond()
        st = s.solve(tol=tol, stop_rule=capi.STOP_TRUE_REL, max_iter=self.max_pcg)
        self.stats["pcg"] += st.iterations
        return st

    # ---- forward ----------------------------------------------------------------
    def x_tilde_of(self, x_t, v_t):
        xt = x_t + self.dt * v_t + (self.dt * self.dt) * self.grav.repeat(self.n)
        return self.free * xt + (1.0 - self.free) * x_t

    def step(self, x_t, v_t):
        """One backward-Euler step; x_t, v_t flat (3n) FP64 device tensors."""
        s = self.sys
        s.x_tilde.copy_(self.x_tilde_of(x_t, v_t))
        s.x.copy_(x_t)
        prev = math.inf
        for it in range(self.max_newton):
            self._detect()
            self._assemble()
            self._solve(self.pcg_tol)
            # Dirichlet vertices stay put: the MAS coarse levels aggregate
            # pinned with free vertices and leave round-off on pinned DOFs
            s.dx.mul_(self.free)
            self.stats["newton"] += 1
            dxn = float(s.dx.abs().max())
            alpha = min(1.0, self._step_bound(s.dx))
            if dxn > 1e-6:
                # backtracking on the FP64 incremental potential (far from
                # the solution; the quadratic phase takes full steps)
                e0 = self.energy()
                x0 = s.x.clone()
                while True:
                    s.x.copy_(x0 - alpha * s.dx)
                    if self.energy() <= e0 or alpha < 1e-10:
                        break
                    alpha *= 0.5
                    self.stats["line_search"] += 1
            else:
                s.x.sub_(alpha * s.dx)
            if alpha == 1.0 and (dxn <= self.newton_tol or (dxn < self.stall_tol and dxn > 0.5 * prev)):
                break                       # converged, or stalled at round-off
            prev = dxn
        else:
            raise RuntimeError("Newton did not converge in %d iterations (|dx|_inf %.3e)" % (self.max_newton, dxn))
        self.stats["steps"] += 1
        x_n = s.x.clone()
        return x_n, (x_n - x_t) / self.dt

    def rollout(self, log_E, v0, steps, ckpt=5):
        """Returns (x_T, checkpoints {t: (x_t, v_t)} at t = 0, ckpt, 2 ckpt ...)."""
        self.set_log_E(log_E)
        x = torch.tensor(self.sc.x_rest.reshape(-1), dtype=torch.float64, device=self.dev)
        v = v0.detach().to(self.dev, torch.float64).reshape(-1) * self.free
        cps = {}
        for t in range(steps):
            if t % ckpt == 0:
                cps[t] = (x.clone(), v.clone())
            x, v = self.step(x, v)
        return x, cps

    # ---- adjoint ----------------------------------------------------------------
    def adjoint_step(self, x_t, v_t, x_n, gx, gv, gmu, glam):
        """Adjoints of (x_t, v_t) from those of (x_{t+1}, v_{t+1}); adds the
        parameter adjoints of this step into gmu, glam."""
        s = self.sys
        s.x_tilde.copy_(self.x_tilde_of(x_t, v_t))
        s.x.copy_(x_n)
        self._detect()
        self._call("sgipc_set_fem_projection", 0)
        try:
            self._assemble()
        finally:
            self._call("sgipc_set_fem_projection", 1)
        xbar = (gx + gv / self.dt) * self.free
        s.b.copy_(xbar)
        st = self._solve(self.adj_tol)
        lam = s.dx.mul_(self.free)
        self._call("sgipc_fem_param_adjoint", self.T, _p(s.x), _p(s.tets), _p(s.DmInv), _p(s.vol), self.dt,
                   _p(s.btype), _p(lam), _p(self.dmu), _p(self.dlam))
        gmu.sub_(self.dmu)
        glam.sub_(self.dlam)
        ml = self.mass3 * lam * self.free
        return ml - (gv / self.dt) * self.free, self.dt * ml, st

    def gradient(self, log_E, cps, gxT, steps, ckpt=5):
        """(dL/d ln E per tet, dL/d v0) from dL/dx_T, recomputing each
        checkpoint segment forward before walking it backward."""
        self.set_log_E(log_E)
        gx = gxT.detach().to(self.dev, torch.float64).reshape(-1) * self.free
        gv = torch.zeros_like(gx)
        gmu = torch.zeros(self.T, dtype=torch.float64, device=self.dev)
        glam = torch.zeros_like(gmu)
        for t0 in sorted(cps, reverse=True):
            t1 = min(t0 + ckpt, steps)
            states = [cps[t0]]
            for t in range(t0, t1):
                states.append(self.step(*states[-1]))
            for t in range(t1 - 1, t0 - 1, -1):
                x_t, v_t = states[t - t0]
                gx, gv, _ = self.adjoint_step(x_t, v_t, states[t + 1 - t0][0], gx, gv, gmu, glam)
        g_logE = gmu * self.sys.mu + glam * self.sys.lam       # d mu / d ln E = mu
        return g_logE, gv.reshape(-1, 3)


class RolloutFunction(torch.autograd.Function):
    """x_T = rollout(ln E, v0): `steps` implicit steps from rest with
    checkpoints every `ckpt` steps; backward runs the adjoint MAS-PCG solves."""

    @staticmethod
    def forward(ctx, log_E, v0, sim, steps, ckpt):
        xT, cps = sim.rollout(log_E, v0, steps, ckpt)
        ctx.sim, ctx.steps, ctx.ckpt, ctx.cps = sim, steps, ckpt, cps
        ctx.save_for_backward(log_E)
        return xT.reshape(-1, 3)

    @staticmethod
    def backward(ctx, gxT):
        (log_E,) = ctx.saved_tensors
        g_logE, g_v0 = ctx.sim.gradient(log_E, ctx.cps, gxT, ctx.steps, ctx.ckpt)
        return g_logE.to(log_E.dtype), g_v0, None, None, None


def loss_and_gradient(sim, log_E, v0, target, steps, ckpt=5):
    """L = sum |x_T - target|^2 and its gradient in (ln E, v0) via autograd."""
    le = log_E.detach().clone().requires_grad_(True)
    vv = v0.detach().clone().requires_grad_(True)
    xT = RolloutFunction.apply(le, vv, sim, steps, ckpt)
    L = ((xT - target) ** 2).sum()
    L.backward()
    return float(L.detach()), le.grad.detach(), vv.grad.detach()


def loss_only(sim, log_E, v0, target, steps):
    xT, _ = sim.rollout(log_E, v0, steps, ckpt=max(steps, 1))
    return float(((xT.reshape(-1, 3) - target) ** 2).sum())


def final_positions(sim, log_E, v0, steps):
    xT, _ = sim.rollout(log_E, v0, steps, ckpt=max(steps, 1))
    return xT.reshape(-1, 3)


def ridders(D, h0, con=2.0, levels=6, safe=2.0):
    """Derivative from central differences D(h) by Ridders' extrapolation
    (C. J. F. Ridders, Adv. Eng. Software 4(2), 1982): D(h) is taken at
    h0, h0/con, h0/con^2, ... and extrapolated to h -> 0 in a Neville
    tableau, each column cancelling the next even power of h.  Every entry
    carries an error estimate (its distance to the two entries it came
    from); the entry with the smallest estimate is returned, and the tableau
    stops growing once its diagonal moves by more than `safe` times that
    estimate (round-off has taken over).  Returns (estimate, error estimate,
    smallest h used)."""
    c2 = con * con
    prev = [D(h0)]
    best, err, h = prev[0], float("inf"), h0
    for i in range(1, levels):
        h /= con
        row = [D(h)]
        fac = c2
        for j in range(1, i + 1):
            row.append((row[j - 1] * fac - prev[j - 1]) / (fac - 1.0))
            fac *= c2
            e = max(abs(row[j] - row[j - 1]), abs(row[j] - prev[j - 1]))
            if e <= err:
                best, err = row[j], e
        if abs(row[i] - prev[i - 1]) >= safe * err:
            break
        prev = row
    return best, err, h


def fd_check(sim, log_E, v0, target, steps, ckpt=5, n_dir=32, h=1e-6, seed=0, progress=None, ridders_above=None,
             ridders_h0=1e-4):
    """Central differences of L along n_dir seeded unit directions in
    (ln E, free v0) space against the adjoint gradient.  Returns a list of
    (fd, adjoint, relative error) and the loss.  With ridders_above set,
    directions whose error exceeds it are also differentiated by Ridders'
    extrapolation of central differences from ridders_h0 down (the step
    chosen per direction by the tableau's error estimate): entries then carry
    (fd_r, rel_r, err_est_r, h_min_r).  ridders_above=0 does so for every
    direction."""
    L, gE, gv = loss_and_gradient(sim, log_E, v0, target, steps, ckpt)
    rng = np.random.default_rng(seed)
    free = sim.free.reshape(-1, 3)
    out = []
    for _ in range(n_dir):
        dE = torch.tensor(rng.normal(size=sim.T), dtype=torch.float64, device=sim.dev)
        dv = torch.tensor(rng.normal(size=(sim.n, 3)), dtype=torch.float64, device=sim.dev) * free
        nrm = math.sqrt(float((dE * dE).sum() + (dv * dv).sum()))
        dE /= nrm
        dv /= nrm
        def central(hh):
            xp = final_positions(sim, log_E + hh * dE, v0 + hh * dv, steps)
            xm = final_positions(sim, log_E - hh * dE, v0 - hh * dv, steps)
            # (L+ - L-) / 2h with L = |x_T - target|^2, summed as
            # (x+ - x-) . (x+ + x- - 2 target) so that no two O(L) numbers
            # are subtracted (their round-off would swamp small derivatives)
            return float(((xp - xm) * (xp + xm - 2.0 * target)).sum()) / (2 * hh)
        fd = central(h)
        ad = float((gE * dE).sum() + (gv * dv).sum())
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        if ridders_above is not None and (ridders_above <= 0 or rel > ridders_above):
            fr, er, hmin = ridders(central, ridders_h0)
            out.append((fd, ad, rel, fr, abs(fr - ad) / max(abs(ad), 1e-300), er, hmin))
        else:
            out.append((fd, ad, rel))
        if progress:
            progress(len(out), *out[-1][:3])
    return out, L
