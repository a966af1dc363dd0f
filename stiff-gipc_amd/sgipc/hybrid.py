"""Newton step of the hybrid stiff/soft system (affine bodies + FEM) through
the C ABI.

Caller-side sequence of the reference's GIPC::computeGradientAndHessian
with ABD (StiffGIPC/GIPC.cu:10110-10344): the contact barrier in vertex
space over all vertices (body vertices first), its blocks mapped into DOF
space (ABDSystem::setup_abd_system_gradient_hessian,
abd_system/abd_system.h:180-189), the FEM kinetic/elastic/mass terms on the
FEM block rows (offset 4 * bodies), the body gradient and Hessian, the RHS
(ABDLinearSubsystem / FEMLinearSubsystem::assemble,
linear_system/subsystem/abd_linear_subsystem.cu:15-56,
fem_linear_subsystem.cu:56-84), then convert, MAS over the FEM rows plus the
bodies' 12x12 block-Jacobi, and PCG.  torch provides device buffers and the
stream.

Triplet store layout per step: [mapped contact blocks | FEM tets | FEM mass
| body blocks] are converted; the vertex-space barrier blocks are written
past them (at the end of the store) and mapped from there.
"""
import ctypes as C

import numpy as np
import torch

from . import capi
from .system import Context, _p, host_bank_map, host_neighbors


class HybridSystem:
    def __init__(self, hs, device="cuda:0", precond=capi.PRECOND_MAS):
        self.hs = hs
        fem = hs.fem
        self.dev = torch.device(device)
        self.ctx = Context(self.dev.index or 0)
        f64 = dict(dtype=torch.float64, device=self.dev)
        i32 = dict(dtype=torch.int32, device=self.dev)
        self.nb, self.nav, self.nf, self.T = hs.nb, hs.nav, fem.n, fem.T
        self.nv = self.nav + self.nf
        self.nbr = 4 * self.nb + self.nf                  # block rows
        self.dt = float(fem.dt)
        self.dhat, self.kappa = float(fem.dhat), float(fem.kappa)
        body_x = np.asarray(hs.meta["positions"], np.float64)
        self.x = torch.tensor(np.concatenate([body_x, fem.x]).reshape(-1), **f64)
        self.x_rest = torch.tensor(np.concatenate([body_x, fem.x_rest]).reshape(-1), **f64)
        self.xf = self.x[3 * self.nav:]                   # FEM view of the vertex positions
        self.x_tilde = torch.tensor(fem.x_tilde.reshape(-1), **f64)
        self.q = torch.tensor(hs.q.reshape(-1), **f64)
        self.q_tilde = torch.tensor(hs.q_tilde.reshape(-1), **f64)
        self.mass = torch.tensor(fem.mass, **f64)
        self.tets = torch.tensor(fem.tets.astype(np.int32).reshape(-1), **i32)
        self.DmInv = torch.tensor(fem.DmInv.reshape(-1), **f64)
        self.vol = torch.tensor(fem.vol, **f64)
        self.mu = torch.tensor(fem.mu, **f64)
        self.lam = torch.tensor(fem.lam, **f64)
        self.btype_f = torch.tensor(fem.btype, **i32)
        self.btype = torch.tensor(np.concatenate([np.zeros(self.nav, np.int32), fem.btype]), **i32)
        self.shape = torch.zeros(3 * self.nf, **f64)
        self.fb = torch.zeros(3 * self.nv, **f64)
        self.b = torch.zeros(3 * self.nbr, **f64)
        self.dx = torch.zeros(3 * self.nbr, **f64)
        self.faces = torch.tensor(hs.faces.astype(np.int32).reshape(-1), **i32)
        self.edges = torch.tensor(hs.edges.astype(np.int32).reshape(-1), **i32)
        self.surf = torch.tensor(hs.surf.astype(np.int32), **i32)
        self.body_id = torch.tensor(hs.body_id, **i32)
        self.precond = precond
        c = self.ctx
        c.call("sgipc_set_fem_mesh", self.nf, self.T, np.ascontiguousarray(fem.tets, np.uint32).ctypes.data)
        self.ns, self.nn, self.nl = host_neighbors(self.nf, fem.tets)
        self.part_map, self.real_map = host_bank_map(fem.part, fem.bank)
        # contact pair ids are vertex ids: the MAS subtracts nav
        c.call("sgipc_set_mas_topology", self.nf, self.nav, self.ns.ctypes.data, self.nn.ctypes.data,
               self.nl.ctypes.data, self.nl.shape[0], self.part_map.ctypes.data, self.part_map.shape[0],
               self.real_map.ctypes.data, fem.bank)
        self._h = dict(vstart=np.ascontiguousarray(hs.vstart, np.int32),
                       xbar=np.ascontiguousarray(hs.xbar, np.float64),
                       mass=np.ascontiguousarray(hs.mass12, np.float64),
                       vol=np.ascontiguousarray(hs.volume, np.float64),
                       fixed=np.ascontiguousarray(hs.fixed, np.int32))
        h = self._h
        c.call("sgipc_abd_setup", self.nb, self.nav, h["vstart"].ctypes.data, h["xbar"].ctypes.data,
               h["mass"].ctypes.data, h["vol"].ctypes.data, h["fixed"].ctypes.data)
        self.pairs = None
        self.mat_index = None
        self.counts = np.zeros(3, np.int32)
        self.n_pairs = 0
        self.n_src = 0
        self.n_mapped = 0
        self.nt = 0
        self.K = 0
        self.cap = 0
        self._set_capacity(10 * self.T + self.nf + 10 * self.nb)

    def _set_capacity(self, cap):
        if cap <= self.cap:
            return
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.cap = int(cap)
        self.tv = torch.zeros(9 * self.cap, **f64)
        self.tr = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.tc = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.ctx.call("sgipc_triplets_view", _p(self.tv), _p(self.tr), _p(self.tc), self.cap)

    # ---- state ----------------------------------------------------------------
    def update_body_positions(self):
        """Body vertex positions from q (x_i = J_i q)."""
        self.ctx.call("sgipc_abd_positions", _p(self.q), _p(self.x))

    def vertex_step(self, d):
        """Vertex-space step of a DOF-space step d: J_i dq for the body
        vertices, dx for the FEM vertices."""
        out = torch.empty(3 * self.nv, dtype=torch.float64, device=self.dev)
        self.ctx.call("sgipc_abd_positions", _p(d), _p(out))
        out[3 * self.nav:] = d[12 * self.nb:]
        return out

    # ---- contacts -------------------------------------------------------------
    def detect(self, cap=None):
        """LBVH broad phase over the union surface (pairs within one body
        skipped), then the narrow phase at the current positions."""
        cap = cap or 4 * (len(self.hs.surf) + len(self.hs.edges)) + 1024

        def run(pt, cp, ee, ce, npt, nee):
            self.ctx.call("sgipc_broad_phase", len(self.hs.surf), _p(self.surf), len(self.hs.faces), _p(self.faces),
                          len(self.hs.edges), _p(self.edges), _p(self.x), _p(self.btype), _p(self.body_id),
                          self.dhat, _p(pt), cp, C.byref(npt), _p(ee), ce, C.byref(nee))
        pt, npt, ee, nee = self._candidates(run, cap)
        return self.set_candidates(pt, npt, ee, nee)

    def _candidates(self, run, cap):
        """Candidate quadruples from a broad-phase call; grown and rerun once
        when the counts exceed the capacity (the library then writes none)."""
        dev = dict(dtype=torch.int32, device=self.dev)
        npt, nee = C.c_int(), C.c_int()
        cp = ce = cap
        for _ in range(2):
            pt = torch.zeros(4 * max(cp, 1), **dev)
            ee = torch.zeros(4 * max(ce, 1), **dev)
            try:
                run(pt, cp, ee, ce, npt, nee)
                return pt[:4 * npt.value], npt.value, ee[:4 * nee.value], nee.value
            except capi.SgipcError as e:
                if e.code != 3:
                    raise
                cp, ce = max(cp, npt.value), max(ce, nee.value)
        raise RuntimeError("broad phase: candidate counts kept exceeding the capacity")

    def set_candidates(self, pt, npt, ee, nee):
        m = npt + nee
        pairs = torch.zeros(4 * max(m, 1), dtype=torch.int32, device=self.dev)
        mi = torch.zeros(max(m, 1), dtype=torch.int32, device=self.dev)
        npair = C.c_int(0)
        counts = np.zeros(3, np.int32)
        self.ctx.call("sgipc_classify_contacts", npt, _p(pt) if npt else None, nee, _p(ee) if nee else None,
                      _p(self.x), _p(self.x_rest), self.dhat, _p(pairs), _p(mi), m, C.byref(npair),
                      counts.ctypes.data)
        self.n_pairs = npair.value
        self.pairs = pairs[:4 * self.n_pairs].clone() if self.n_pairs else None
        self.mat_index = mi[:self.n_pairs].clone() if self.n_pairs else None
        self.counts = counts
        self.n_src = 4 * int(counts[0]) + 9 * int(counts[1]) + 16 * int(counts[2])
        # mapped blocks: at most 16 per source block
        self._set_capacity(17 * self.n_src + 10 * self.T + self.nf + 10 * self.nb)
        return self.n_pairs

    # ---- one Newton step ------------------------------------------------------
    def assemble(self):
        c = self.ctx
        nb3 = 12 * self.nb
        src = self.cap - self.n_src
        self.fb.zero_()
        if self.n_pairs:
            c.call("sgipc_assemble_barrier", self.n_pairs, _p(self.pairs), _p(self.mat_index),
                   self.counts.ctypes.data, _p(self.x), _p(self.x_rest), self.dhat, self.kappa, _p(self.btype), 0,
                   src, _p(self.fb))
        n = C.c_longlong(0)
        c.call("sgipc_abd_map_contacts", src, self.n_src, 0, C.byref(n))
        self.n_mapped = off = n.value
        c.call("sgipc_kinetic_gradient", self.nf, _p(self.xf), _p(self.x_tilde), _p(self.mass), _p(self.shape))
        c.call("sgipc_assemble_fem", self.T, _p(self.xf), _p(self.tets), _p(self.DmInv), _p(self.vol), _p(self.mu),
               _p(self.lam), self.dt, _p(self.btype_f), 4 * self.nb, off, _p(self.shape))
        off += 10 * self.T
        c.call("sgipc_mass_triplets", self.nf, _p(self.mass), 4 * self.nb, off)
        off += self.nf
        bq = self.b[:nb3]
        c.call("sgipc_abd_assemble", _p(self.q), _p(self.q_tilde), self.hs.kappa_abd, self.dt, off, _p(bq))
        off += 10 * self.nb
        c.call("sgipc_abd_contact_gradient", _p(self.fb), _p(bq))
        c.call("sgipc_fem_rhs", self.nf, _p(self.fb[3 * self.nav:]), _p(self.shape), _p(self.btype_f),
               _p(self.b[nb3:]))
        self.nt = off

    def convert(self):
        K = C.c_int(0)
        self.ctx.call("sgipc_convert", self.nt, self.nbr, C.byref(K))
        self.K = K.value
        return self.K

    def build_precond(self, kind=None):
        kind = self.precond if kind is None else kind
        self.ctx.call("sgipc_precond_build", kind, 4 * self.nb, self.nf, _p(self.pairs) if self.n_pairs else None,
                      self.n_pairs)

    def solve(self, tol=1e-6, stop_rule=capi.STOP_TRUE_REL, precision=capi.PRECISION_FP64, max_iter=0,
              check_every=8, max_iter_ratio=0.3):
        o = capi.SolveOpts(precision, stop_rule, tol, max_iter_ratio, max_iter, check_every)
        st = capi.SolveStats()
        self.ctx.call("sgipc_solve", _p(self.b), _p(self.dx), C.byref(o), C.byref(st))
        return st

    def newton_step(self, **solve_kw):
        self.assemble()
        self.convert()
        self.build_precond()
        return self.solve(**solve_kw)

    def energy(self):
        """FP64 incremental potential at the current state: bodies (kinetic +
        shape), FEM kinetic + elastic, contact barrier of the current pairs."""
        c = self.ctx
        ea, ek, ef, eb = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        c.call("sgipc_abd_energy", _p(self.q), _p(self.q_tilde), self.hs.kappa_abd, self.dt, C.byref(ea))
        c.call("sgipc_kinetic_energy", self.nf, _p(self.xf), _p(self.x_tilde), _p(self.mass), C.byref(ek))
        c.call("sgipc_fem_energy", self.T, _p(self.xf), _p(self.tets), _p(self.DmInv), _p(self.vol), _p(self.mu),
               _p(self.lam), self.dt, C.byref(ef))
        if self.n_pairs:
            c.call("sgipc_barrier_energy", self.n_pairs, _p(self.pairs), _p(self.x), _p(self.x_rest), self.dhat,
                   self.kappa, C.byref(eb))
        return ea.value + ek.value + ef.value + eb.value

    def apply_step(self, alpha, d=None):
        """q -= alpha dq, x_FEM -= alpha dx, body vertices recomputed."""
        d = self.dx if d is None else d
        self.q.sub_(alpha * d[:12 * self.nb])
        self.xf.sub_(alpha * d[12 * self.nb:])
        self.update_body_positions()

    def step_bound(self, dv, slack=0.8, cap=None):
        """Largest alpha <= 1 along -dv (vertex space) that CCD admits:
        swept-box candidates, then additive CCD with the given slackness."""
        cap = cap or 4 * (len(self.hs.surf) + len(self.hs.edges)) + 1024

        def run(pt, cp, ee, ce, npt, nee):
            self.ctx.call("sgipc_broad_phase_ccd", len(self.hs.surf), _p(self.surf), len(self.hs.faces),
                          _p(self.faces), len(self.hs.edges), _p(self.edges), _p(self.x), _p(dv), 1.0, _p(self.btype),
                          _p(self.body_id), self.dhat, _p(pt), cp, C.byref(npt), _p(ee), ce, C.byref(nee))
        pt, npt, ee, nee = self._candidates(run, cap)
        a = C.c_double(1.0)
        self.ctx.call("sgipc_ccd_step_bound", npt, _p(pt) if npt else None, nee, _p(ee) if nee else None,
                      _p(self.x), _p(dv), slack, C.byref(a))
        return a.value

    def newton_solve(self, newton_tol=1e-9, max_newton=30, pcg_tol=1e-8, precision=capi.PRECISION_FP64, slack=0.8):
        """Minimise the incremental potential from the current state: per
        iteration detection, assembly, MAS-PCG, a CCD-bounded step and
        backtracking on the FP64 energy; stops when |dq, dx|_inf <=
        newton_tol.  Returns per-iteration (|d|_inf, alpha, pcg iterations)."""
        hist = []
        nq = 12 * self.nb
        for _ in range(max_newton):
            self.detect()
            st = self.newton_step(tol=pcg_tol, precision=precision)
            dn = float(self.dx.abs().max())
            alpha = min(1.0, self.step_bound(self.vertex_step(self.dx), slack))
            e0 = self.energy()
            q0, x0 = self.q.clone(), self.xf.clone()
            while True:
                self.q.copy_(q0 - alpha * self.dx[:nq])
                self.xf.copy_(x0 - alpha * self.dx[nq:])
                self.update_body_positions()
                if self.energy() <= e0 or alpha < 1e-10:
                    break
                alpha *= 0.5
            hist.append((dn, alpha, st.iterations))
            if dn <= newton_tol:
                break
        return hist

    # ---- inspection -----------------------------------------------------------
    def triplets(self):
        torch.cuda.synchronize(self.dev)
        nt = self.nt
        return (self.tr[:nt].cpu().numpy(), self.tc[:nt].cpu().numpy(), self.tv[:9 * nt].cpu().numpy().reshape(nt, 9))

    def source_triplets(self):
        """The vertex-space barrier blocks of the last assembly."""
        torch.cuda.synchronize(self.dev)
        s, m = self.cap - self.n_src, self.n_src
        return (self.tr[s:s + m].cpu().numpy(), self.tc[s:s + m].cpu().numpy(),
                self.tv[9 * s:9 * (s + m)].cpu().numpy().reshape(m, 9))

    def bcoo(self):
        torch.cuda.synchronize(self.dev)
        K = self.K
        return (self.tr[:K].cpu().numpy(), self.tc[:K].cpu().numpy(), self.tv[:9 * K].cpu().numpy().reshape(K, 9))

    def precond_apply(self, r):
        rt = torch.as_tensor(np.asarray(r, np.float64), device=self.dev)
        zt = torch.zeros_like(rt)
        self.ctx.call("sgipc_precond_apply", _p(rt), _p(zt))
        torch.cuda.synchronize(self.dev)
        return zt.cpu().numpy()
