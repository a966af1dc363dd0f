"""Python driver of one Newton step through the C ABI (tests and benchmarks).

The caller-side sequence mirrors GIPC::computeGradientAndHessian
(S/GIPC.cu:10110-10344) followed by GlobalLinearSystem::solve_linear_system
(S/linear_system/linear_system/global_linear_system.cu:124-133): kinetic
gradient, FEM triplets, mass diagonal, RHS, convert, preconditioner build,
PCG.  torch provides the device buffers and the stream.
"""
import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import check, lib


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def host_neighbors(n, tets):
    tets = np.ascontiguousarray(tets, dtype=np.uint32)
    T = tets.shape[0]
    start = np.zeros(max(n, 1), np.uint32); num = np.zeros(max(n, 1), np.uint32)
    lst = np.zeros(12 * T + 1, np.uint32)
    L = C.c_int(0)
    check(lib().sgipc_build_vertex_neighbors(n, T, tets.ctypes.data, start.ctypes.data, num.ctypes.data,
                                             lst.ctypes.data, lst.shape[0], C.byref(L)))
    return start[:n], num[:n], lst[:L.value].copy()


def host_bank_map(part, bank=16):
    part = np.ascontiguousarray(part, dtype=np.int32)
    n = part.shape[0]
    cap = bank * (int(part[-1]) + 1) if n else 0
    pm = np.zeros(max(cap, 1), np.int32); rm = np.zeros(max(n, 1), np.int32)
    M = C.c_int(0)
    check(lib().sgipc_build_bank_map(n, part.ctypes.data, bank, pm.ctypes.data, cap, C.byref(M), rm.ctypes.data))
    return pm[:M.value].copy(), rm[:n].copy()


class Context:
    def __init__(self, device=0, stream=None):
        h = C.c_void_p()
        check(lib().sgipc_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        check(lib().sgipc_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)), self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().sgipc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        return check(getattr(lib(), name)(self.h, *args), self.h)


class NewtonSystem:
    """Device-resident Newton-step system for one Scene."""

    def __init__(self, scene, device="cuda:0", precond=capi.PRECOND_MAS, extra_triplets=0):
        self.sc = scene
        self.dev = torch.device(device)
        self.ctx = Context(self.dev.index or 0)
        f64 = dict(dtype=torch.float64, device=self.dev)
        s = scene
        self.n, self.T = s.n, s.T
        self.x = torch.tensor(s.x.reshape(-1), **f64)
        self.x_tilde = torch.tensor(s.x_tilde.reshape(-1), **f64)
        self.x_rest = torch.tensor(s.x_rest.reshape(-1), **f64)
        self.mass = torch.tensor(s.mass, **f64)
        self.tets = torch.tensor(s.tets.astype(np.int32).reshape(-1), dtype=torch.int32, device=self.dev)
        self.DmInv = torch.tensor(s.DmInv.reshape(-1), **f64)
        self.vol = torch.tensor(s.vol, **f64)
        self.mu = torch.tensor(s.mu, **f64)
        self.lam = torch.tensor(s.lam, **f64)
        self.btype = torch.tensor(s.btype, dtype=torch.int32, device=self.dev)
        self.shape = torch.zeros(3 * self.n, **f64)
        self.fb = torch.zeros(3 * self.n, **f64)
        self.b = torch.zeros(3 * self.n, **f64)
        self.dx = torch.zeros(3 * self.n, **f64)
        self.cap = 10 * self.T + self.n + int(extra_triplets)
        self.tv = torch.zeros(9 * self.cap, **f64)
        self.tr = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.tc = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.ctx.call("sgipc_triplets_view", _p(self.tv), _p(self.tr), _p(self.tc), self.cap)
        self.ctx.call("sgipc_set_fem_mesh", self.n, self.T, np.ascontiguousarray(s.tets, np.uint32).ctypes.data)
        self.precond = precond
        self.ns, self.nn, self.nl = host_neighbors(self.n, s.tets)
        self.part_map, self.real_map = host_bank_map(s.part, s.bank)
        self.ctx.call("sgipc_set_mas_topology", self.n, 0, self.ns.ctypes.data, self.nn.ctypes.data,
                      self.nl.ctypes.data, self.nl.shape[0], self.part_map.ctypes.data, self.part_map.shape[0],
                      self.real_map.ctypes.data, s.bank)
        self.pairs = None
        self.mat_index = None
        self.counts = np.zeros(3, np.int32)
        self.n_pairs = 0
        self.n_contact_triplets = 0
        self.nt = 0
        self.K = 0

    def _set_capacity(self, cap):
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.cap = int(cap)
        self.tv = torch.zeros(9 * self.cap, **f64)
        self.tr = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.tc = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self.ctx.call("sgipc_triplets_view", _p(self.tv), _p(self.tr), _p(self.tc), self.cap)

    def setup_contacts(self, pt, ee):
        """Classify candidate quadruples on the GPU (narrow phase) at the
        current iterate and keep the resulting pairs for the Newton step."""
        pt = np.ascontiguousarray(pt, np.int32).reshape(-1, 4)
        ee = np.ascontiguousarray(ee, np.int32).reshape(-1, 4)
        m = pt.shape[0] + ee.shape[0]
        dpt = torch.tensor(pt.reshape(-1), dtype=torch.int32, device=self.dev) if len(pt) else None
        dee = torch.tensor(ee.reshape(-1), dtype=torch.int32, device=self.dev) if len(ee) else None
        pairs = torch.zeros(4 * max(m, 1), dtype=torch.int32, device=self.dev)
        mi = torch.zeros(max(m, 1), dtype=torch.int32, device=self.dev)
        npair = C.c_int(0)
        counts = np.zeros(3, np.int32)
        self.ctx.call("sgipc_classify_contacts", len(pt), _p(dpt), len(ee), _p(dee), _p(self.x), _p(self.x_rest),
                      float(self.sc.dhat), _p(pairs), _p(mi), m, C.byref(npair), counts.ctypes.data)
        self.n_pairs = npair.value
        self.pairs = pairs[:4 * self.n_pairs].clone() if self.n_pairs else None
        self.mat_index = mi[:self.n_pairs].clone() if self.n_pairs else None
        self.counts = counts
        self.n_contact_triplets = 4 * int(counts[0]) + 9 * int(counts[1]) + 16 * int(counts[2])
        self._set_capacity(10 * self.T + self.n + self.n_contact_triplets)
        return self.n_pairs

    # ---- steps --------------------------------------------------------------
    def assemble(self):
        s = self.sc
        self.ctx.call("sgipc_kinetic_gradient", self.n, _p(self.x), _p(self.x_tilde), _p(self.mass), _p(self.shape))
        self.fb.zero_()
        if self.n_pairs:
            self.ctx.call("sgipc_assemble_barrier", self.n_pairs, _p(self.pairs), _p(self.mat_index),
                          self.counts.ctypes.data, _p(self.x), _p(self.x_rest), float(s.dhat), float(s.kappa),
                          _p(self.btype), 0, 0, _p(self.fb))
        off = self.n_contact_triplets
        self.ctx.call("sgipc_assemble_fem", self.T, _p(self.x), _p(self.tets), _p(self.DmInv), _p(self.vol),
                      _p(self.mu), _p(self.lam), float(s.dt), _p(self.btype), 0, off, _p(self.shape))
        off += 10 * self.T
        self.ctx.call("sgipc_mass_triplets", self.n, _p(self.mass), 0, off)
        off += self.n
        self.nt = off
        self.ctx.call("sgipc_fem_rhs", self.n, _p(self.fb), _p(self.shape), _p(self.btype), _p(self.b))

    def convert(self):
        K = C.c_int(0)
        self.ctx.call("sgipc_convert", self.nt, self.n, C.byref(K))
        self.K = K.value
        return self.K

    def build_precond(self, kind=None):
        kind = self.precond if kind is None else kind
        self.ctx.call("sgipc_precond_build", kind, 0, self.n, _p(self.pairs) if self.n_pairs else None, self.n_pairs)

    def solve(self, tol=1e-4, stop_rule=capi.STOP_REF_RZ, precision=capi.PRECISION_FP64, max_iter=0,
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

    def energies(self):
        """FP64 line-search energies at the current x: kinetic, elastic
        (dt^2-scaled SNK1) and barrier (current contact pairs)."""
        s = self.sc
        ek = C.c_double(0.0); ef = C.c_double(0.0); eb = C.c_double(0.0)
        self.ctx.call("sgipc_kinetic_energy", self.n, _p(self.x), _p(self.x_tilde), _p(self.mass), C.byref(ek))
        self.ctx.call("sgipc_fem_energy", self.T, _p(self.x), _p(self.tets), _p(self.DmInv), _p(self.vol),
                      _p(self.mu), _p(self.lam), float(s.dt), C.byref(ef))
        if self.n_pairs:
            self.ctx.call("sgipc_barrier_energy", self.n_pairs, _p(self.pairs), _p(self.x), _p(self.x_rest),
                          float(s.dhat), float(s.kappa), C.byref(eb))
        return {"kinetic": ek.value, "elastic": ef.value, "barrier": eb.value}

    # ---- inspection -----------------------------------------------------------
    def triplets(self):
        torch.cuda.synchronize(self.dev)
        nt = self.nt
        return (self.tr[:nt].cpu().numpy(), self.tc[:nt].cpu().numpy(), self.tv[:9 * nt].cpu().numpy().reshape(nt, 9))

    def bcoo(self):
        torch.cuda.synchronize(self.dev)
        K = self.K
        return (self.tr[:K].cpu().numpy(), self.tc[:K].cpu().numpy(), self.tv[:9 * K].cpu().numpy().reshape(K, 9))

    def spmv(self, xvec, a=1.0, b=0.0, y=None):
        xt = torch.as_tensor(np.asarray(xvec, np.float64), device=self.dev)
        yt = torch.zeros(3 * self.n, dtype=torch.float64, device=self.dev) if y is None else \
            torch.as_tensor(np.asarray(y, np.float64), device=self.dev).clone()
        self.ctx.call("sgipc_spmv", float(a), _p(xt), float(b), _p(yt))
        torch.cuda.synchronize(self.dev)
        return yt.cpu().numpy()

    def precond_apply(self, r):
        rt = torch.as_tensor(np.asarray(r, np.float64), device=self.dev)
        zt = torch.zeros_like(rt)
        self.ctx.call("sgipc_precond_apply", _p(rt), _p(zt))
        torch.cuda.synchronize(self.dev)
        return zt.cpu().numpy()

    def mas_tables(self):
        L = C.c_int(0); tot = C.c_int(0)
        ls = np.zeros(16, np.int32)
        self.ctx.call("sgipc_mas_info", C.byref(L), C.byref(tot), ls.ctypes.data)
        L, tot, n = L.value, tot.value, self.n
        cs = np.zeros(L * n, np.int32); gn = np.zeros(max(tot, n), np.int32); ct = np.zeros(6 * n, np.int32)
        fm = np.zeros(n, np.uint32); p0 = np.zeros(len(self.part_map) // self.sc.bank, np.int32)
        self.ctx.call("sgipc_mas_tables", cs.ctypes.data, gn.ctypes.data, ct.ctypes.data, fm.ctypes.data,
                      p0.ctypes.data)
        return dict(levelnum=L, total=tot, level_size=ls[:2 * (L + 1)].reshape(-1, 2), coarse_space=cs.reshape(L, n),
                    going_next=gn, coarse_table=ct.reshape(n, 6), fine_mask=fm, prefix0=p0)

    def mas_levels(self):
        L = C.c_int(0)
        try:
            self.ctx.call("sgipc_mas_info", C.byref(L), None, None)
        except capi.SgipcError:
            return 0
        return L.value

    def bank_inverse(self, b):
        D = 3 * self.sc.bank
        out = np.zeros(D * D)
        self.ctx.call("sgipc_mas_bank_inverse", b, out.ctypes.data)
        return out.reshape(D, D)

    def bank_inverse_f32(self, b):
        D = 3 * self.sc.bank
        out = np.zeros(D * D, np.float32)
        self.ctx.call("sgipc_mas_bank_inverse_f32", b, out.ctypes.data)
        return out.reshape(D, D)

    def bank_inverse_f16(self, b):
        """D S D rebuilt from the half storage (after a build with
        sgipc_set_bank_storage(SGIPC_BANK_STORE_F16))."""
        D = 3 * self.sc.bank
        out = np.zeros(D * D, np.float32)
        self.ctx.call("sgipc_mas_bank_inverse_f16", b, out.ctypes.data)
        return out.reshape(D, D)
