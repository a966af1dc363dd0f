"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB)
        L.or_svd3.argtypes = [f64p, f64p, f64p, f64p]
        L.or_syev3.argtypes = [f64p, f64p, f64p]
        L.or_snk1_energy.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double]
        L.or_snk1_energy.restype = C.c_double
        L.or_snk1_tet.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_double, f64p, f64p]
        L.or_snk1_hq.argtypes = [f64p, C.c_double, C.c_double, f64p]
        L.or_fem_precompute.argtypes = [C.c_int, C.c_int, f64p, u32p, C.c_double, C.c_double, C.c_double,
                                        f64p, f64p, f64p, f64p, f64p]
        L.or_fem_assemble.argtypes = [C.c_int, C.c_int, f64p, u32p, f64p, f64p, f64p, f64p, C.c_double,
                                      C.c_void_p, C.c_int, f64p, f64p, i32p, i32p]
        L.or_kinetic_gradient.argtypes = [C.c_int, f64p, f64p, f64p, f64p]
        L.or_mass_triplets.argtypes = [C.c_int, f64p, C.c_int, f64p, i32p, i32p]
        L.or_fem_rhs.argtypes = [C.c_int, f64p, f64p, C.c_void_p, f64p]
        L.or_convert.argtypes = [C.c_int, i32p, i32p, f64p, i32p, i32p, f64p]
        L.or_convert.restype = C.c_int
        L.or_spmv_sym.argtypes = [C.c_int, i32p, i32p, f64p, C.c_int, C.c_double, f64p, C.c_double, f64p]
        L.or_vertex_neighbors.argtypes = [C.c_int, C.c_int, u32p, u32p, u32p, u32p]
        L.or_vertex_neighbors.restype = C.c_int
        L.or_bank_map.argtypes = [C.c_int, i32p, C.c_int, i32p, i32p]
        L.or_bank_map.restype = C.c_int
        L.or_mas_num_levels.argtypes = [C.c_int, C.c_int]
        L.or_mas_num_levels.restype = C.c_int
        L.or_mas_create.argtypes = [C.c_int, C.c_int, i32p, i32p, u32p, u32p, u32p, C.c_int]
        L.or_mas_create.restype = C.c_void_p
        L.or_mas_destroy.argtypes = [C.c_void_p]
        L.or_mas_build.argtypes = [C.c_void_p, C.c_int, i32p, i32p, f64p, C.c_int, C.c_void_p, C.c_int,
                                   C.c_int, C.c_int]
        L.or_mas_build.restype = C.c_int
        L.or_mas_apply.argtypes = [C.c_void_p, f64p, f64p, C.c_int]
        L.or_mas_levelnum.argtypes = [C.c_void_p]
        L.or_mas_levelnum.restype = C.c_int
        L.or_mas_total_clusters.argtypes = [C.c_void_p]
        L.or_mas_total_clusters.restype = C.c_int
        L.or_mas_tables.argtypes = [C.c_void_p, i32p, i32p, i32p, i32p, u32p, i32p]
        L.or_mas_bank_matrix.argtypes = [C.c_void_p, C.c_int, f64p, f64p]
        L.or_pcg.argtypes = [C.c_int, i32p, i32p, f64p, C.c_int, f64p, f64p, C.c_int, C.c_void_p, C.c_int,
                             C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.or_pcg.restype = C.c_int
        L.or_barrier_pair.argtypes = [f64p, f64p, i32p, C.c_double, C.c_double, i32p, C.POINTER(C.c_int), f64p, f64p,
                                      C.POINTER(C.c_double)]
        L.or_barrier_pair.restype = C.c_int
        L.or_barrier_assemble.argtypes = [C.c_int, i32p, i32p, i32p, f64p, f64p, C.c_double, C.c_double, C.c_void_p,
                                          f64p, f64p, i32p, i32p]
        L.or_barrier_energy.argtypes = [C.c_int, i32p, f64p, f64p, C.c_double, C.c_double]
        L.or_barrier_energy.restype = C.c_double
        L.or_classify.argtypes = [C.c_int, i32p, C.c_int, i32p, f64p, f64p, C.c_double, i32p, i32p, i32p]
        L.or_classify.restype = C.c_int
        L.or_set_fem_projection.argtypes = [C.c_int]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_accd.argtypes = [C.c_int, f64p, f64p, C.c_double]
        L.or_accd.restype = C.c_double
        L.or_ccd_step_bound.argtypes = [C.c_int, i32p, C.c_int, i32p, f64p, f64p, C.c_double]
        L.or_ccd_step_bound.restype = C.c_double
        L.or_ground_detect.argtypes = [C.c_int, u32p, f64p, f64p, C.c_double, C.c_double, u32p]
        L.or_ground_detect.restype = C.c_int
        L.or_ground_assemble.argtypes = [C.c_int, u32p, f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                         C.c_int, f64p, f64p, i32p, i32p]
        L.or_ground_energy.argtypes = [C.c_int, u32p, f64p, f64p, C.c_double, C.c_double, C.c_double]
        L.or_ground_energy.restype = C.c_double
        L.or_ground_step_bound.argtypes = [C.c_int, u32p, f64p, f64p, f64p, C.c_double, C.c_double]
        L.or_ground_step_bound.restype = C.c_double
        L.or_fem_param_adjoint.argtypes = [C.c_int, f64p, u32p, f64p, f64p, C.c_double, C.c_void_p, f64p, f64p,
                                           f64p]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def svd3(A):
    A = _c(A, np.float64).reshape(9)
    U = np.zeros(9); S = np.zeros(3); V = np.zeros(9)
    lib().or_svd3(A, U, S, V)
    return U.reshape(3, 3), S, V.reshape(3, 3)


def syev3(A):
    A = _c(A, np.float64).reshape(9)
    w = np.zeros(3); Q = np.zeros(9)
    lib().or_syev3(A, w, Q)
    return w, Q.reshape(3, 3)


def snk1_energy(xt, DmInv, vol, mu, lam):
    return lib().or_snk1_energy(_c(xt, np.float64).reshape(12), _c(DmInv, np.float64).reshape(9), vol, mu, lam)


def snk1_tet(xt, DmInv, vol, mu, lam, dt):
    g = np.zeros(12); H = np.zeros(144)
    lib().or_snk1_tet(_c(xt, np.float64).reshape(12), _c(DmInv, np.float64).reshape(9), vol, mu, lam, dt, g, H)
    return g, H.reshape(12, 12)


def snk1_hq(F, mu, lam):
    H = np.zeros(81)
    lib().or_snk1_hq(_c(F, np.float64).reshape(9), mu, lam, H)
    return H.reshape(9, 9)


def fem_precompute(xrest, tets, E, nu, density):
    n = xrest.shape[0]; T = tets.shape[0]
    DmInv = np.zeros((T, 9)); vol = np.zeros(T); mass = np.zeros(n); mu = np.zeros(T); lam = np.zeros(T)
    lib().or_fem_precompute(n, T, _c(xrest, np.float64).reshape(-1), _c(tets, np.uint32).reshape(-1), E, nu,
                            density, DmInv.reshape(-1), vol, mass, mu, lam)
    return DmInv, vol, mass, mu, lam


def fem_assemble(x, tets, DmInv, vol, mu, lam, dt, btype=None, index_offset=0, grad=None):
    n = x.shape[0]; T = tets.shape[0]
    if grad is None:
        grad = np.zeros((n, 3))
    vals = np.zeros((10 * T, 9)); rows = np.zeros(10 * T, np.int32); cols = np.zeros(10 * T, np.int32)
    bt = None if btype is None else _c(btype, np.int32)
    lib().or_fem_assemble(n, T, _c(x, np.float64).reshape(-1), _c(tets, np.uint32).reshape(-1),
                          _c(DmInv, np.float64).reshape(-1), _c(vol, np.float64), _c(mu, np.float64),
                          _c(lam, np.float64), dt, None if bt is None else bt.ctypes.data, index_offset,
                          grad.reshape(-1), vals.reshape(-1), rows, cols)
    return grad, vals, rows, cols


def kinetic_gradient(x, xt, mass):
    g = np.zeros_like(x)
    lib().or_kinetic_gradient(x.shape[0], _c(x, np.float64).reshape(-1), _c(xt, np.float64).reshape(-1),
                              _c(mass, np.float64), g.reshape(-1))
    return g


def mass_triplets(mass, index_offset=0):
    n = mass.shape[0]
    vals = np.zeros((n, 9)); rows = np.zeros(n, np.int32); cols = np.zeros(n, np.int32)
    lib().or_mass_triplets(n, _c(mass, np.float64), index_offset, vals.reshape(-1), rows, cols)
    return vals, rows, cols


def fem_rhs(fb, shape, btype):
    b = np.zeros_like(fb)
    bt = _c(btype, np.int32)
    lib().or_fem_rhs(fb.shape[0], _c(fb, np.float64).reshape(-1), _c(shape, np.float64).reshape(-1),
                     bt.ctypes.data, b.reshape(-1))
    return b


def convert(rows, cols, vals):
    nt = rows.shape[0]
    orow = np.zeros(max(nt, 1), np.int32); ocol = np.zeros(max(nt, 1), np.int32)
    oval = np.zeros((max(nt, 1), 9))
    K = lib().or_convert(nt, _c(rows, np.int32), _c(cols, np.int32), _c(vals, np.float64).reshape(-1),
                         orow, ocol, oval.reshape(-1))
    return orow[:K].copy(), ocol[:K].copy(), oval[:K].copy()


def spmv_sym(rows, cols, vals, nb, x, a=1.0, b=0.0, y=None):
    y = np.zeros(3 * nb) if y is None else np.array(y, dtype=np.float64).reshape(-1)
    lib().or_spmv_sym(rows.shape[0], _c(rows, np.int32), _c(cols, np.int32), _c(vals, np.float64).reshape(-1),
                      nb, a, _c(x, np.float64).reshape(-1), b, y)
    return y


def vertex_neighbors(n, tets):
    T = tets.shape[0]
    start = np.zeros(n, np.uint32); num = np.zeros(n, np.uint32); lst = np.zeros(12 * T + 1, np.uint32)
    L = lib().or_vertex_neighbors(n, T, _c(tets, np.uint32).reshape(-1), start, num, lst)
    return start, num, lst[:L].copy()


def bank_map(part_id, bank=16):
    part_id = _c(part_id, np.int32)
    n = part_id.shape[0]
    M = bank * (int(part_id[-1]) + 1)
    pm = np.zeros(M, np.int32); rm = np.zeros(n, np.int32)
    r = lib().or_bank_map(n, part_id, bank, pm, rm)
    if r < 0:
        raise ValueError("invalid partition (part larger than bank or unsorted)")
    return pm, rm


class Mas:
    """Oracle MAS preconditioner (MASPreconditioner.cu restated)."""

    def __init__(self, n, part_map, real_map, nstart, nnum, nlist, bank=16):
        self.n = n
        self.bank = bank
        self._keep = [_c(part_map, np.int32), _c(real_map, np.int32), _c(nstart, np.uint32),
                      _c(nnum, np.uint32), _c(np.append(nlist, 0), np.uint32)]
        self.h = lib().or_mas_create(n, len(part_map), *self._keep, bank)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mas_destroy(self.h)
            self.h = None

    def build(self, rows, cols, vals, offset=0, pairs=None, node_offset=0, fp32_store=True):
        p = None if pairs is None or len(pairs) == 0 else _c(pairs, np.int32).reshape(-1)
        np_ = 0 if p is None else p.shape[0] // 4
        return lib().or_mas_build(self.h, rows.shape[0], _c(rows, np.int32), _c(cols, np.int32),
                                  _c(vals, np.float64).reshape(-1), offset,
                                  None if p is None else p.ctypes.data, np_, node_offset, int(fp32_store))

    def apply(self, r, fp32=False):
        z = np.zeros(3 * self.n)
        lib().or_mas_apply(self.h, _c(r, np.float64).reshape(-1), z, int(fp32))
        return z

    @property
    def levelnum(self):
        return lib().or_mas_levelnum(self.h)

    @property
    def total(self):
        return lib().or_mas_total_clusters(self.h)

    def tables(self):
        L = self.levelnum; n = self.n; M = len(self._keep[0])
        ls = np.zeros(2 * (L + 1), np.int32); cs = np.zeros(L * n, np.int32)
        gn = np.zeros(max(self.total, n), np.int32); ct = np.zeros(6 * n, np.int32)
        fm = np.zeros(n, np.uint32); p0 = np.zeros(M // self.bank, np.int32)
        lib().or_mas_tables(self.h, ls, cs, gn, ct, fm, p0)
        return dict(level_size=ls.reshape(-1, 2), coarse_space=cs.reshape(L, n), going_next=gn,
                    coarse_table=ct.reshape(n, 6), fine_mask=fm, prefix0=p0)

    def bank_matrix(self, b):
        D = 3 * self.bank
        A = np.zeros(D * D); Ai = np.zeros(D * D)
        lib().or_mas_bank_matrix(self.h, b, A, Ai)
        return A.reshape(D, D), Ai.reshape(D, D)


def pcg(rows, cols, vals, nb, b, precond=2, mas=None, mas_fp32=False, tol=1e-4, rule=0, max_iter=None):
    x = np.zeros(3 * nb)
    rr = C.c_double(0.0)
    if max_iter is None:
        max_iter = int(0.3 * 3 * nb)
    it = lib().or_pcg(rows.shape[0], _c(rows, np.int32), _c(cols, np.int32), _c(vals, np.float64).reshape(-1),
                      nb, _c(b, np.float64).reshape(-1), x, precond, None if mas is None else mas.h,
                      int(mas_fp32), tol, rule, max_iter, C.byref(rr))
    return x, it, rr.value


def barrier_pair(x, xrest, q, dhat, kappa):
    vid = np.zeros(4, np.int32); cls = C.c_int(0); g = np.zeros(12); H = np.zeros(144); e = C.c_double(0)
    nv = lib().or_barrier_pair(_c(x, np.float64).reshape(-1), _c(xrest, np.float64).reshape(-1),
                               _c(q, np.int32).reshape(4), dhat, kappa, vid, C.byref(cls), g, H, C.byref(e))
    nd = 3 * nv
    return vid[:nv].copy(), cls.value, g[:nd].copy(), H[:nd * nd].reshape(nd, nd).copy(), e.value


def barrier_assemble(pairs, mat_index, counts, x, xrest, dhat, kappa, btype=None, fb=None):
    pairs = _c(pairs, np.int32).reshape(-1, 4)
    np_ = pairs.shape[0]
    counts = _c(counts, np.int32)
    nt = 16 * int(counts[2]) + 9 * int(counts[1]) + 4 * int(counts[0])
    fb = np.zeros(x.size) if fb is None else fb.reshape(-1).copy()
    vals = np.zeros((max(nt, 1), 9)); rows = np.zeros(max(nt, 1), np.int32); cols = np.zeros(max(nt, 1), np.int32)
    bt = None if btype is None else _c(btype, np.int32)
    lib().or_barrier_assemble(np_, pairs.reshape(-1), _c(mat_index, np.int32), counts,
                              _c(x, np.float64).reshape(-1), _c(xrest, np.float64).reshape(-1), dhat, kappa,
                              None if bt is None else bt.ctypes.data, fb, vals.reshape(-1), rows, cols)
    return fb, vals[:nt], rows[:nt], cols[:nt]


def barrier_energy(pairs, x, xrest, dhat, kappa):
    pairs = _c(pairs, np.int32).reshape(-1)
    return lib().or_barrier_energy(pairs.shape[0] // 4, pairs, _c(x, np.float64).reshape(-1),
                                   _c(xrest, np.float64).reshape(-1), dhat, kappa)


def classify(pt, ee, x, xrest, dhat):
    pt = _c(pt, np.int32).reshape(-1, 4); ee = _c(ee, np.int32).reshape(-1, 4)
    m = pt.shape[0] + ee.shape[0]
    pairs = np.zeros((max(m, 1), 4), np.int32); mi = np.zeros(max(m, 1), np.int32); counts = np.zeros(3, np.int32)
    npair = lib().or_classify(pt.shape[0], pt.reshape(-1), ee.shape[0], ee.reshape(-1),
                              _c(x, np.float64).reshape(-1), _c(xrest, np.float64).reshape(-1), dhat,
                              pairs.reshape(-1), mi, counts)
    return pairs[:npair].copy(), mi[:npair].copy(), counts


# ---- ground plane and parameter adjoint ---------------------------------------
class fem_projection:
    """context manager: SNK1 Hessian projection on/off in the oracle"""

    def __init__(self, on):
        self.on = on

    def __enter__(self):
        lib().or_set_fem_projection(int(self.on))

    def __exit__(self, *a):
        lib().or_set_fem_projection(1)


def set_threads(t):
    """Threads of the heavy stages (FEM assembly, convert, MAS build, PCG);
    1 runs the serial restatement."""
    lib().or_set_threads(int(t))


def get_threads():
    return lib().or_get_threads()


class threads:
    """with threads(t): ... runs the oracle's heavy stages on t threads."""

    def __init__(self, t):
        self.t = int(t)

    def __enter__(self):
        self.prev = get_threads()
        set_threads(self.t)

    def __exit__(self, *a):
        set_threads(self.prev)


def ground_detect(cand, x, normal, offset, dhat):
    cand = _c(cand, np.uint32)
    out = np.zeros(max(len(cand), 1), np.uint32)
    c = lib().or_ground_detect(len(cand), cand, _c(x, np.float64).reshape(-1), _c(normal, np.float64), offset,
                               dhat, out)
    return out[:c]


def ground_assemble(gp, x, normal, offset, dhat, kappa, btype=None, index_offset=0, fb=None):
    gp = _c(gp, np.uint32)
    x = _c(x, np.float64).reshape(-1)
    fb = np.zeros(x.shape[0]) if fb is None else _c(fb, np.float64).reshape(-1).copy()
    m = len(gp)
    vals = np.zeros(9 * max(m, 1)); rows = np.zeros(max(m, 1), np.int32); cols = np.zeros(max(m, 1), np.int32)
    bt = None if btype is None else _c(btype, np.int32)
    lib().or_ground_assemble(m, gp, x, _c(normal, np.float64), offset, dhat, kappa,
                             None if bt is None else bt.ctypes.data, index_offset, fb, vals, rows, cols)
    return fb, vals[:9 * m].reshape(m, 9), rows[:m], cols[:m]


def ground_energy(ids, x, normal, offset, dhat, kappa):
    ids = _c(ids, np.uint32)
    return lib().or_ground_energy(len(ids), ids, _c(x, np.float64).reshape(-1), _c(normal, np.float64), offset,
                                  dhat, kappa)


def ground_step_bound(ids, x, d, normal, offset, slack):
    ids = _c(ids, np.uint32)
    return lib().or_ground_step_bound(len(ids), ids, _c(x, np.float64).reshape(-1), _c(d, np.float64).reshape(-1),
                                      _c(normal, np.float64), offset, slack)


def fem_param_adjoint(x, tets, DmInv, vol, dt, adj, btype=None):
    T = tets.shape[0]
    dmu = np.zeros(max(T, 1)); dlam = np.zeros(max(T, 1))
    bt = None if btype is None else _c(btype, np.int32)
    lib().or_fem_param_adjoint(T, _c(x, np.float64).reshape(-1), _c(tets, np.uint32).reshape(-1),
                               _c(DmInv, np.float64).reshape(-1), _c(vol, np.float64),
                               dt, None if bt is None else bt.ctypes.data, _c(adj, np.float64).reshape(-1), dmu, dlam)
    return dmu[:T], dlam[:T]


def accd(pt, v, u, eta):
    """time of impact (fraction of the step) of 4 vertices v moving by u"""
    return lib().or_accd(int(pt), _c(v, np.float64).reshape(12), _c(u, np.float64).reshape(12), eta)


def ccd_step_bound(pt, ee, x, d, slack):
    pt = _c(pt, np.int32).reshape(-1, 4); ee = _c(ee, np.int32).reshape(-1, 4)
    return lib().or_ccd_step_bound(pt.shape[0], pt.reshape(-1), ee.shape[0], ee.reshape(-1),
                                   _c(x, np.float64).reshape(-1), _c(d, np.float64).reshape(-1), slack)
