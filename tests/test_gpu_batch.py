"""GPU parity of the batched PCG (independent scenes, one launch sequence)
through the C ABI: one scene against the oracle's PCG, several scenes
against a per-scene PCG written here over the oracle's MAS of the union
(its block-diagonality per scene checked on the way), the mixed-precision
form against a direct solve, the coupling check, and the autotuner."""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _csr(kr, kc, kv, n):
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for k in range(len(kr)):
        M = kv[k].reshape(3, 3, order="F")
        for p in range(3):
            for q in range(3):
                rows.append(3 * kr[k] + p); cols.append(3 * kc[k] + q); vals.append(M[p, q])
                if kr[k] != kc[k]:
                    rows.append(3 * kc[k] + q); cols.append(3 * kr[k] + p); vals.append(M[p, q])
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * n, 3 * n))


def _solve_batched(sysm, row_start, precision, tol, rule, max_iter=0):
    from sgipc import capi
    from sgipc.system import _p
    S = len(row_start) - 1
    st = (capi.SolveStats * S)()
    o = capi.SolveOpts(precision, rule, tol, 0.3, max_iter, 8)
    rs = np.ascontiguousarray(row_start, np.int32)
    sysm.ctx.call("sgipc_solve_batched", S, rs.ctypes.data, _p(sysm.b), _p(sysm.dx), C.byref(o), st)
    return st, sysm.dx.cpu().numpy()


def test_batched_single_scene_matches_oracle_pcg(torch_cuda, oracle, cube10):
    from sgipc import capi
    from sgipc.system import NewtonSystem
    print("case: batched S=1 cube10 vs oracle pcg", flush=True)
    ref = oracle_system(oracle, cube10)
    sysm = NewtonSystem(cube10, "cuda:0", precond=capi.PRECOND_MAS)
    sysm.assemble(); sysm.convert(); sysm.build_precond()
    kr, kc, kv = ref["bcoo"]
    mas = oracle.Mas(cube10.n, sysm.part_map, sysm.real_map, sysm.ns, sysm.nn, sysm.nl, cube10.bank)
    mas.build(kr, kc, kv)
    st, x = _solve_batched(sysm, [0, cube10.n], capi.PRECISION_FP64, 1e-4, 0)
    xo, ito, _ = oracle.pcg(kr, kc, kv, cube10.n, ref["b"], precond=2, mas=mas, tol=1e-4, rule=0)
    assert st[0].iterations == ito and st[0].converged
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    # the same budget cut as the oracle
    st, x = _solve_batched(sysm, [0, cube10.n], capi.PRECISION_FP64, 1e-30, 0, max_iter=6)
    xo, ito, _ = oracle.pcg(kr, kc, kv, cube10.n, ref["b"], precond=2, mas=mas, tol=1e-30, rule=0, max_iter=6)
    assert st[0].iterations == ito == 6 and not st[0].converged
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.fixture(scope="module")
def five_scenes(torch_cuda):
    from sgipc import scene
    from sgipc.batch import BatchedNewton
    from sgipc import capi
    scs = scene.batch_scenes(5, seed=11, n_min=150, n_max=900)
    bn = BatchedNewton(scs, "cuda:0", precond=capi.PRECOND_MAS)
    bn.assemble(); bn.convert(); bn.build_precond()
    return scs, bn


def test_batched_scenes_match_per_scene_pcg(torch_cuda, oracle, five_scenes):
    from sgipc import capi
    scs, bn = five_scenes
    sysm = bn.sys
    print("case: batched S=5 vs per-scene PCG over the union MAS", flush=True)
    kr, kc, kv = sysm.bcoo()
    n = bn.sc.n
    A = _csr(kr, kc, kv, n)
    b = sysm.b.cpu().numpy()
    assert bn.sim.n_close > 0                                    # ground contacts in the system
    mas = oracle.Mas(n, sysm.part_map, sysm.real_map, sysm.ns, sysm.nn, sysm.nl, bn.sc.bank)
    mas.build(kr, kc, kv)
    tol = 1e-4
    st, x = _solve_batched(sysm, bn.row_start, capi.PRECISION_FP64, tol, 0)
    for s in range(bn.S):
        sl = bn.scene_slice(s)
        # per-scene PCG (the reference loop) with the union MAS restricted to the scene
        def M(r):
            full = np.zeros(3 * n); full[sl] = r
            z = mas.apply(full)
            assert np.abs(np.delete(z, np.r_[sl])).max() == 0.0   # block-diagonal per scene
            return z[sl]
        As = A[sl, sl]; bs = b[sl]
        xs = np.zeros_like(bs); r = bs.copy(); z = M(r); rz = r @ z; rz0 = rz; p = z.copy()
        max_iter = int(0.3 * 3 * n)
        k = 1
        while k < max_iter:
            Ap = As @ p; al = rz / (p @ Ap)
            xs += al * p; r -= al * Ap
            if abs(rz) <= tol * rz0:
                break
            z = M(r); rzn = r @ z; p = z + (rzn / rz) * p; rz = rzn
            k += 1
        assert st[s].iterations == k, (s, st[s].iterations, k)
        assert np.linalg.norm(x[sl] - xs) <= 1e-10 * np.linalg.norm(xs)
    assert len(set(st[s].iterations for s in range(bn.S))) > 1    # scenes stop independently


def test_batched_mixed_reaches_tolerance_per_scene(torch_cuda, five_scenes):
    import scipy.sparse.linalg as spla
    from sgipc import capi
    scs, bn = five_scenes
    print("case: batched S=5 mixed FP32/FP64 to 1e-6", flush=True)
    bn.sys.build_precond(capi.PRECOND_MAS_FP32)
    st, x = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_MIXED, 1e-6, 1)
    kr, kc, kv = bn.sys.bcoo()
    A = _csr(kr, kc, kv, bn.sc.n)
    b = bn.sys.b.cpu().numpy()
    for s in range(bn.S):
        sl = bn.scene_slice(s)
        rel = np.linalg.norm(b[sl] - A[sl, sl] @ x[sl]) / np.linalg.norm(b[sl])
        assert st[s].converged and st[s].rel_residual <= 1e-6 and rel <= 1.5e-6
        xs = spla.spsolve(A[sl, sl].tocsc(), b[sl])
        assert np.linalg.norm(x[sl] - xs) <= 1e-3 * np.linalg.norm(xs)
    bn.sys.build_precond(capi.PRECOND_MAS)


def test_batched_rejects_coupled_split(torch_cuda, five_scenes):
    from sgipc import capi
    from sgipc.capi import SgipcError
    scs, bn = five_scenes
    print("case: batched coupling check", flush=True)
    rs = bn.row_start.copy()
    bad = np.array([0, rs[1] // 2] + list(rs[1:]), np.int32)     # cuts scene 0 in two
    with pytest.raises(SgipcError, match="couple two scenes"):
        _solve_batched(bn.sys, bad, capi.PRECISION_FP64, 1e-4, 0)


def test_autotune_then_batched_solve(torch_cuda, five_scenes):
    from sgipc import capi
    scs, bn = five_scenes
    print("case: autotune + batched solve", flush=True)
    st0, x0 = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-4, 0)
    it0 = [st0[s].iterations for s in range(bn.S)]
    g, pg = bn.autotune(capi.PRECISION_FP64)
    assert g in (4, 8, 16) and pg in (2, 4, 8)
    st, x = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-4, 0)
    # another lane split of the SpMV only reorders row sums (round-off)
    for s in range(bn.S):
        assert abs(st[s].iterations - it0[s]) <= 1 and st[s].converged
    assert np.linalg.norm(x - x0) <= 1e-3 * np.linalg.norm(x0)
    bn.sys.ctx.call("sgipc_set_schedule", 8, 4)
