"""GPU parity of the batched PCG (independent scenes) through the C ABI, in
both forms -- the graph-captured launch sequence and the persistent launch
(a workgroup per scene): one scene against the oracle's PCG, several scenes
against a per-scene PCG written here over the oracle's MAS of the union
(its block-diagonality per scene checked on the way), the mixed-precision
form against a direct solve, the persistent form against the launch
sequence and against itself with fewer workgroups than scenes, the coupling
check, and the autotuner."""
import contextlib
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


@contextlib.contextmanager
def _mode(sysm, form, threads=1024, max_wg=0):
    """Run the enclosed solves in the given batched form ("sequence" or
    "persistent"), then restore the launch sequence."""
    sysm.ctx.call("sgipc_set_batch_mode", int(form == "persistent"), threads, max_wg)
    try:
        yield
    finally:
        sysm.ctx.call("sgipc_set_batch_mode", 0, 1024, 0)


FORMS = ["sequence", "persistent"]


@pytest.mark.parametrize("form", FORMS)
def test_batched_single_scene_matches_oracle_pcg(torch_cuda, oracle, cube10, form):
    from sgipc import capi
    from sgipc.system import NewtonSystem
    print("case: batched S=1 cube10 vs oracle pcg (%s)" % form, flush=True)
    ref = oracle_system(oracle, cube10)
    sysm = NewtonSystem(cube10, "cuda:0", precond=capi.PRECOND_MAS)
    sysm.assemble(); sysm.convert(); sysm.build_precond()
    sysm.ctx.call("sgipc_set_batch_mode", int(form == "persistent"), 1024, 0)
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


@pytest.mark.parametrize("form", FORMS)
def test_batched_scenes_match_per_scene_pcg(torch_cuda, oracle, five_scenes, form):
    from sgipc import capi
    scs, bn = five_scenes
    sysm = bn.sys
    print("case: batched S=5 vs per-scene PCG over the union MAS (%s)" % form, flush=True)
    kr, kc, kv = sysm.bcoo()
    n = bn.sc.n
    A = _csr(kr, kc, kv, n)
    b = sysm.b.cpu().numpy()
    assert bn.sim.n_close > 0                                    # ground contacts in the system
    mas = oracle.Mas(n, sysm.part_map, sysm.real_map, sysm.ns, sysm.nn, sysm.nl, bn.sc.bank)
    mas.build(kr, kc, kv)
    tol = 1e-4
    with _mode(sysm, form):
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


@pytest.mark.parametrize("form", FORMS)
def test_batched_mixed_reaches_tolerance_per_scene(torch_cuda, five_scenes, form):
    import scipy.sparse.linalg as spla
    from sgipc import capi
    scs, bn = five_scenes
    print("case: batched S=5 mixed FP32/FP64 to 1e-6 (%s)" % form, flush=True)
    bn.sys.build_precond(capi.PRECOND_MAS_FP32)
    with _mode(bn.sys, form):
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


def test_batched_half_bank_storage(torch_cuda, five_scenes):
    """The launch sequence's FP32 bank solves reading the FP16 storage
    (sgipc_set_bank_storage): every scene reaches the 1e-6 true residual
    (checked against scipy's matrix), with iteration counts within a few of
    the FP32 storage's; back in FP32 storage after a rebuild the iteration
    counts are the FP32 ones again (the solve's cached graph follows the
    storage)."""
    from sgipc import capi
    scs, bn = five_scenes
    print("case: batched S=5 mixed, FP16 bank storage (sequence)", flush=True)
    bn.sys.build_precond(capi.PRECOND_MAS_FP32)
    st32, _ = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_MIXED, 1e-6, 1)
    its32 = [st32[s].iterations for s in range(bn.S)]
    kr, kc, kv = bn.sys.bcoo()
    A = _csr(kr, kc, kv, bn.sc.n)
    b = bn.sys.b.cpu().numpy()
    try:
        bn.sys.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F16)
        bn.sys.build_precond(capi.PRECOND_MAS_FP32)
        st, x = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_MIXED, 1e-6, 1)
        its16 = [st[s].iterations for s in range(bn.S)]
        print("  iterations FP32 storage %s, FP16 storage %s" % (its32, its16), flush=True)
        for s in range(bn.S):
            sl = bn.scene_slice(s)
            rel = np.linalg.norm(b[sl] - A[sl, sl] @ x[sl]) / np.linalg.norm(b[sl])
            assert st[s].converged and st[s].rel_residual <= 1e-6 and rel <= 1.5e-6, (s, rel)
            assert abs(its16[s] - its32[s]) <= 3, (s, its16[s], its32[s])
    finally:
        bn.sys.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F32)
    bn.sys.build_precond(capi.PRECOND_MAS_FP32)
    st2, _ = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_MIXED, 1e-6, 1)
    # (the banked projection's atomic sums may reorder between builds: one
    # iteration of slack)
    assert all(abs(st2[s].iterations - its32[s]) <= 1 for s in range(bn.S))
    bn.sys.build_precond(capi.PRECOND_MAS)


@pytest.mark.parametrize("precond", ["MAS", "MAS_FP32"])
@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_persistent_matches_launch_sequence(torch_cuda, five_scenes, precond, precision):
    from sgipc import capi
    scs, bn = five_scenes
    if precond == "MAS" and precision == "MIXED":
        pytest.skip("mixed precision needs the FP32 MAS inverses")
    print("case: persistent vs launch sequence, S=5, %s, %s" % (precond, precision), flush=True)
    bn.sys.build_precond(getattr(capi, "PRECOND_" + precond))
    prec = getattr(capi, "PRECISION_" + precision)
    rule = 1 if precision == "MIXED" else 0
    tol = 1e-6 if precision == "MIXED" else 1e-5
    st0, x0 = _solve_batched(bn.sys, bn.row_start, prec, tol, rule)
    ref = [(st0[s].iterations, st0[s].converged) for s in range(bn.S)]
    with _mode(bn.sys, "persistent", 512):
        st1, x1 = _solve_batched(bn.sys, bn.row_start, prec, tol, rule)
        # fewer workgroups than scenes: each workgroup takes several scenes
        # from the queue; a scene's arithmetic does not depend on which
        with _mode(bn.sys, "persistent", 512, 2):
            st2, x2 = _solve_batched(bn.sys, bn.row_start, prec, tol, rule)
    kr, kc, kv = bn.sys.bcoo()
    A = _csr(kr, kc, kv, bn.sc.n)
    b = bn.sys.b.cpu().numpy()
    for s in range(bn.S):
        sl = bn.scene_slice(s)
        assert st1[s].converged and ref[s][1]
        assert abs(st1[s].iterations - ref[s][0]) <= 1, (s, st1[s].iterations, ref[s][0])
        # the true residual the kernel reports is the scene's FP64 b - A x
        rel = np.linalg.norm(b[sl] - A[sl, sl] @ x1[sl]) / np.linalg.norm(b[sl])
        assert abs(st1[s].rel_residual - rel) <= 1e-6 * rel + 1e-15, (s, st1[s].rel_residual, rel)
        if precision == "MIXED":
            assert rel <= 1.5 * tol
        scale = np.linalg.norm(x0[sl])
        assert np.linalg.norm(x1[sl] - x0[sl]) <= (1e-4 if precision == "MIXED" else 1e-8) * scale
        assert st2[s].iterations == st1[s].iterations and st2[s].refinements == st1[s].refinements
    assert np.array_equal(x1, x2)                                    # bitwise: same arithmetic per scene
    bn.sys.build_precond(capi.PRECOND_MAS)


def test_persistent_graph_replays(torch_cuda, five_scenes):
    """The persistent form is captured once per argument set and replayed:
    three replays, with eager launch-sequence solves between them, give the
    same bits each time."""
    from sgipc import capi
    scs, bn = five_scenes
    print("case: persistent graph replays", flush=True)
    outs = []
    for rep in range(3):
        with _mode(bn.sys, "persistent"):
            st, x = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-5, 0)
        outs.append((x, [st[s].iterations for s in range(bn.S)]))
        _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-3, 1)   # eager work in between
    for x, its in outs[1:]:
        assert np.array_equal(x, outs[0][0]) and its == outs[0][1]


def test_persistent_budget_cut(torch_cuda, five_scenes):
    from sgipc import capi
    scs, bn = five_scenes
    print("case: persistent iteration budget", flush=True)
    st0, _ = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-30, 0, max_iter=7)
    with _mode(bn.sys, "persistent"):
        st1, _ = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-30, 0, max_iter=7)
    for s in range(bn.S):
        assert st0[s].iterations == st1[s].iterations == 7
        assert not st0[s].converged and not st1[s].converged


@pytest.mark.parametrize("form", FORMS)
def test_batched_empty_scene(torch_cuda, five_scenes, form):
    """A scene with no block rows (row_start repeats a value): zero
    iterations, converged, and the other scenes unchanged."""
    from sgipc import capi
    scs, bn = five_scenes
    print("case: batched with an empty scene (%s)" % form, flush=True)
    rs = [int(v) for v in bn.row_start]
    rs_e = rs[:2] + [rs[1]] + rs[2:]                               # empty scene at index 1
    with _mode(bn.sys, form):
        st0, x0 = _solve_batched(bn.sys, rs, capi.PRECISION_FP64, 1e-5, 0)
        st1, x1 = _solve_batched(bn.sys, rs_e, capi.PRECISION_FP64, 1e-5, 0)
    assert st1[1].iterations == 0 and st1[1].converged and st1[1].rel_residual == 0.0
    others = [0] + list(range(2, bn.S + 1))
    for s0, s1 in zip(range(bn.S), others):
        assert st1[s1].iterations == st0[s0].iterations and st1[s1].converged
    assert np.array_equal(x0, x1)


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
    persistent, threads = bn.mode()
    assert threads in (512, 1024)
    print("autotune: group %d, iterations per graph %d, persistent %s (%d threads)" % (g, pg, persistent, threads),
          flush=True)
    st, x = _solve_batched(bn.sys, bn.row_start, capi.PRECISION_FP64, 1e-4, 0)
    # another lane split of the SpMV only reorders row sums (round-off)
    for s in range(bn.S):
        assert abs(st[s].iterations - it0[s]) <= 1 and st[s].converged
    assert np.linalg.norm(x - x0) <= 1e-3 * np.linalg.norm(x0)
    bn.sys.ctx.call("sgipc_set_schedule", 8, 4)
    bn.set_mode(False)
