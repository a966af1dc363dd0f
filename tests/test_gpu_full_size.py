"""The benchmark workloads at their full sizes, checked through properties
that do not need the oracle at that size: every solve's true relative
residual ||b - A dx|| / ||b|| recomputed here with the library's FP64 SpMV
of the converted system (not the solver's own report), the converted BCOO
sorted, unique and upper, the mixed and FP32-apply directions against the
FP64 one, every batched scene to its own tolerance, and the cantilever
adjoint against Ridders-extrapolated central differences.

Workloads (the Newton-step path of StiffGIPC/GIPC.cu:10110-10344 ->
linear_system/linear_system/global_linear_system.cu:124-133):
  two stacked soft boxes, 2 x 40^3 cells (137,842 vertices, 640,000 tets),
  gap 0.5 dhat, contact pairs from the facing-surface candidates;
  64 affine cubes over a 100x20x100-cell slab (214,221 slab vertices);
  256 free boxes of 450-20,128 vertices over a ground plane;
  a 60x12x12-cell cantilever over the ground, 20 implicit steps.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def true_rel_residual(torch, ctx, b, dx):
    """||b - A dx|| / ||b|| with the library's FP64 SpMV (sgipc_spmv)."""
    y = b.clone()
    ctx.call("sgipc_spmv", -1.0, C.c_void_p(dx.data_ptr()), 1.0, C.c_void_p(y.data_ptr()))
    torch.cuda.synchronize()
    return float(y.norm() / b.norm()), y


def check_bcoo(torch, rows, cols, K):
    r = rows[:K].to(torch.int64)
    c = cols[:K].to(torch.int64)
    assert bool((r <= c).all())
    key = r * (int(c.max()) + 1) + c
    assert bool((key[1:] > key[:-1]).all())     # sorted by (row, col), no duplicates


@pytest.fixture(scope="module")
def two_box_40(torch_cuda):
    from sgipc import scene, capi
    from sgipc.system import NewtonSystem
    sc = scene.two_box_scene(40)
    pt, ee = scene.facing_contact_candidates(sc)
    s = NewtonSystem(sc, "cuda:0", precond=capi.PRECOND_MAS)
    s.setup_contacts(pt, ee)
    return s


def test_two_box_40_fp64_and_mixed(torch_cuda, two_box_40):
    from sgipc import capi
    torch = torch_cuda
    s = two_box_40
    print("case: two boxes 40^3 n=%d pairs=%d" % (s.n, s.n_pairs), flush=True)
    assert s.n == 137842 and s.n_pairs > 30000
    dx = {}
    for name, prec in (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED)):
        st = s.newton_step(tol=TOL, stop_rule=capi.STOP_TRUE_REL, precision=prec, max_iter=100000)
        rel, _ = true_rel_residual(torch, s.ctx, s.b, s.dx)
        print("  %s: %d iterations, reported %.3e, recomputed %.3e" % (name, st.iterations, st.rel_residual, rel),
              flush=True)
        assert st.converged and rel <= TOL * 1.001
        dx[name] = s.dx.clone()
    check_bcoo(torch, s.tr, s.tc, s.K)
    assert s.K > 900000
    diff = float((dx["mixed"] - dx["fp64"]).norm() / dx["fp64"].norm())
    print("  mixed vs fp64 direction: %.3e" % diff, flush=True)
    assert diff <= 1e-3


def test_hybrid_64_bodies_fp64_and_fp32_apply(torch_cuda):
    from sgipc import scene, capi
    from sgipc.hybrid import HybridSystem
    torch = torch_cuda
    hs = scene.hybrid_scene()
    h = HybridSystem(hs, "cuda:0")
    h.detect()
    print("case: hybrid 64 bodies, slab n=%d pairs=%d" % (h.nf, h.n_pairs), flush=True)
    assert h.nb == 64 and h.nf == 214221 and h.n_pairs > 0
    dx = {}
    for name, precond in (("fp64", capi.PRECOND_MAS), ("fp32apply", capi.PRECOND_MAS_FP32)):
        h.precond = precond
        st = h.newton_step(tol=TOL, stop_rule=capi.STOP_TRUE_REL, precision=capi.PRECISION_FP64, max_iter=100000)
        rel, _ = true_rel_residual(torch, h.ctx, h.b, h.dx)
        print("  %s: %d iterations, recomputed residual %.3e" % (name, st.iterations, rel), flush=True)
        assert st.converged and rel <= TOL * 1.001
        dx[name] = h.dx.clone()
    check_bcoo(torch, h.tr, h.tc, h.K)
    diff = float((dx["fp32apply"] - dx["fp64"]).norm() / dx["fp64"].norm())
    print("  fp32apply vs fp64 direction: %.3e" % diff, flush=True)
    assert diff <= 1e-3


def test_batched_256_scenes_each_converged(torch_cuda):
    from sgipc import scene, capi
    from sgipc.batch import BatchedNewton
    torch = torch_cuda
    scs = scene.batch_scenes(256)
    bn = BatchedNewton(scs, "cuda:0", precond=capi.PRECOND_MAS_FP32)
    print("case: batched 256 scenes, %d vertices" % bn.sc.n, flush=True)
    assert bn.S == 256 and bn.sc.n > 1000000
    st = bn.newton_step(precision=capi.PRECISION_MIXED, tol=TOL)
    s = bn.sys
    _, res = true_rel_residual(torch, s.ctx, s.b, s.dx)
    worst = 0.0
    for i in range(bn.S):
        sl = bn.scene_slice(i)
        bi = s.b[sl]
        rel = float(res[sl].norm() / bi.norm()) if float(bi.norm()) > 0 else 0.0
        worst = max(worst, rel)
        assert st[i].converged, i
        assert rel <= TOL * 1.001, (i, rel)
    its = [st[i].iterations for i in range(bn.S)]
    print("  worst scene residual %.3e, iterations %d-%d" % (worst, min(its), max(its)), flush=True)


def test_cantilever_60x12x12_adjoint_vs_extrapolated_fd(torch_cuda):
    from sgipc import scene
    from sgipc.sim import GroundSim, fd_check
    torch = torch_cuda
    sc = scene.cantilever_scene(60, 12, 12, seed=4)
    sim = GroundSim(sc, "cuda:0", newton_tol=1e-13, stall_tol=1e-10, pcg_tol=1e-12, adj_tol=1e-12)
    m = sc.meta
    dev = "cuda:0"
    logE = torch.tensor(m["log_E"], dtype=torch.float64, device=dev)
    v0 = torch.tensor(m["v0"], dtype=torch.float64, device=dev)
    target = torch.tensor(m["target"], dtype=torch.float64, device=dev)
    print("case: cantilever 60x12x12 n=%d T=%d, 20 steps, 4 directions" % (sc.n, sc.T), flush=True)

    def progress(i, fd, ad, rel):
        print("  direction %d: fd(h=1e-6) %.9e adjoint %.9e rel %.2e" % (i, fd, ad, rel), flush=True)
    # every direction also by Ridders' extrapolation of central differences
    res, L = fd_check(sim, logE, v0, target, 20, ckpt=5, n_dir=4, h=1e-6, seed=0, progress=progress,
                      ridders_above=0)
    assert sim.n_close > 0                      # the bar reaches the ground barrier
    for fd, ad, rel, fr, rr, er, hmin in res:
        print("  extrapolated fd %.9e rel %.2e (error estimate %.1e, smallest h %.1e)" % (fr, rr, er, hmin),
              flush=True)
        assert rr <= 1e-4
