"""GPU parity of the ground-plane barrier, the exact-Hessian SNK1 mode and
the per-tet parameter adjoint against the oracle, through the C ABI, and the
checkpointed rollout adjoint (torch.autograd.Function) against central
finite differences of the GPU rollout."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _dev(torch, a, dt=None):
    return torch.tensor(np.ascontiguousarray(a).reshape(-1), dtype=dt or torch.float64, device="cuda:0")


def _plane(sc, gap):
    n = np.array([0.1, 1.0, 0.2]); n /= np.linalg.norm(n)
    return n, float((sc.x @ n).min() - gap)


def test_ground_kernels_match_oracle(torch_cuda, oracle, cube4):
    import torch
    from sgipc.system import NewtonSystem, _p
    sc = cube4
    print("case: ground kernels cube4", flush=True)
    dhat, kappa = 0.05 ** 2, 1e3
    nrm, off = _plane(sc, 0.02)
    sysm = NewtonSystem(sc, "cuda:0", extra_triplets=sc.n)
    ids = np.arange(sc.n, dtype=np.uint32)[::-1].copy()          # candidate order is kept
    d_ids = _dev(torch, ids.astype(np.int32), torch.int32)
    out = torch.zeros(sc.n, dtype=torch.int32, device="cuda:0")
    k = C.c_int()
    sysm.ctx.call("sgipc_ground_detect", sc.n, _p(d_ids), _p(sysm.x), nrm.ctypes.data, off, dhat, _p(out), C.byref(k))
    gp = oracle.ground_detect(ids, sc.x, nrm, off, dhat)
    assert k.value == len(gp) > 0
    assert np.array_equal(out[:k.value].cpu().numpy().astype(np.uint32), gp)
    # assembly: gradient into fb, one 3x3 block per close vertex (pinned zeroed)
    sysm.fb.zero_()
    sysm.ctx.call("sgipc_assemble_ground", k.value, _p(out), _p(sysm.x), nrm.ctypes.data, off, dhat, kappa,
                  _p(sysm.btype), 0, 5, _p(sysm.fb))
    torch.cuda.synchronize()
    fbo, vo, ro, co = oracle.ground_assemble(gp, sc.x, nrm, off, dhat, kappa, btype=sc.btype)
    fb = sysm.fb.cpu().numpy()
    assert np.abs(fb - fbo).max() <= 1e-13 * np.abs(fbo).max()
    m = k.value
    v = sysm.tv[45:45 + 9 * m].cpu().numpy().reshape(m, 9)
    assert np.array_equal(sysm.tr[5:5 + m].cpu().numpy(), ro) and np.array_equal(sysm.tc[5:5 + m].cpu().numpy(), co)
    assert np.abs(v - vo).max() <= 1e-13 * np.abs(vo).max()
    # energy and step bound
    e = C.c_double()
    sysm.ctx.call("sgipc_ground_energy", sc.n, _p(d_ids), _p(sysm.x), nrm.ctypes.data, off, dhat, kappa, C.byref(e))
    eo = oracle.ground_energy(ids, sc.x, nrm, off, dhat, kappa)
    assert eo > 0 and abs(e.value - eo) <= 1e-13 * eo
    rng = np.random.default_rng(2)
    for scale in (1e-3, 0.05, 1.0):
        d = rng.normal(size=3 * sc.n) * scale
        a = C.c_double()
        sysm.ctx.call("sgipc_ground_step_bound", sc.n, _p(d_ids), _p(sysm.x), _p(_dev(torch, d)), nrm.ctypes.data,
                      off, 0.8, C.byref(a))
        assert a.value == oracle.ground_step_bound(ids, sc.x, d, nrm, off, 0.8)


def test_exact_hessian_mode_and_param_adjoint_match_oracle(torch_cuda, oracle, cube4):
    import torch
    from sgipc.system import NewtonSystem, _p
    sc = cube4
    print("case: exact hessian + param adjoint cube4", flush=True)
    rng = np.random.default_rng(4)
    x = sc.x_rest * np.array([1.0, 0.55, 1.3]) + rng.normal(size=sc.x.shape) * 0.03
    sysm = NewtonSystem(sc, "cuda:0")
    sysm.x.copy_(_dev(torch, x))
    sysm.ctx.call("sgipc_set_fem_projection", 0)
    try:
        sysm.assemble()
    finally:
        sysm.ctx.call("sgipc_set_fem_projection", 1)
    R, Cc, V = sysm.triplets()
    with oracle.fem_projection(False):
        g, Vo, Ro, Co = oracle.fem_assemble(x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype)
    nT = 10 * sc.T
    assert np.array_equal(R[:nT], Ro) and np.array_equal(Cc[:nT], Co)
    assert np.abs(V[:nT] - Vo).max() <= 1e-12 * np.abs(Vo).max()
    # projection is back on for the next assembly
    sysm.assemble()
    _, Vp, _, _ = oracle.fem_assemble(x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype)
    assert np.abs(sysm.triplets()[2][:nT] - Vp).max() <= 1e-12 * np.abs(Vp).max()
    assert np.abs(Vp - Vo).max() > 1e-6 * np.abs(Vo).max()
    # per-tet parameter adjoint
    adj = rng.normal(size=3 * sc.n)
    dmu = torch.zeros(sc.T, dtype=torch.float64, device="cuda:0")
    dlam = torch.zeros_like(dmu)
    sysm.ctx.call("sgipc_fem_param_adjoint", sc.T, _p(sysm.x), _p(sysm.tets), _p(sysm.DmInv), _p(sysm.vol),
                  float(sc.dt), _p(sysm.btype), _p(_dev(torch, adj)), _p(dmu), _p(dlam))
    om, ol = oracle.fem_param_adjoint(x, sc.tets, sc.DmInv, sc.vol, sc.dt, adj, sc.btype)
    assert np.abs(dmu.cpu().numpy() - om).max() <= 1e-12 * np.abs(om).max()
    assert np.abs(dlam.cpu().numpy() - ol).max() <= 1e-12 * np.abs(ol).max()


@pytest.fixture(scope="module")
def small_bar(torch_cuda):
    from sgipc import scene
    from sgipc.sim import GroundSim
    sc = scene.cantilever_scene(12, 3, 3, seed=4, clearance=0.02)
    return sc, GroundSim(sc, "cuda:0")


def test_rollout_hits_ground_and_is_deterministic(torch_cuda, small_bar):
    import torch
    sc, sim = small_bar
    print("case: rollout cantilever 12x3x3, 12 steps", flush=True)
    m = sc.meta
    logE = torch.tensor(m["log_E"], dtype=torch.float64, device="cuda:0")
    v0 = torch.tensor(m["v0"], dtype=torch.float64, device="cuda:0")
    x1, _ = sim.rollout(logE, v0, 12)
    x2, _ = sim.rollout(logE, v0, 12)
    assert torch.equal(x1, x2)                    # bitwise repeatable (ordered MAS projection)
    y = x1.reshape(-1, 3)[:, 1].cpu().numpy()
    assert y.min() > 0.0                                               # never crosses the ground
    assert sim.n_close > 0                                             # in contact at the end
    assert np.abs(x1.reshape(-1, 3).cpu().numpy()[sc.btype != 0] - sc.x_rest[sc.btype != 0]).max() == 0.0


def test_rollout_adjoint_vs_central_fd(torch_cuda, small_bar):
    import torch
    from sgipc.sim import fd_check
    sc, sim = small_bar
    print("case: adjoint vs FD cantilever 12x3x3, 12 steps, ckpt 5, 6 directions", flush=True)
    m = sc.meta
    logE = torch.tensor(m["log_E"], dtype=torch.float64, device="cuda:0")
    v0 = torch.tensor(m["v0"], dtype=torch.float64, device="cuda:0")
    target = torch.tensor(m["target"], dtype=torch.float64, device="cuda:0")
    res, L = fd_check(sim, logE, v0, target, steps=12, ckpt=5, n_dir=6, h=1e-6, seed=1)
    for fd, ad, rel in res:
        print("  fd %.9e adjoint %.9e rel %.2e" % (fd, ad, rel), flush=True)
    assert L > 0
    assert max(r for _, _, r in res) <= 1e-4


def test_rollout_gradient_matches_cpu_reference(torch_cuda, small_bar):
    """GPU rollout adjoint (MAS-PCG solves, HIP parameter Jacobian) against
    the FP64 CPU reference rollout + adjoint (oracle kernels, dense solves)."""
    import torch
    from sgipc.sim import loss_and_gradient
    from rollout_ref import RolloutRef
    sc, sim = small_bar
    print("case: adjoint vs CPU reference cantilever 12x3x3, 8 steps", flush=True)
    m = sc.meta
    steps = 8
    logE = torch.tensor(m["log_E"], dtype=torch.float64, device="cuda:0")
    v0 = torch.tensor(m["v0"], dtype=torch.float64, device="cuda:0")
    target = torch.tensor(m["target"], dtype=torch.float64, device="cuda:0")
    L, gE, gv = loss_and_gradient(sim, logE, v0, target, steps, ckpt=3)
    Lr, gEr, gvr = RolloutRef(sc).gradient(m["log_E"], m["v0"], m["target"], steps)
    assert abs(L - Lr) <= 1e-10 * Lr
    gE, gv = gE.cpu().numpy(), gv.cpu().numpy()
    print("  |gE| %.6e rel err %.2e   |gv| %.6e rel err %.2e" % (
        np.linalg.norm(gEr), np.linalg.norm(gE - gEr) / np.linalg.norm(gEr),
        np.linalg.norm(gvr), np.linalg.norm(gv - gvr) / np.linalg.norm(gvr)), flush=True)
    assert np.linalg.norm(gE - gEr) <= 1e-7 * np.linalg.norm(gEr)
    assert np.linalg.norm(gv - gvr) <= 1e-7 * np.linalg.norm(gvr)
