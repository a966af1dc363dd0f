"""The reference-shaped C++ API (include/gipc_linear_system.hpp) as a GIPC
build would use it: stiff-gipc_amd/newton_step registers an FEM subsystem,
the PCG solver and a preconditioner with GlobalLinearSystem::create<T>()
and runs one Newton step through solve_linear_system(); hybrid_step does
the same for affine bodies over a soft body, with a create_LinearSystem
that makes the calls of the reference's GIPC::create_LinearSystem
(StiffGIPC/gipc/gipc.cu:42-70).  On the GPU their results must match the
Python driver's (same kernels, same order); on the CPU their error paths
must report cleanly."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "stiff-gipc_amd", "newton_step")


def write_scene(path, sc, precond=2, stop_rule=0, tol=1e-4, max_iter=0):
    with open(path, "wb") as f:
        f.write(b"SGIPCSC1")
        f.write(struct.pack("<6i2d", sc.n, sc.T, sc.bank, precond, stop_rule, max_iter, sc.dt, tol))
        for a, dt in ((sc.x, "<f8"), (sc.x_tilde, "<f8"), (sc.mass, "<f8"), (sc.btype, "<i4"), (sc.tets, "<u4"),
                      (sc.DmInv, "<f8"), (sc.vol, "<f8"), (sc.mu, "<f8"), (sc.lam, "<f8"), (sc.part, "<i4")):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def read_out(path, n):
    with open(path, "rb") as f:
        it, conv = struct.unpack("<2i", f.read(8))
        rr, = struct.unpack("<d", f.read(8))
        dx = np.frombuffer(f.read(), dtype="<f8")
    assert dx.size == 3 * n
    return it, conv, rr, dx


def have_gpu():
    import torch
    return torch.cuda.device_count() > 0


def test_facade_rejects_bad_scene_file(tmp_path):
    assert os.path.exists(EXE), "build first (make -C stiff-gipc_amd newton_step)"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTASCENE" + b"\0" * 64)
    p = subprocess.run([EXE, str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "bad scene" in p.stderr


def test_facade_reports_missing_device(tmp_path, cube4):
    if have_gpu():
        pytest.skip("a HIP device is present; the GPU test covers the facade")
    path = tmp_path / "cube4.bin"
    write_scene(path, cube4)
    p = subprocess.run([EXE, str(path), str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 3 and "sgipc_ctx_create" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("stop_rule,tol,precond", [(0, 1e-4, 2), (1, 1e-6, 2), (0, 1e-4, 1), (1, 1e-6, 3)])
def test_facade_newton_step_matches_python_driver(tmp_path, cube10, stop_rule, tol, precond):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sgipc import capi
    from sgipc.system import NewtonSystem
    print("case: C++ facade newton step cube10 rule=%d precond=%d" % (stop_rule, precond), flush=True)
    path = tmp_path / "cube10.bin"
    out = tmp_path / "out.bin"
    write_scene(path, cube10, precond=precond, stop_rule=stop_rule, tol=tol, max_iter=5000)
    p = subprocess.run([EXE, str(path), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    it, conv, rr, dx = read_out(out, cube10.n)
    sysm = NewtonSystem(cube10, "cuda:0", precond=precond)
    st = sysm.newton_step(tol=tol, stop_rule=stop_rule, max_iter=5000)
    ref = sysm.dx.cpu().numpy()
    assert it == st.iterations and conv == st.converged
    # same kernels and order; the coarse-level Galerkin projection of the MAS
    # build accumulates with FP64 atomics, so last bits may differ per build
    assert np.linalg.norm(dx - ref) <= 1e-10 * np.linalg.norm(ref)


HYB = os.path.join(ROOT, "stiff-gipc_amd", "hybrid_step")


def write_hybrid(path, hs, h, stop_rule=0, tol=1e-4, max_iter=0):
    """The hybrid scene and the contact pairs the Python driver classified."""
    fem = hs.fem
    npairs = h.n_pairs
    pairs = h.pairs.cpu().numpy() if npairs else np.zeros(0, np.int32)
    mi = h.mat_index.cpu().numpy() if npairs else np.zeros(0, np.int32)
    with open(path, "wb") as f:
        f.write(b"SGIPCHY1")
        f.write(struct.pack("<11i5d", hs.nb, hs.nav, fem.n, fem.T, fem.bank, npairs, *[int(c) for c in h.counts],
                            stop_rule, max_iter, fem.dt, fem.dhat, fem.kappa, hs.kappa_abd, tol))
        body_x = h.x[:3 * hs.nav].cpu().numpy()
        x_all = np.concatenate([body_x, fem.x.reshape(-1)])
        for a, dt in ((hs.vstart, "<i4"), (hs.xbar, "<f8"), (hs.mass12, "<f8"), (hs.volume, "<f8"), (hs.fixed, "<i4"),
                      (hs.q, "<f8"), (hs.q_tilde, "<f8"), (x_all, "<f8"), (h.x_rest.cpu().numpy(), "<f8"),
                      (h.btype.cpu().numpy(), "<i4"), (fem.x_tilde, "<f8"), (fem.mass, "<f8"), (fem.tets, "<u4"),
                      (fem.DmInv, "<f8"), (fem.vol, "<f8"), (fem.mu, "<f8"), (fem.lam, "<f8"), (fem.part, "<i4"),
                      (pairs, "<i4"), (mi, "<i4")):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def test_hybrid_facade_rejects_bad_scene_file(tmp_path):
    assert os.path.exists(HYB), "build first (make -C stiff-gipc_amd hybrid_step)"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"SGIPCSC1" + b"\0" * 64)
    p = subprocess.run([HYB, str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "bad scene" in p.stderr


@pytest.mark.gpu
def test_hybrid_facade_matches_python_driver(tmp_path):
    """Affine bodies through the reference-shaped API (ABD and FEM
    subsystems, PCGSolver with the PCG threshold, ABD + MAS preconditioners
    created as GIPC::create_LinearSystem does; the reference stop rule) give
    the Python driver's Newton direction."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sgipc import capi, scene
    from sgipc.hybrid import HybridSystem
    print("case: C++ facade hybrid step", flush=True)
    hs = scene.hybrid_scene(cells=(8, 2, 8), lattice=2, cube=0.2, jitter=0.02, max_turn=0.2, dt=0.01)
    h = HybridSystem(hs, "cuda:0")
    h.update_body_positions()
    assert h.detect() > 0
    path, out = tmp_path / "hyb.bin", tmp_path / "out.bin"
    write_hybrid(path, hs, h)
    p = subprocess.run([HYB, str(path), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    with open(out, "rb") as f:
        it, conv = struct.unpack("<2i", f.read(8))
        rr, = struct.unpack("<d", f.read(8))
        dx = np.frombuffer(f.read(), dtype="<f8")
    assert dx.size == 12 * hs.nb + 3 * hs.fem.n
    st = h.newton_step(tol=1e-4, stop_rule=capi.STOP_REF_RZ)
    ref = h.dx.cpu().numpy()
    assert conv == 1 and st.converged == 1 and it == st.iterations
    assert np.linalg.norm(dx - ref) <= 1e-10 * np.linalg.norm(ref)
