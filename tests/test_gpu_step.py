"""GPU parity of the line search's step and CFL speed through the C ABI
against numpy: x = x_temp - alpha d on free vertices (all with
move_boundary; StiffGIPC/GIPC.cu:7937-7954) to one rounding (the kernel may
fuse the multiply-add), untouched elsewhere bit for bit, and the largest |d|
over a vertex list (cfl_largestSpeed, :8916-8947) equal to numpy's."""
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


@pytest.mark.parametrize("move_boundary", [0, 1])
def test_step_forward(torch_cuda, move_boundary):
    from sgipc.system import Context
    torch = torch_cuda
    print("case: step_forward move_boundary=%d" % move_boundary, flush=True)
    ctx = Context(0)
    rng = np.random.default_rng(3)
    n = 100003
    xt = rng.standard_normal((n, 3))
    d = rng.standard_normal((n, 3))
    bt = (rng.random(n) < 0.2).astype(np.int32) * rng.choice([-1, 1, 2], n).astype(np.int32)
    x0 = rng.standard_normal((n, 3))
    alpha = 0.37
    tx = torch.from_numpy(x0.copy()).cuda()
    txt = torch.from_numpy(xt).cuda()
    td = torch.from_numpy(d).cuda()
    tb = torch.from_numpy(bt).cuda()
    ctx.call("sgipc_step_forward", n, tx.data_ptr(), txt.data_ptr(), td.data_ptr(), tb.data_ptr(), alpha,
             move_boundary)
    got = tx.cpu().numpy()
    sel = np.ones(n, bool) if move_boundary else (bt == 0)
    tol = 4e-16 * (np.abs(xt) + np.abs(alpha * d))
    assert np.all(np.abs(got[sel] - (xt[sel] - alpha * d[sel])) <= tol[sel])
    assert np.array_equal(got[~sel], x0[~sel])             # pinned vertices untouched
    assert sel.sum() < n or move_boundary
    # no boundary array: every vertex moves
    ctx.call("sgipc_step_forward", n, tx.data_ptr(), txt.data_ptr(), td.data_ptr(), None, alpha, 0)
    assert np.all(np.abs(tx.cpu().numpy() - (xt - alpha * d)) <= tol)


@pytest.mark.parametrize("m", [0, 1, 777, 300000])
def test_cfl_largest_speed(torch_cuda, m):
    from sgipc.system import Context
    torch = torch_cuda
    print("case: cfl_largest_speed m=%d" % m, flush=True)
    ctx = Context(0)
    rng = np.random.default_rng(m)
    n = max(m, 10) * 2
    d = rng.standard_normal((n, 3))
    ids = rng.choice(n, size=m, replace=False).astype(np.uint32)
    d[ids[m // 2] if m else 0] *= 50.0            # one dominant vertex among the listed ones
    td = torch.from_numpy(d).cuda()
    tids = torch.from_numpy(ids.view(np.int32)).cuda()
    out = C.c_double(-1.0)
    ctx.call("sgipc_cfl_largest_speed", m, tids.data_ptr() if m else None, td.data_ptr(), C.byref(out))
    ref = float(np.max(np.sqrt((d[ids] ** 2).sum(1)))) if m else 0.0
    assert abs(out.value - ref) <= 1e-15 * max(ref, 1.0)
