"""The allocation-free SNK1 energy (sgipc_fem_energy_noalloc) against the
allocating entry point (sgipc_fem_energy), through the C ABI: bit-identical
results on two calls in a row, the second on a smaller input, with the
scratch left holding the first call's partials."""
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


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_energy_equals_allocating_entry_point(torch_cuda):
    from sgipc import scene
    from sgipc.system import NewtonSystem
    torch = torch_cuda
    sc = scene.cube_scene(10, seed=1)
    s = NewtonSystem(sc, "cuda:0")
    # a deformed state so every tet has nonzero energy
    x = s.x + 0.02 * torch.sin(3.0 * s.x)
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=s.dev)
    for T in (s.T, s.T // 3):      # the second call is smaller than the first
        print("case: fem energy noalloc, T=%d" % T, flush=True)
        ref = C.c_double()
        s.ctx.call("sgipc_fem_energy", T, _p(x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam), sc.dt,
                   C.byref(ref))
        s.ctx.call("sgipc_fem_energy_noalloc", T, _p(x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam),
                   sc.dt, C.c_void_p(out.data_ptr() + 8 * (T != s.T)))
        torch.cuda.synchronize()
        got = float(out[int(T != s.T)])
        assert np.isfinite(ref.value) and ref.value > 0
        assert got == ref.value, (T, got, ref.value)     # same grid, same summation order: same bits
    # invalid arguments are rejected before any launch
    from sgipc import capi
    with pytest.raises(capi.SgipcError):
        s.ctx.call("sgipc_fem_energy_noalloc", -1, _p(x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam),
                   sc.dt, C.c_void_p(out.data_ptr()))
    with pytest.raises(capi.SgipcError):
        s.ctx.call("sgipc_fem_energy_noalloc", s.T, _p(x), _p(s.tets), _p(s.DmInv), _p(s.vol), _p(s.mu), _p(s.lam),
                   sc.dt, None)
