"""GPU parity of the library's own stable radix sort (and, through the
convert, its exclusive scan) through the C ABI: against rocPRIM's radix
sort bit for bit (keys and the values that travel with them) and against a
stable numpy sort by the low `bits` bits, over tile edges (4096 keys per
tile), heavy duplication (stability), partial last digits and full 64-bit
keys; and a whole Newton-step convert with each sort giving the same BCOO
bits.  Timings of both sorts are printed."""
import ctypes as C
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    from sgipc.system import Context
    return Context(0)


def _sort(ctx, torch, keys, vals, bits, impl):
    n = keys.numel()
    ko = torch.empty_like(keys)
    vo = torch.empty(n, dtype=torch.int32, device=keys.device)
    ctx.call("sgipc_sort_pairs", keys.data_ptr() if n else None, ko.data_ptr() if n else None,
             vals.data_ptr() if vals is not None else None, vo.data_ptr() if n else None, n,
             keys.element_size(), bits, impl)
    return ko, vo


CASES = [(4, 0, 8), (4, 1, 8), (4, 100, 5), (4, 4095, 13), (4, 4096, 16), (4, 4097, 30), (4, 200000, 32),
         (8, 4097, 36), (8, 123457, 40), (8, 300001, 64), (8, 2000000, 36)]


@pytest.mark.parametrize("key_bytes,n,bits", CASES)
def test_own_sort_equals_rocprim_and_numpy(torch_cuda, ctx, key_bytes, n, bits):
    from sgipc import capi
    torch = torch_cuda
    print("case: sort n=%d key_bytes=%d bits=%d" % (n, key_bytes, bits), flush=True)
    rng = np.random.default_rng(n + bits)
    # few distinct low parts (stability matters), random high bits above `bits` (must be ignored)
    distinct = max(1, min(n // 7 + 1, 1 << min(bits, 20)))
    low = rng.integers(0, distinct, size=n, dtype=np.uint64) * np.uint64(((1 << bits) - 1) // max(distinct - 1, 1))
    low &= np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    top = 8 * key_bytes
    high = (rng.integers(0, 1 << 16, size=n, dtype=np.uint64) << np.uint64(bits)) if bits + 16 <= top else 0
    k = (low | high) if bits < top else low
    dt = np.uint32 if key_bytes == 4 else np.uint64
    k = k.astype(dt)
    vals = rng.integers(-2 ** 31, 2 ** 31 - 1, size=n, dtype=np.int64).astype(np.int32)
    tk = torch.from_numpy(k.view(np.int32 if key_bytes == 4 else np.int64)).to("cuda:0")
    tv = torch.from_numpy(vals).to("cuda:0")
    ko, vo = _sort(ctx, torch, tk, tv, bits, capi.SORT_OWN)
    kr, vr = _sort(ctx, torch, tk, tv, bits, capi.SORT_ROCPRIM)
    assert torch.equal(ko, kr) and torch.equal(vo, vr)
    # identity values (null input) -> the stable permutation
    _, po = _sort(ctx, torch, tk, None, bits, capi.SORT_OWN)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    perm = np.argsort(k.astype(np.uint64) & mask, kind="stable")
    assert np.array_equal(po.cpu().numpy(), perm.astype(np.int32))
    assert np.array_equal(ko.cpu().numpy().view(dt), k[perm])
    assert np.array_equal(vo.cpu().numpy(), vals[perm])
    if n >= 100000:
        torch.cuda.synchronize()
        times = {}
        for name, impl in (("own", capi.SORT_OWN), ("rocprim", capi.SORT_ROCPRIM)):
            _sort(ctx, torch, tk, tv, bits, impl)
            t0 = time.perf_counter()
            for _ in range(10):
                _sort(ctx, torch, tk, tv, bits, impl)
            times[name] = (time.perf_counter() - t0) / 10 * 1e6
        print("sort n=%d key_bytes=%d bits=%d: own %.1f us, rocPRIM %.1f us (host-timed, synchronised)" %
              (n, key_bytes, bits, times["own"], times["rocprim"]), flush=True)


def test_newton_step_same_bits_with_either_sort(torch_cuda):
    """Contact Newton step in the deterministic mode (ordered barrier
    gradient and MAS projection sums, both sorted): convert, SpMV structure,
    preconditioner and solve give the same bits with either sort."""
    from sgipc import capi, scene
    from sgipc.system import NewtonSystem
    print("case: two-box 8 Newton step, own sort vs rocPRIM", flush=True)
    sc = scene.two_box_scene(8, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    out = {}
    for impl in (capi.SORT_OWN, capi.SORT_ROCPRIM):
        sysm = NewtonSystem(sc, "cuda:0")
        sysm.ctx.call("sgipc_set_deterministic", 1)
        sysm.ctx.call("sgipc_set_sort", impl)
        sysm.setup_contacts(pt, ee)
        sysm.assemble(); sysm.convert(); sysm.build_precond()
        st = sysm.solve(tol=1e-6, stop_rule=1, max_iter=20000)
        kr, kc, kv = sysm.bcoo()
        out[impl] = (kr, kc, kv, sysm.b.cpu().numpy(), sysm.dx.cpu().numpy(), st.iterations)
    a, b = out[capi.SORT_OWN], out[capi.SORT_ROCPRIM]
    assert len(a[0]) > 0 and a[5] == b[5]
    for u, v in zip(a[:5], b[:5]):
        assert np.array_equal(u, v)
