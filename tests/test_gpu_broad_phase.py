"""GPU LBVH broad phase through the C ABI: candidate sets equal to the
brute-force all-pairs box test (point-triangle and edge-edge, with the
reference's filters), repeatable order, the capacity protocol, and the
contact pairs it yields through the GPU narrow phase."""
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


def _canon(pairs, n):
    q = np.array(pairs, np.int64).reshape(-1, 4).copy()
    q[q[:, 3] < 0, 3] = -1     # negative codes: only the sign is read (parallel edge pairs carry a candidate index)
    return set(map(tuple, q))


def _broad(torch, sysm, faces, edges, surf, dhat, btype=None, body=None, cap=None):
    from sgipc.system import _p
    dev = dict(device="cuda:0")
    f = torch.tensor(faces.astype(np.int32).reshape(-1), dtype=torch.int32, **dev)
    e = torch.tensor(edges.astype(np.int32).reshape(-1), dtype=torch.int32, **dev)
    v = torch.tensor(surf.astype(np.int32), dtype=torch.int32, **dev)
    bt = None if btype is None else torch.tensor(btype.astype(np.int32), dtype=torch.int32, **dev)
    bd = None if body is None else torch.tensor(body.astype(np.int32), dtype=torch.int32, **dev)
    cp, ce = (cap if cap is not None else (200000, 200000))
    pt = torch.zeros(4 * max(cp, 1), dtype=torch.int32, **dev)
    ee = torch.zeros(4 * max(ce, 1), dtype=torch.int32, **dev)
    npt, nee = C.c_int(), C.c_int()
    sysm.ctx.call("sgipc_broad_phase", len(surf), _p(v), len(faces), _p(f), len(edges), _p(e), _p(sysm.x), _p(bt),
                  _p(bd), float(dhat), _p(pt), cp, C.byref(npt), _p(ee), ce, C.byref(nee))
    return pt[:4 * npt.value].cpu().numpy().reshape(-1, 4), ee[:4 * nee.value].cpu().numpy().reshape(-1, 4)


@pytest.fixture(scope="module")
def twobox8(torch_cuda):
    from sgipc import scene
    from sgipc.system import NewtonSystem
    sc = scene.two_box_scene(8, seed=2)
    return sc, NewtonSystem(sc, "cuda:0"), scene.surface_primitives(sc.tets)


def test_candidates_equal_brute_force(torch_cuda, twobox8):
    from broad_phase_ref import candidates
    sc, sysm, (faces, edges, surf) = twobox8
    print("case: broad phase two_box 8 vs brute force", flush=True)
    pt, ee = _broad(torch_cuda, sysm, faces, edges, surf, sc.dhat)
    rpt, ree = candidates(sc.x, faces, edges, surf, sc.dhat)
    assert len(pt) == len(rpt) > 0 and len(ee) == len(ree) > 0
    assert set(map(tuple, pt)) == set(map(tuple, rpt))
    assert set(map(tuple, ee)) == set(map(tuple, ree))
    # query-major order, identical on a second run
    assert (np.diff(np.searchsorted(surf, pt[:, 0])) >= 0).all()
    pt2, ee2 = _broad(torch_cuda, sysm, faces, edges, surf, sc.dhat)
    assert np.array_equal(pt, pt2) and np.array_equal(ee, ee2)


def test_filters_and_capacity(torch_cuda, twobox8):
    from sgipc.capi import SgipcError
    from broad_phase_ref import candidates
    sc, sysm, (faces, edges, surf) = twobox8
    print("case: broad phase filters + capacity", flush=True)
    rng = np.random.default_rng(3)
    btype = (rng.random(sc.n) < 0.7).astype(np.int32) * 2     # many static vertices
    body = np.where(np.arange(sc.n) < sc.meta["n_lower"], 0, -1).astype(np.int32)
    pt, ee = _broad(torch_cuda, sysm, faces, edges, surf, 4 * sc.dhat, btype, body)
    rpt, ree = candidates(sc.x, faces, edges, surf, 4 * sc.dhat, btype, body)
    assert set(map(tuple, pt)) == set(map(tuple, rpt)) and set(map(tuple, ee)) == set(map(tuple, ree))
    with pytest.raises(SgipcError, match="exceed the capacities"):
        _broad(torch_cuda, sysm, faces, edges, surf, sc.dhat, cap=(3, 3))


def test_pairs_from_bvh_cover_grid_candidates(torch_cuda, oracle, twobox8):
    from sgipc import scene
    sc, sysm, (faces, edges, surf) = twobox8
    print("case: broad phase -> narrow phase pairs", flush=True)
    pt, ee = _broad(torch_cuda, sysm, faces, edges, surf, sc.dhat)
    n_bvh = sysm.setup_contacts(pt, ee)
    bvh_pairs = _canon(sysm.pairs.cpu().numpy(), sc.n)
    # the oracle's narrow phase on the same candidates gives the same pairs
    po, _, _ = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)
    assert n_bvh == len(po) > 0 and bvh_pairs == _canon(po, sc.n)
    # the grid-cell candidates of the benchmark find a subset
    gpt, gee = scene.facing_contact_candidates(sc)
    sysm.setup_contacts(gpt, gee)
    grid_pairs = _canon(sysm.pairs.cpu().numpy(), sc.n)
    assert grid_pairs <= bvh_pairs


def _broad_ccd(torch, sysm, faces, edges, surf, dhat, d, alpha, cap=(400000, 400000)):
    from sgipc.system import _p
    dev = dict(device="cuda:0")
    f = torch.tensor(faces.astype(np.int32).reshape(-1), dtype=torch.int32, **dev)
    e = torch.tensor(edges.astype(np.int32).reshape(-1), dtype=torch.int32, **dev)
    v = torch.tensor(surf.astype(np.int32), dtype=torch.int32, **dev)
    dd = torch.tensor(np.ascontiguousarray(d).reshape(-1), dtype=torch.float64, **dev)
    pt = torch.zeros(4 * cap[0], dtype=torch.int32, **dev)
    ee = torch.zeros(4 * cap[1], dtype=torch.int32, **dev)
    npt, nee = C.c_int(), C.c_int()
    sysm.ctx.call("sgipc_broad_phase_ccd", len(surf), _p(v), len(faces), _p(f), len(edges), _p(e), _p(sysm.x), _p(dd),
                  float(alpha), None, None, float(dhat), _p(pt), cap[0], C.byref(npt), _p(ee), cap[1], C.byref(nee))
    return pt[:4 * npt.value], ee[:4 * nee.value], dd, npt.value, nee.value


def test_ccd_candidates_and_step_bound(torch_cuda, oracle, twobox8):
    from broad_phase_ref import candidates
    from sgipc.system import _p
    sc, sysm, (faces, edges, surf) = twobox8
    print("case: CCD swept candidates + ACCD step bound", flush=True)
    rng = np.random.default_rng(7)
    d = rng.normal(size=(sc.n, 3)) * 0.002
    d[:, 1] += 0.004 * (np.arange(sc.n) >= sc.meta["n_lower"])   # the upper box moves down onto the lower
    pt, ee, dd, npt, nee = _broad_ccd(torch_cuda, sysm, faces, edges, surf, sc.dhat, d, 1.0)
    ptn = pt.cpu().numpy().reshape(-1, 4); een = ee.cpu().numpy().reshape(-1, 4)
    rpt, ree = candidates(sc.x, faces, edges, surf, sc.dhat, dir=d, alpha=1.0)
    assert set(map(tuple, ptn)) == set(map(tuple, rpt)) and set(map(tuple, een)) == set(map(tuple, ree))
    s0, _, _, _, _ = candidates(sc.x, faces, edges, surf, sc.dhat) + (None,) * 3
    assert npt > len(s0)                                          # sweeping adds candidates
    a = C.c_double()
    sysm.ctx.call("sgipc_ccd_step_bound", npt, _p(pt), nee, _p(ee), _p(sysm.x), _p(dd), 0.8, C.byref(a))
    ao = oracle.ccd_step_bound(ptn, een, sc.x, d, 0.8)
    assert 0 < a.value < 1 and abs(a.value - ao) <= 1e-12 * ao
    # conservative advancement keeps every candidate at >= eta^2 of its start
    # distance over the bounded step (eta = 1 - slackness = 0.2)
    from test_oracle_ground_adjoint import _pt_d, _seg_seg_d
    x0 = sc.x; x1 = sc.x - a.value * d
    for q in ptn[:: max(1, len(ptn) // 400)]:
        d0 = _pt_d(*x0[q])
        assert _pt_d(*x1[q]) >= 0.04 * d0 * (1 - 1e-9)
    for q in een[:: max(1, len(een) // 400)]:
        d0 = _seg_seg_d(*x0[q])
        assert _seg_seg_d(*x1[q]) >= 0.04 * d0 * (1 - 1e-9)
