"""GPU parity of the affine-body (ABD) slice against the oracle
(oracle/abd_ref.py), through the C ABI: body positions, body gradient and
Hessian triplets, the DOF-space mapping of the contact blocks (J^T H J),
the coupled system and its MAS + 12x12 block-Jacobi solve, the FP64
energy, fixed bodies, and a Newton solve of the hybrid scene."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _scene(fixed=None, kappa_abd=1e11, seed=11):
    import abd_ref as A
    from sgipc import scene
    hs = scene.hybrid_scene(cells=(8, 2, 8), lattice=2, cube=0.2, jitter=0.02, max_turn=0.2, dt=0.01,
                            kappa_abd=kappa_abd)
    rng = np.random.default_rng(seed)
    hs.q[:, 3:] += 5e-4 * rng.normal(size=(hs.nb, 9))        # off the rotation manifold
    hs.q[:, [0, 2]] += 1e-4 * rng.normal(size=(hs.nb, 2))
    hs.meta["positions"] = A.positions(hs.q, np.repeat(np.arange(hs.nb), 8), hs.xbar)
    if fixed is not None:
        hs.fixed[:] = fixed
    return hs


def _mat(v):
    """(m, 9) column-major blocks -> (m, 3, 3)"""
    return np.asarray(v).reshape(-1, 3, 3).transpose(0, 2, 1)


def _system(hs):
    from sgipc.hybrid import HybridSystem
    s = HybridSystem(hs, "cuda:0")
    n = s.detect()
    assert n > 0
    s.assemble()
    return s


def _body_of(hs):
    return np.repeat(np.arange(hs.nb), np.diff(hs.vstart))


def test_positions_and_body_terms(torch_cuda):
    import abd_ref as A
    print("case: abd positions / body terms, 8 bodies", flush=True)
    hs = _scene()
    s = _system(hs)
    x = s.x[:3 * s.nav].cpu().numpy().reshape(-1, 3)
    np.testing.assert_allclose(x, A.positions(hs.q, _body_of(hs), hs.xbar), rtol=0, atol=1e-15)
    G, H = A.body_system(hs.q, hs.q_tilde, hs.mass12, hs.volume, hs.fixed, hs.kappa_abd, hs.fem.dt)
    fb = s.fb.cpu().numpy()
    Gc = A.contact_gradient(fb[:3 * s.nav], _body_of(hs), hs.xbar, hs.nb, hs.fixed)
    bq = s.b[:12 * s.nb].cpu().numpy().reshape(-1, 12)
    scale = np.abs(G + Gc).max()
    np.testing.assert_allclose(bq, G + Gc, rtol=0, atol=1e-12 * scale)
    assert np.abs(Gc).max() > 0                     # contacts reach the bodies
    r, c, v = s.triplets()
    nb10 = 10 * s.nb
    br, bc, bv = A.body_triplets(H)
    np.testing.assert_array_equal(r[-nb10:], br)
    np.testing.assert_array_equal(c[-nb10:], bc)
    np.testing.assert_allclose(_mat(v[-nb10:]), bv, rtol=0, atol=1e-12 * np.abs(H).max())


def test_contact_mapping(torch_cuda):
    import abd_ref as A
    print("case: abd contact mapping", flush=True)
    hs = _scene()
    s = _system(hs)
    sr, sc, sv = s.source_triplets()
    R, Cc, V = A.map_contacts(sr, sc, _mat(sv), s.nav, s.nb, _body_of(hs), hs.xbar, hs.fixed)
    r, c, v = s.triplets()
    m = s.n_mapped
    assert 0 < m <= len(R)          # duplicate vertex pairs are merged before the mapping
    nbr = s.nbr
    want = A.dense_upper(R, Cc, V, nbr)
    got = A.dense_upper(r[:m], c[:m], _mat(v[:m]), nbr)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # the scene exercises every class: body-body, body-FEM and FEM-FEM blocks
    kinds = set()
    for a, b in zip(r[:m], c[:m]):
        kinds.add((a < 4 * s.nb, b < 4 * s.nb))
    assert (True, True) in kinds and (True, False) in kinds
    assert np.all(r[:m] <= c[:m])


def _dense(s):
    from conftest import dense_from_bcoo
    kr, kc, kv = s.bcoo()
    return dense_from_bcoo(kr, kc, kv, s.nbr)


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_coupled_solve(torch_cuda, precision, tol):
    import abd_ref as A
    from sgipc import capi
    print("case: abd coupled solve precision=%d" % precision, flush=True)
    hs = _scene()
    s = _system(hs)
    r, c, v = s.triplets()
    s.convert()
    Ad = _dense(s)
    # the converted system is the sum of its triplets: mapped contacts (oracle
    # map of the GPU barrier blocks), FEM + mass (GPU), bodies (oracle)
    sr, sc, sv = s.source_triplets()
    R, Cc, V = A.map_contacts(sr, sc, _mat(sv), s.nav, s.nb, _body_of(hs), hs.xbar, hs.fixed)
    G, H = A.body_system(hs.q, hs.q_tilde, hs.mass12, hs.volume, hs.fixed, hs.kappa_abd, hs.fem.dt)
    br, bc, bv = A.body_triplets(H)
    m, nf10 = s.n_mapped, 10 * s.T + s.nf
    want = (A.dense_upper(R, Cc, V, s.nbr) + A.dense_upper(r[m:m + nf10], c[m:m + nf10], _mat(v[m:m + nf10]), s.nbr)
            + A.dense_upper(br, bc, bv, s.nbr))
    np.testing.assert_allclose(Ad, want, rtol=0, atol=1e-12 * np.abs(want).max())
    s.build_precond()
    st = s.solve(tol=tol, stop_rule=capi.STOP_TRUE_REL, precision=precision, max_iter=20000)
    assert st.converged
    b = s.b.cpu().numpy()
    dx = s.dx.cpu().numpy()
    res = np.linalg.norm(Ad @ dx - b) / np.linalg.norm(b)
    assert res <= 1.01 * tol, res
    # the library's own residual agrees with the dense one
    assert st.rel_residual == pytest.approx(res, rel=1e-3, abs=1e-3 * tol)


def test_block_jacobi(torch_cuda):
    import abd_ref as A
    print("case: abd block-Jacobi apply", flush=True)
    hs = _scene()
    s = _system(hs)
    s.convert()
    s.build_precond()
    Ad = _dense(s)
    rng = np.random.default_rng(3)
    rv = rng.normal(size=3 * s.nbr)
    z = s.precond_apply(rv)
    # backward error of each body's 12x12 solve (the blocks are stiff:
    # kappa_abd = 1e11 gives condition numbers near 1e8), and the forward
    # error against the oracle's inverse within that conditioning
    inv = A.block_jacobi_inverse(Ad, s.nb)
    for b in range(s.nb):
        Bb = Ad[12 * b:12 * b + 12, 12 * b:12 * b + 12]
        zb, rb = z[12 * b:12 * b + 12], rv[12 * b:12 * b + 12]
        assert np.linalg.norm(Bb @ zb - rb) <= 1e-13 * np.linalg.norm(Bb, 2) * np.linalg.norm(zb)
        want = inv[b] @ rb
        assert np.linalg.norm(zb - want) <= 1e-14 * np.linalg.cond(Bb) * np.linalg.norm(want)
    # identity outside the bodies and the MAS range: none here (all FEM rows are MAS rows)


def test_energy_and_gradient_consistency(torch_cuda):
    import ctypes as C
    import abd_ref as A
    from sgipc.system import _p
    print("case: abd energy / directional derivative", flush=True)
    hs = _scene(kappa_abd=1e5)
    s = _system(hs)
    e = C.c_double()
    s.ctx.call("sgipc_abd_energy", _p(s.q), _p(s.q_tilde), hs.kappa_abd, hs.fem.dt, C.byref(e))
    want = A.energy(hs.q, hs.q_tilde, hs.mass12, hs.volume, hs.fixed, hs.kappa_abd, hs.fem.dt)
    assert e.value == pytest.approx(want, rel=1e-13)
    # b is the gradient of the whole incremental potential: central
    # differences of the FP64 energy along a random DOF direction
    rng = np.random.default_rng(5)
    import torch
    d = torch.tensor(rng.normal(size=3 * s.nbr), dtype=torch.float64, device="cuda:0")
    d[12 * s.nb:] *= torch.tensor(np.repeat(hs.fem.btype == 0, 3), dtype=torch.float64, device="cuda:0")
    d *= 1e-3 / float(d.abs().max())
    gd = float(torch.dot(s.b, d))
    q0, x0 = s.q.clone(), s.xf.clone()
    h = 1e-3

    def e_at(t):
        s.q.copy_(q0 - t * d[:12 * s.nb]); s.xf.copy_(x0 - t * d[12 * s.nb:]); s.update_body_positions()
        return s.energy()
    fd = (e_at(-h) - e_at(h)) / (2 * h)
    assert fd == pytest.approx(gd, rel=1e-5)


def test_fixed_body(torch_cuda):
    import abd_ref as A
    from sgipc import capi
    print("case: abd fixed body", flush=True)
    fixed = np.zeros(8, np.int32); fixed[0] = 1
    hs = _scene(fixed=fixed)
    s = _system(hs)
    r, c, v = s.triplets()
    m = s.n_mapped
    assert not np.any((r[:m] < 4) | (c[:m] < 4))        # no contact block touches body 0
    assert np.all(s.b[:12].cpu().numpy() == 0.0)
    nb10 = 10 * s.nb
    np.testing.assert_allclose(A.dense_upper(r[-nb10:-nb10 + 10], c[-nb10:-nb10 + 10], _mat(v[-nb10:-nb10 + 10]), 4),
                               hs.mass12[0].reshape(12, 12), rtol=0, atol=1e-14)
    s.convert(); s.build_precond()
    st = s.solve(tol=1e-10, stop_rule=capi.STOP_TRUE_REL)
    assert st.converged
    assert np.abs(s.dx[:12].cpu().numpy()).max() <= 1e-12 * np.abs(s.dx.cpu().numpy()).max()


def test_newton_solve_converges(torch_cuda):
    print("case: hybrid Newton solve", flush=True)
    hs = _scene(kappa_abd=1e11)
    from sgipc.hybrid import HybridSystem
    s = HybridSystem(hs, "cuda:0")
    hist = s.newton_solve(newton_tol=1e-9, max_newton=30, pcg_tol=1e-10)
    assert hist[-1][0] <= 1e-9, hist
    # converged: the bodies stay (nearly) rigid under kappa_abd = 1e11
    q = s.q.cpu().numpy().reshape(-1, 12)
    for b in range(s.nb):
        Am = q[b, 3:].reshape(3, 3)
        assert np.abs(Am @ Am.T - np.eye(3)).max() < 1e-6
