"""Pins of the affine-body (ABD) oracle (oracle/abd_ref.py) on the CPU:
closed forms of the shape energy, central differences tying energy,
gradient and Hessian together, and the DOF-space contact mapping against a
dense J^T H J built from the global Jacobian."""
import numpy as np
import pytest

import abd_ref as A


def _rot(rng):
    Q, R = np.linalg.qr(rng.normal(size=(3, 3)))
    return Q * np.sign(np.diag(R))


def test_shape_energy_closed_forms():
    rng = np.random.default_rng(0)
    for _ in range(5):
        R = _rot(rng)
        assert abs(A.shape_energy(R.reshape(9))) < 1e-24 + 1e-14
        assert np.abs(A.shape_gradient(R.reshape(9))).max() < 1e-12
    for s in (0.5, 1.0, 1.3):
        assert A.shape_energy((s * np.eye(3)).reshape(9)) == pytest.approx(3.0 * (s * s - 1.0) ** 2, rel=1e-14, abs=1e-30)
    # |A A^T - I|_F^2 for a general matrix
    M = rng.normal(size=(3, 3))
    assert A.shape_energy(M.reshape(9)) == pytest.approx(np.sum((M @ M.T - np.eye(3)) ** 2), rel=1e-12)


@pytest.mark.parametrize("seed", [1, 2])
def test_shape_derivatives_by_central_differences(seed):
    rng = np.random.default_rng(seed)
    a = (np.eye(3) + 0.3 * rng.normal(size=(3, 3))).reshape(9)
    g = A.shape_gradient(a)
    H = A.shape_hessian(a)
    h = 1e-6
    for k in range(9):
        e = np.zeros(9); e[k] = h
        fd = (A.shape_energy(a + e) - A.shape_energy(a - e)) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-7, abs=1e-8)
        fdg = (A.shape_gradient(a + e) - A.shape_gradient(a - e)) / (2 * h)
        np.testing.assert_allclose(fdg, H[:, k], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(H, H.T, atol=1e-13)


def test_make_psd():
    rng = np.random.default_rng(3)
    B = rng.normal(size=(9, 9))
    P = B @ B.T
    np.testing.assert_allclose(A.make_psd(P), P, rtol=1e-12, atol=1e-12)
    S = B + B.T
    Q = A.make_psd(S)
    w = np.linalg.eigvalsh(Q)
    assert w.min() > -1e-12
    # the projection keeps the nonnegative part of the spectrum
    ws, V = np.linalg.eigh(S)
    np.testing.assert_allclose(Q @ V[:, ws > 0], V[:, ws > 0] * ws[ws > 0], atol=1e-10)


def _bodies(rng, nb=3, per=8, fixed=None):
    xbar = rng.normal(size=(nb * per, 3)) * 0.1
    vstart = per * np.arange(nb + 1)
    mass = rng.uniform(0.5, 2.0, nb * per)
    M = A.dyadic_mass(xbar, mass, vstart)
    q = np.zeros((nb, 12))
    for b in range(nb):
        q[b, :3] = rng.normal(size=3)
        q[b, 3:] = (_rot(rng) + 0.05 * rng.normal(size=(3, 3))).reshape(9)
    qt = q + 0.01 * rng.normal(size=q.shape)
    vol = rng.uniform(0.5, 1.5, nb)
    fx = np.zeros(nb, np.int32) if fixed is None else np.asarray(fixed, np.int32)
    return xbar, vstart, mass, M, q, qt, vol, fx


def test_body_gradient_is_energy_derivative():
    rng = np.random.default_rng(4)
    xbar, vstart, mass, M, q, qt, vol, fx = _bodies(rng, fixed=[0, 1, 0])
    kappa, dt = 1e3, 0.01
    G, H = A.body_system(q, qt, M, vol, fx, kappa, dt)
    h = 1e-6
    for b in range(q.shape[0]):
        for k in range(12):
            e = np.zeros_like(q); e[b, k] = h
            fd = (A.energy(q + e, qt, M, vol, fx, kappa, dt) - A.energy(q - e, qt, M, vol, fx, kappa, dt)) / (2 * h)
            if fx[b]:
                assert G[b, k] == 0.0     # fixed bodies: zero gradient (energy keeps the shape term)
            else:
                assert fd == pytest.approx(G[b, k], rel=1e-6, abs=1e-7)
    # fixed body Hessian is its mass; free bodies: M + PSD part, symmetric, >= M
    np.testing.assert_array_equal(H[1], M[1])
    for b in (0, 2):
        np.testing.assert_allclose(H[b], H[b].T, atol=1e-12)
        assert np.linalg.eigvalsh(H[b] - M[b]).min() > -1e-9


def test_dyadic_mass_is_dense_jt_m_j():
    rng = np.random.default_rng(5)
    xbar, vstart, mass, M, q, qt, vol, fx = _bodies(rng)
    body_of = np.repeat(np.arange(3), 8)
    Jg = A.global_jacobian(24, 0, 3, body_of, xbar)
    Md = Jg.T @ np.diag(np.repeat(mass, 3)) @ Jg
    for b in range(3):
        np.testing.assert_allclose(Md[12 * b:12 * b + 12, 12 * b:12 * b + 12], M[b], rtol=1e-13, atol=1e-14)
    # J q reproduces the positions
    x = A.positions(q, body_of, xbar)
    np.testing.assert_allclose(x.reshape(-1), Jg @ q.reshape(-1), atol=1e-14)


def _vertex_hessian(rng, nv, npairs):
    """Random symmetric vertex-space blocks as the barrier stores them:
    upper (r <= c) with data, lower positions as zero fillers."""
    rows, cols, vals = [], [], []
    for _ in range(npairs):
        r, c = sorted(rng.integers(0, nv, 2))
        B = rng.normal(size=(3, 3))
        if r == c:
            B = B + B.T
        rows.append(r); cols.append(c); vals.append(B)
        rows.append(c); cols.append(r); vals.append(np.zeros((3, 3)))   # filler
    return np.array(rows), np.array(cols), np.array(vals)


@pytest.mark.parametrize("fixed", [None, [0, 1, 0]])
def test_mapping_exact_equals_dense_jthj(fixed):
    rng = np.random.default_rng(6)
    nb, per, nf = 3, 8, 10
    xbar, vstart, mass, M, q, qt, vol, fx = _bodies(rng, nb, per, fixed)
    nav = nb * per
    body_of = np.repeat(np.arange(nb), per)
    rows, cols, vals = _vertex_hessian(rng, nav + nf, 80)
    Hv = A.dense_upper(rows, cols, vals, nav + nf)
    if fixed is not None:   # blocks touching a fixed body are dropped
        for i in range(nav):
            if fx[body_of[i]]:
                Hv[3 * i:3 * i + 3, :] = 0.0
                Hv[:, 3 * i:3 * i + 3] = 0.0
    Jg = A.global_jacobian(nav, nf, nb, body_of, xbar)
    want = Jg.T @ Hv @ Jg
    R, Cc, V = A.map_contacts(rows, cols, vals, nav, nb, body_of, xbar, fx)
    got = A.dense_upper(R, Cc, V, 4 * nb + nf)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.all(R <= Cc)


def test_upper_storage_mapping_equals_reference_full_storage_procedure():
    """The reference maps full vertex-space storage ((i, j) and (j, i) for
    every pair, SymGH off); the library maps upper storage.  Same matrix."""
    rng = np.random.default_rng(7)
    nb, per, nf = 3, 8, 6
    xbar, vstart, mass, M, q, qt, vol, fx = _bodies(rng, nb, per, fixed=[0, 0, 1])
    nav = nb * per
    body_of = np.repeat(np.arange(nb), per)
    rows, cols, vals = _vertex_hessian(rng, nav + nf, 60)
    # full storage: each upper entry also as its transpose at (c, r)
    fr, fc, fv = [], [], []
    for r, c, H in zip(rows, cols, vals):
        if not np.any(H):
            continue
        fr.append(r); fc.append(c); fv.append(H)
        if r != c:
            fr.append(c); fc.append(r); fv.append(H.T)
    Rr, Cr, Vr = A.map_contacts_full_storage(np.array(fr), np.array(fc), np.array(fv), nav, nb, body_of, xbar, fx)
    Ru, Cu, Vu = A.map_contacts(rows, cols, vals, nav, nb, body_of, xbar, fx)
    nbr = 4 * nb + nf
    want = A.dense_upper(Rr, Cr, Vr, nbr)
    np.testing.assert_allclose(A.dense_upper(Ru, Cu, Vu, nbr), want, rtol=0, atol=1e-12 * np.abs(want).max())
    # and both are the dense J^T H J
    Hv = A.dense_upper(rows, cols, vals, nav + nf)
    for i in range(nav):
        if fx[body_of[i]]:
            Hv[3 * i:3 * i + 3, :] = 0.0
            Hv[:, 3 * i:3 * i + 3] = 0.0
    Jg = A.global_jacobian(nav, nf, nb, body_of, xbar)
    np.testing.assert_allclose(want, Jg.T @ Hv @ Jg, rtol=0, atol=1e-12 * np.abs(want).max())


def test_contact_gradient_is_dense_jt_g():
    rng = np.random.default_rng(8)
    nb, per = 3, 8
    xbar, vstart, mass, M, q, qt, vol, fx = _bodies(rng, nb, per, fixed=[0, 0, 1])
    nav = nb * per
    body_of = np.repeat(np.arange(nb), per)
    g = rng.normal(size=3 * nav)
    Jg = A.global_jacobian(nav, 0, nb, body_of, xbar)
    want = (Jg.T @ g).reshape(nb, 12)
    want[2] = 0.0
    np.testing.assert_allclose(A.contact_gradient(g, body_of, xbar, nb, fx), want, atol=1e-13)
