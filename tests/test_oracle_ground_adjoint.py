"""Pins of the oracle's ground-plane barrier, SNK1 exact-Hessian mode and
parameter adjoint, without a GPU: central finite differences of the oracle's
own energy and gradient, continuity at the barrier edge, and the step-bound
property."""
import numpy as np
import pytest

from conftest import dense_from_bcoo


def _plane(sc, gap):
    n = np.array([0.1, 1.0, 0.2]); n /= np.linalg.norm(n)
    off = float((sc.x @ n).min() - gap)
    return n, off


def _ground_grad(oracle, ids, x, n, off, dhat, kappa):
    gp = oracle.ground_detect(ids, x, n, off, dhat)
    fb, vals, rows, cols = oracle.ground_assemble(gp, x, n, off, dhat, kappa)
    return gp, fb, vals, rows


def test_ground_gradient_and_hessian_vs_fd(oracle, cube4):
    sc = cube4
    dhat, kappa = 0.05 ** 2, 1e3
    n, off = _plane(sc, 0.02)
    ids = np.arange(sc.n, dtype=np.uint32)
    x0 = sc.x.reshape(-1).copy()
    gp, fb, vals, rows = _ground_grad(oracle, ids, x0, n, off, dhat, kappa)
    assert 0 < len(gp) < sc.n
    assert np.array_equal(rows, gp.astype(np.int32))
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(4):
        v = rng.normal(size=3 * sc.n) * 1e-2
        # gradient: central FD of the energy over all vertices (energy filters s < dhat)
        ep = oracle.ground_energy(ids, x0 + h * v, n, off, dhat, kappa)
        em = oracle.ground_energy(ids, x0 - h * v, n, off, dhat, kappa)
        fd = (ep - em) / (2 * h)
        assert abs(fd - fb @ v) <= 1e-6 * max(abs(fb @ v), 1e-8)
        # Hessian: central FD of the gradient (same close set on both sides)
        _, fp, _, _ = _ground_grad(oracle, gp, x0 + h * v, n, off, dhat, kappa)
        _, fm, _, _ = _ground_grad(oracle, gp, x0 - h * v, n, off, dhat, kappa)
        Hv = np.zeros(3 * sc.n)
        for i, vid in enumerate(gp):
            B = vals[i].reshape(3, 3).T
            Hv[3 * vid:3 * vid + 3] += B @ v[3 * vid:3 * vid + 3]
        assert np.abs((fp - fm) / (2 * h) - Hv).max() <= 1e-5 * np.abs(Hv).max()


def test_ground_barrier_edge_and_sign(oracle):
    # one vertex above the plane y = 0: b, db, d2b vanish at s = dhat (C2
    # edge), the force pushes away from the plane, the Hessian is n n^T >= 0
    dhat, kappa = 1e-4, 1.0
    n = np.array([0.0, 1.0, 0.0])
    ids = np.array([0], np.uint32)
    for d in (0.2e-2, 0.5e-2, 0.9e-2):
        x = np.array([0.3, d, -0.1])
        fb, vals, _, _ = oracle.ground_assemble(ids, x, n, 0.0, dhat, kappa)
        assert fb[1] < 0 and fb[0] == fb[2] == 0          # gradient points into the plane (force away)
        assert vals[0, 4] > 0 and np.count_nonzero(vals[0]) == 1
    x = np.array([0.0, np.sqrt(dhat) * (1 - 1e-9), 0.0])
    fb, vals, _, _ = oracle.ground_assemble(ids, x, n, 0.0, dhat, kappa)
    assert abs(fb[1]) < 1e-20 and abs(vals[0, 4]) < 1e-7
    assert oracle.ground_energy(ids, np.array([0.0, 0.011, 0.0]), n, 0.0, dhat, kappa) == 0.0
    # closed form at s = dhat / 4: -(s - dhat)^2 ln(1/4)
    s = dhat / 4
    e = oracle.ground_energy(ids, np.array([0.0, np.sqrt(s), 0.0]), n, 0.0, dhat, 2.0)
    assert abs(e - 2.0 * (-(s - dhat) ** 2 * np.log(0.25))) <= 1e-15


def test_ground_step_bound_property(oracle, cube4):
    sc = cube4
    n, off = _plane(sc, 0.02)
    ids = np.arange(sc.n, dtype=np.uint32)
    x0 = sc.x.reshape(-1)
    rng = np.random.default_rng(9)
    slack = 0.8
    d0 = (x0.reshape(-1, 3) @ n) - off
    for scale in (1e-3, 0.05, 1.0):
        dirv = rng.normal(size=3 * sc.n) * scale
        a = oracle.ground_step_bound(ids, x0, dirv, n, off, slack)
        assert 0 < a <= 1
        d1 = ((x0 - a * dirv).reshape(-1, 3) @ n) - off
        assert (d1 >= (1 - slack) * d0 - 1e-15).all()       # nobody crosses the slack margin
        if a < 1:                                              # and the bound is tight
            assert np.isclose(d1 / d0, 1 - slack).any()


def test_snk1_exact_mode_hessian_vs_fd(oracle, cube4):
    sc = cube4
    # a strongly deformed state so that projection clamps some modes
    rng = np.random.default_rng(4)
    x = sc.x_rest * np.array([1.0, 0.55, 1.3]) + rng.normal(size=sc.x.shape) * 0.03
    h = 1e-6
    with oracle.fem_projection(False):
        g0, V, R, Cc = oracle.fem_assemble(x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt)
        kr, kc, kv = oracle.convert(R, Cc, V)
    A = dense_from_bcoo(kr, kc, kv, sc.n)
    gp0, Vp, Rp, Cp = oracle.fem_assemble(x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt)
    Ap = dense_from_bcoo(*oracle.convert(Rp, Cp, Vp), sc.n)
    assert np.abs(A - Ap).max() > 1e-6 * np.abs(A).max()      # projection is active here
    assert np.linalg.eigvalsh(Ap).min() >= -1e-9 * np.abs(Ap).max()
    for _ in range(3):
        v = rng.normal(size=3 * sc.n)
        gp_ = oracle.fem_assemble(x + h * v.reshape(-1, 3), sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt)[0]
        gm_ = oracle.fem_assemble(x - h * v.reshape(-1, 3), sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt)[0]
        fd = (gp_ - gm_).reshape(-1) / (2 * h)
        assert np.abs(fd - A @ v).max() <= 1e-6 * np.abs(A @ v).max()


def test_fem_param_adjoint_vs_fd(oracle, cube4):
    sc = cube4
    rng = np.random.default_rng(8)
    x = sc.x_rest * np.array([1.05, 0.9, 1.0]) + rng.normal(size=sc.x.shape) * 0.01
    adj = rng.normal(size=3 * sc.n)
    btype = sc.btype.astype(np.int32)
    dmu, dlam = oracle.fem_param_adjoint(x, sc.tets, sc.DmInv, sc.vol, sc.dt, adj, btype)
    free = np.repeat(btype == 0, 3)

    def objective(mu, lam):
        g = oracle.fem_assemble(x, sc.tets, sc.DmInv, sc.vol, mu, lam, sc.dt)[0].reshape(-1)
        return (adj * free) @ g
    for t in (0, 17, sc.T - 1):
        for which, ref in (("mu", dmu), ("lam", dlam)):
            mu, lam = sc.mu.copy(), sc.lam.copy()
            arr = mu if which == "mu" else lam
            h = 1e-4 * arr[t]
            arr[t] += h; fp = objective(mu, lam)
            arr[t] -= 2 * h; fm = objective(mu, lam)
            fd = (fp - fm) / (2 * h)
            assert abs(fd - ref[t]) <= 1e-7 * max(abs(ref[t]), 1e-3 * np.abs(ref).max())
    # the gradient is linear in (mu, lam): adj . g = sum_t mu_t dmu_t + lam_t dlam_t
    assert np.isclose(objective(sc.mu, sc.lam), sc.mu @ dmu + sc.lam @ dlam, rtol=1e-10)


def test_reference_rollout_adjoint_vs_fd(oracle):
    """The CPU reference rollout gradient (dense solves over the oracle)
    against central differences of its own loss, on a 4x2x2-cell cantilever
    that reaches the ground."""
    from sgipc import scene
    from rollout_ref import RolloutRef
    sc = scene.cantilever_scene(4, 2, 2, seed=4, clearance=0.01)
    ref = RolloutRef(sc)
    m = sc.meta
    steps = 6
    L, gE, gv = ref.gradient(m["log_E"], m["v0"], m["target"], steps)
    states = ref.rollout(m["log_E"], m["v0"], steps)
    assert (states[-1][0].reshape(-1, 3)[:, 1] < sc.x_rest[:, 1].min()).any()     # it fell
    gp = oracle.ground_detect(ref.surf, states[-1][0], ref.nrm, ref.off, ref.dhat)
    assert len(gp) > 0                                                            # and touches the ground
    rng = np.random.default_rng(0)
    free = (sc.btype == 0)[:, None].astype(float)
    h = 1e-6
    for _ in range(3):
        dE = rng.normal(size=sc.T); dv = rng.normal(size=(sc.n, 3)) * free
        nrm = np.sqrt(dE @ dE + np.sum(dv * dv)); dE /= nrm; dv /= nrm

        def loss(s):
            st = ref.rollout(m["log_E"] + s * dE, m["v0"] + s * dv, steps)
            return np.sum((st[-1][0] - m["target"].reshape(-1)) ** 2)
        fd = (loss(h) - loss(-h)) / (2 * h)
        ad = gE @ dE + np.sum(gv * dv)
        assert abs(fd - ad) <= 1e-5 * abs(ad)


# ---- additive CCD ------------------------------------------------------------
def _seg_seg_d(a0, a1, b0, b1):
    """exact distance between two segments (closest points, clamped)"""
    u, v, w = a1 - a0, b1 - b0, a0 - b0
    a, b, c, d, e = u @ u, u @ v, v @ v, u @ w, v @ w
    cands = []
    D = a * c - b * b
    if D > 1e-300:
        s = (b * e - c * d) / D; t = (a * e - b * d) / D
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append(np.linalg.norm(w + s * u - t * v))
    for p, q0, q1 in ((a0, b0, b1), (a1, b0, b1), (b0, a0, a1), (b1, a0, a1)):
        cands.append(_p_seg_d(p, q0, q1))
    return min(cands)


def _p_seg_d(p, q0, q1):
    r = q1 - q0
    t = np.clip((p - q0) @ r / (r @ r), 0, 1)
    return np.linalg.norm(p - q0 - t * r)


def _pt_d(p, t0, t1, t2):
    """exact point-triangle distance"""
    n = np.cross(t1 - t0, t2 - t0); n2 = n @ n
    q = p - ((p - t0) @ n) / n2 * n
    bc = [np.cross(t2 - t1, q - t1) @ n, np.cross(t0 - t2, q - t2) @ n, np.cross(t1 - t0, q - t0) @ n]
    if min(bc) >= 0:
        return abs((p - t0) @ n) / np.sqrt(n2)
    return min(_p_seg_d(p, a, b) for a, b in ((t0, t1), (t1, t2), (t2, t0)))


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_accd_head_on_is_one_minus_eta_of_impact(oracle):
    rng = np.random.default_rng(21)
    for trial in range(20):
        R, off = _rot(rng), rng.normal(size=3)
        eta = rng.uniform(0.1, 0.5)
        d0, speed = rng.uniform(0.01, 0.5), rng.uniform(0.2, 3.0)   # impact at d0 / speed
        tstar = d0 / speed
        # point above a triangle's interior, moving straight at it; the
        # triangle drifts sideways with it (no change in the gap)
        tri = np.array([[-1.0, -1, 0], [1, -0.5, 0], [0, 1, 0]])
        p = np.array([0.05, 0.1, d0])
        drift = np.array([rng.normal(), rng.normal(), 0.0]) * 0.1
        v = np.concatenate([p, *tri]) .reshape(4, 3) @ R.T + off
        u = np.stack([drift + [0, 0, -speed]] + [drift] * 3) @ R.T
        t = oracle.accd(True, v, u, eta)
        if tstar < 1:
            assert abs(t - (1 - eta) * tstar) <= 1e-9 * tstar
        else:
            assert t == 1.0
        # crossing edges approaching head-on
        a = np.array([[-1.0, 0, d0], [1, 0, d0]]); b = np.array([[0.0, -1, 0], [0, 1, 0]])
        v = np.concatenate([a, b]) @ R.T + off
        u = np.stack([[0, 0, -0.5 * speed]] * 2 + [[0, 0, 0.5 * speed]] * 2) @ R.T
        t = oracle.accd(False, v, u, eta)
        if tstar < 1:
            assert abs(t - (1 - eta) * tstar) <= 1e-9 * tstar
        else:
            assert t == 1.0


def test_accd_never_passes_contact(oracle):
    """motions built to touch at a chosen time t_c (the point reaches a point
    of the moving triangle; the edges cross): the returned time stays below
    t_c, and the exact distance stays positive on [0, toc]"""
    rng = np.random.default_rng(5)
    for trial in range(120):
        pt = trial % 2 == 0
        tc = rng.uniform(0.2, 0.9)
        if pt:
            tri0 = rng.normal(size=(3, 3)); ut = rng.normal(size=(3, 3)) * 0.5
            w = rng.dirichlet([1.0, 1.0, 1.0])
            c = w @ (tri0 + tc * ut)
            p0 = c + rng.normal(size=3)
            v = np.vstack([p0, tri0]); u = np.vstack([(c - p0) / tc, ut])
            dist = lambda y: _pt_d(*y)
        else:
            b0 = rng.normal(size=(2, 3)); ub = rng.normal(size=(2, 3)) * 0.5
            r, sa = rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)
            bt = b0 + tc * ub
            c = bt[0] + r * (bt[1] - bt[0])
            a0 = rng.normal(size=3); ua0 = rng.normal(size=3) * 0.5
            a0t = a0 + tc * ua0
            a1t = a0t + (c - a0t) / sa
            a1 = a1t - rng.normal(size=3) * 0.5 * tc
            v = np.vstack([a0, a1, b0]); u = np.vstack([ua0, (a1t - a1) / tc, ub])
            dist = lambda y: _seg_seg_d(*y)
        d0 = dist(v)
        if d0 < 1e-2:
            continue
        assert dist(v + tc * u) < 1e-9                       # the construction touches at t_c
        toc = oracle.accd(pt, v, u, 0.2)
        assert 0 < toc < tc
        for t in np.linspace(0.0, toc, 401):
            assert dist(v + t * u) > 0
