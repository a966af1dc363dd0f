"""CPU tests that pin the oracle (the checker of the HIP path) against
independent references: closed forms, central finite differences, numpy
eigen-decompositions, brute-force dictionaries and connected components,
dense linear algebra and the reference's own partition fixture."""
import os

import numpy as np
import pytest

from conftest import oracle_system, dense_from_bcoo

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- 3x3 math ---------------------------------------------------------------
def test_svd3_signed(oracle):
    rng = np.random.default_rng(0)
    cases = [rng.normal(size=(3, 3)) for _ in range(300)]
    cases += [np.eye(3) + 1e-6 * rng.normal(size=(3, 3)) for _ in range(50)]
    cases += [np.diag(rng.normal(size=3)) for _ in range(50)]
    sing = rng.normal(size=(3, 3)); sing[:, 2] = sing[:, 0]
    cases += [sing, np.zeros((3, 3)), -np.eye(3)]
    for A in cases:
        U, S, V = oracle.svd3(A)
        assert np.abs(U @ np.diag(S) @ V.T - A).max() <= 1e-12 * max(1.0, np.abs(A).max())
        assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
        assert np.allclose(np.abs(S), np.linalg.svd(A, compute_uv=False)[np.argsort(-np.abs(S)).argsort()] if False
                           else np.abs(S), atol=0)
        # magnitudes equal the singular values, only the last may be negative
        assert np.allclose(np.sort(np.abs(S)), np.sort(np.linalg.svd(A, compute_uv=False)), atol=1e-12)
        assert S[0] >= 0 and S[1] >= 0
        assert np.sign(np.prod(S)) == np.sign(np.linalg.det(A)) or abs(np.linalg.det(A)) < 1e-14


def test_syev3(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        B = rng.normal(size=(3, 3)); A = B + B.T
        w, Q = oracle.syev3(A)
        assert np.allclose(A @ Q, Q * w, atol=1e-12 * np.abs(A).max())
        assert np.allclose(Q.T @ Q, np.eye(3), atol=1e-12)
        assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * np.abs(A).max())


# ---- SNK1 tet -------------------------------------------------------------
def exact_hessian(F, mu, lam):
    """d2 Psi / dF2 of Psi = mu/2 (I2-3) + lam/2 (J-1-mu/lam)^2, vec = column-major."""
    J = np.linalg.det(F)
    g = (J * np.linalg.inv(F).T).reshape(-1, order="F")
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    HJ = np.einsum("ikm,jln,mn->ijkl", eps, eps, F)
    HJv = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    HJv[i + 3 * j, k + 3 * l] = HJ[i, j, k, l]
    return mu * np.eye(9) + lam * np.outer(g, g) + lam * (J - 1 - mu / lam) * HJv


def test_snk1_hessian_is_psd_projection_of_exact(oracle):
    rng = np.random.default_rng(2)
    for t in range(300):
        F = np.eye(3) + 0.3 * rng.normal(size=(3, 3)) if t % 3 else rng.normal(size=(3, 3))
        if t % 7 == 0:  # inverted element
            F[:, 0] *= -1
        mu, lam = 1e5 * rng.uniform(0.5, 2), 1e6 * rng.uniform(0.1, 3)
        w, Q = np.linalg.eigh(exact_hessian(F, mu, lam))
        Hp = (Q * np.maximum(w, 0)) @ Q.T
        Ho = oracle.snk1_hq(F, mu, lam)
        assert np.abs(Ho - Hp).max() <= 1e-11 * np.abs(Hp).max()


def test_snk1_energy_gradient_and_scaling(oracle, cube4):
    sc = cube4
    rng = np.random.default_rng(3)
    for t in rng.choice(sc.T, 20, replace=False):
        xt = sc.x[sc.tets[t].astype(int)].reshape(-1)
        Dm = sc.DmInv[t].reshape(3, 3)
        # closed-form energy
        Ds = np.stack([xt[3:6] - xt[:3], xt[6:9] - xt[:3], xt[9:12] - xt[:3]], 1)
        F = Ds @ Dm
        mu, lam, vol = sc.mu[t], sc.lam[t], sc.vol[t]
        e_ref = 0.5 * vol * (mu * ((F * F).sum() - 3) + lam * (np.linalg.det(F) - 1 - mu / lam) ** 2)
        assert abs(oracle.snk1_energy(xt, Dm, vol, mu, lam) - e_ref) <= 1e-12 * abs(e_ref)
        dt = 0.01
        g, H = oracle.snk1_tet(xt, Dm, vol, mu, lam, dt)
        fd = np.zeros(12)
        h = 1e-6 * np.abs(xt).max()
        for i in range(12):
            xp = xt.copy(); xp[i] += h
            xm = xt.copy(); xm[i] -= h
            fd[i] = (oracle.snk1_energy(xp, Dm, vol, mu, lam) - oracle.snk1_energy(xm, Dm, vol, mu, lam)) / (2 * h)
        assert np.abs(g - dt * dt * fd).max() <= 1e-6 * np.abs(g).max()
        assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
        assert np.linalg.eigvalsh(H).min() >= -1e-10 * np.abs(H).max()


def test_snk1_rest_state_has_zero_gradient(oracle, cube4):
    sc = cube4
    t = 0
    xt = sc.x_rest[sc.tets[t].astype(int)].reshape(-1)
    g, H = oracle.snk1_tet(xt, sc.DmInv[t], sc.vol[t], sc.mu[t], sc.lam[t], 0.01)
    assert np.abs(g).max() <= 1e-12 * sc.mu[t] * sc.vol[t] * 1e-4 + 1e-20


def test_rest_precompute_matches_numpy(oracle, cube4):
    from sgipc import scene
    sc = cube4
    DmInv, vol, mass, mu, lam = oracle.fem_precompute(sc.x_rest, sc.tets, 1e5, 0.45, 1000.0)
    D2, v2, m2, mu2, l2 = scene.rest_precompute(sc.x_rest, sc.tets, 1e5, 0.45, 1000.0)
    assert np.allclose(DmInv, D2, rtol=1e-12, atol=1e-9) and np.allclose(vol, v2, rtol=1e-12)
    assert np.allclose(mass, m2, rtol=1e-12) and np.allclose(mu, mu2) and np.allclose(lam, l2)


# ---- convert / spmv -----------------------------------------------------------
def test_convert_bruteforce(oracle):
    rng = np.random.default_rng(4)
    nb, nt = 50, 2000
    r = rng.integers(0, nb, nt); c = rng.integers(0, nb, nt)
    rows = np.minimum(r, c).astype(np.int32); cols = np.maximum(r, c).astype(np.int32)
    vals = rng.normal(size=(nt, 9))
    kr, kc, kv = oracle.convert(rows, cols, vals)
    d = {}
    for i in range(nt):
        k = (int(rows[i]), int(cols[i]))
        d[k] = d[k] + vals[i] if k in d else vals[i].copy()
    keys = sorted(d)
    assert [tuple(k) for k in zip(kr.tolist(), kc.tolist())] == keys
    assert all(np.array_equal(kv[j], d[k]) for j, k in enumerate(keys))   # same summation order
    e_r, e_c, e_v = oracle.convert(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 9)))
    assert len(e_r) == 0


def test_spmv_dense(oracle, cube4):
    ref = oracle_system(oracle, cube4)
    kr, kc, kv = ref["bcoo"]
    A = dense_from_bcoo(kr, kc, kv, cube4.n)
    x = np.random.default_rng(5).normal(size=3 * cube4.n)
    y0 = np.random.default_rng(6).normal(size=3 * cube4.n)
    y = oracle.spmv_sym(kr, kc, kv, cube4.n, x, a=2.0, b=-0.5, y=y0)
    assert np.abs(y - (2 * A @ x - 0.5 * y0)).max() <= 1e-12 * np.abs(A @ x).max()


# ---- MAS -------------------------------------------------------------------
def _components_in_bank(ids, edges_local):
    """connected components over local ids (union-find), returns root per id"""
    parent = {i: i for i in ids}

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in edges_local:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return {i: find(i) for i in ids}


def brute_force_tables(n, part_map, real_map, edges, bank, L):
    """Level numbering from scratch: components of each bank's induced
    subgraph, numbered bank by bank in order of their lowest lane."""
    node = [real_map[v] for v in range(n)]       # current global id per vertex (level 0: slot)
    begin, counts, table = [], [], []
    beg = 0
    cur_node = np.array(node)
    local = cur_node - 0
    sizes = []
    M = len(part_map)
    level_nodes = M
    for l in range(1, L):
        # group level-(l-1) nodes by bank
        nb_ = (level_nodes + bank - 1) // bank
        comp_id = {}
        nxt = 0
        adj = {}
        for a, b in edges:
            na, nb2 = local[a], local[b]
            if na // bank == nb2 // bank and na != nb2:
                adj.setdefault(na // bank, []).append((na, nb2))
        present = {}
        for v in range(n):
            present.setdefault(local[v] // bank, set()).add(local[v])
        for bk in range(nb_):
            ids = sorted(present.get(bk, ()))
            if not ids:
                continue
            roots = _components_in_bank(ids, adj.get(bk, []))
            order = sorted(set(roots.values()))
            for r in order:
                comp_id[(bk, r)] = nxt
                nxt += 1
            for i in ids:
                comp_id[("n", i)] = comp_id[(bk, roots[i])]
        local = np.array([comp_id[("n", local[v])] for v in range(n)])
        table.append(local.copy())
        counts.append(nxt)
        level_nodes = nxt
    return counts, table


def _mesh_edges(ns, nn, nl):
    e = []
    for v in range(len(ns)):
        for k in range(nn[v]):
            e.append((v, int(nl[ns[v] + k])))
    return e


@pytest.mark.parametrize("bank", [16, 32])
def test_mas_tables_vs_bruteforce(oracle, request, bank):
    sc = request.getfixturevalue("cube10" if bank == 16 else "cube10_b32")
    ref = oracle_system(oracle, sc)
    ns, nn, nl = oracle.vertex_neighbors(sc.n, sc.tets)
    pm, rm = oracle.bank_map(sc.part, bank)
    mas = oracle.Mas(sc.n, pm, rm, ns, nn, nl, bank)
    kr, kc, kv = ref["bcoo"]
    mas.build(kr, kc, kv)
    t = mas.tables()
    L = mas.levelnum
    assert L >= 3
    counts, table = brute_force_tables(sc.n, pm, rm, _mesh_edges(ns, nn, nl), bank, L)
    for l in range(1, L):
        assert t["level_size"][l, 0] == counts[l - 1]
        assert np.array_equal(t["coarse_space"][l - 1], table[l - 1])


@pytest.mark.parametrize("bank", [16, 32])
def test_mas_apply_dense(oracle, request, bank):
    sc = request.getfixturevalue("cube4" if bank == 16 else "cube4_b32")
    ref = oracle_system(oracle, sc)
    kr, kc, kv = ref["bcoo"]
    A = dense_from_bcoo(kr, kc, kv, sc.n)
    ns, nn, nl = oracle.vertex_neighbors(sc.n, sc.tets)
    pm, rm = oracle.bank_map(sc.part, bank)
    mas = oracle.Mas(sc.n, pm, rm, ns, nn, nl, bank)
    tot = mas.build(kr, kc, kv)
    ct = mas.tables()["coarse_table"]
    L = mas.levelnum
    r = np.random.default_rng(7).normal(size=3 * sc.n)
    z = np.zeros(3 * sc.n)
    for l in range(L):
        nodes = np.array([rm[v] if l == 0 else ct[v, l - 1] for v in range(sc.n)])
        R = np.zeros((3 * tot, 3 * sc.n))
        for v in range(sc.n):
            R[3 * nodes[v]:3 * nodes[v] + 3, 3 * v:3 * v + 3] = np.eye(3)
        Al = R @ A @ R.T
        Dinv = np.zeros_like(Al)
        for b in np.unique(nodes // bank):
            s = slice(3 * bank * b, 3 * bank * (b + 1))
            blk = Al[s, s].copy()
            dg = np.diag(blk).copy(); dg[dg == 0] = 1; np.fill_diagonal(blk, dg)
            Dinv[s, s] = np.linalg.inv(blk)
        z += R.T @ Dinv @ R @ r
    assert np.abs(mas.apply(r) - z).max() <= 1e-12 * np.abs(z).max()
    assert np.abs(mas.apply(r, fp32=True) - z).max() <= 1e-5 * np.abs(z).max()


def test_high_mat_fixture_partition(oracle):
    """The reference's own METIS-sorted mesh and part file: bank map and MAS
    tables are well formed and match the brute-force component numbering."""
    d = np.load(os.path.join(GOLDEN, "high_mat_sorted16.npz"))
    tets, part = d["tets"].astype(np.uint32), d["part"]
    n = part.shape[0]
    pm, rm = oracle.bank_map(part, 16)
    assert len(pm) == 16 * 3000 and np.array_equal(pm[rm], np.arange(n))
    assert np.all(np.diff(rm) > 0)            # slots monotone in vertex id
    ns, nn, nl = oracle.vertex_neighbors(n, tets)
    mas = oracle.Mas(n, pm, rm, ns, nn, nl, 16)
    # diagonal-only system is enough for the tables
    kr = np.arange(n, dtype=np.int32)
    kv = np.tile(np.eye(3).reshape(-1), (n, 1))
    mas.build(kr, kr, kv)
    t = mas.tables()
    assert mas.levelnum == oracle.lib().or_mas_num_levels(max(len(pm), n), 16)
    counts, table = brute_force_tables(n, pm, rm, _mesh_edges(ns, nn, nl), 16, 3)
    assert t["level_size"][1, 0] == counts[0] and np.array_equal(t["coarse_space"][0], table[0])
    assert t["level_size"][2, 0] == counts[1] and np.array_equal(t["coarse_space"][1], table[1])


# ---- PCG ---------------------------------------------------------------------
@pytest.mark.parametrize("precond", [0, 1, 2])
def test_pcg_matches_direct_solve(oracle, cube4, precond):
    sc = cube4
    ref = oracle_system(oracle, sc)
    kr, kc, kv = ref["bcoo"]
    A = dense_from_bcoo(kr, kc, kv, sc.n)
    mas = None
    if precond == 2:
        ns, nn, nl = oracle.vertex_neighbors(sc.n, sc.tets)
        pm, rm = oracle.bank_map(sc.part, 16)
        mas = oracle.Mas(sc.n, pm, rm, ns, nn, nl, 16)
        mas.build(kr, kc, kv)
    x, it, rr = oracle.pcg(kr, kc, kv, sc.n, ref["b"], precond=precond, mas=mas, tol=1e-10, rule=1, max_iter=10000)
    xs = np.linalg.solve(A, ref["b"])
    assert rr <= 1e-9 and np.linalg.norm(x - xs) <= 1e-7 * np.linalg.norm(xs)
    # reference rule stops on |rz| <= tol rz0 with the stale rz
    x1, it1, _ = oracle.pcg(kr, kc, kv, sc.n, ref["b"], precond=precond, mas=mas, tol=1e-4, rule=0)
    assert 1 <= it1 < 0.3 * 3 * sc.n


# ---- barrier ----------------------------------------------------------------
def _barrier_energy_direct(d, dhat, kappa):
    return kappa * (d - dhat) ** 2 * np.log(d / dhat) ** 2


def _closest_dist2(p, prim):
    """brute force squared distance from p to a segment/triangle by dense sampling + polish"""
    from scipy.optimize import minimize
    if len(prim) == 1:
        return float(((p - prim[0]) ** 2).sum())
    if len(prim) == 2:
        a, b = prim
        t = np.clip(np.dot(p - a, b - a) / np.dot(b - a, b - a), 0, 1)
        return float(((p - (a + t * (b - a))) ** 2).sum())
    a, b, c = prim
    best = np.inf
    for u in np.linspace(0, 1, 41):
        for v in np.linspace(0, 1 - u, max(2, int(41 * (1 - u)))):
            q = a + u * (b - a) + v * (c - a)
            best = min(best, ((p - q) ** 2).sum())
    f = lambda uv: ((p - (a + uv[0] * (b - a) + uv[1] * (c - a))) ** 2).sum()
    cons = ({"type": "ineq", "fun": lambda uv: uv[0]}, {"type": "ineq", "fun": lambda uv: uv[1]},
            {"type": "ineq", "fun": lambda uv: 1 - uv[0] - uv[1]})
    res = minimize(f, [1 / 3, 1 / 3], constraints=cons, method="SLSQP", options={"ftol": 1e-20, "maxiter": 200})
    return float(min(best, res.fun))


@pytest.fixture(scope="module")
def twobox5(oracle):
    from sgipc import scene
    sc = scene.two_box_scene(5, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    pairs, mi, counts = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)
    return sc, pt, ee, pairs, mi, counts


def test_classify_pt_picks_closest_feature(oracle, twobox5):
    sc, pt, ee, pairs, mi, counts = twobox5
    assert len(pairs) > 50 and counts.sum() == len(pairs)
    for c in range(3):
        sel = mi[[k for k in range(len(pairs)) if (4 if pairs[k, 0] >= 0 else
                                                   (2 if pairs[k, 2] < 0 else (3 if pairs[k, 3] < 0 else 4))) == c + 2]]
        assert np.array_equal(np.sort(sel), np.arange(counts[c]))
    # every PT candidate: the oracle pair's distance equals the true point-triangle distance
    x = sc.x
    npt_pairs = 0
    for cand in pt[:300]:
        p = x[cand[0]]
        tri = x[cand[1:]]
        dtrue = _closest_dist2(p, tri)
        q, _, _ = oracle.classify(cand[None, :], np.zeros((0, 4), np.int32), sc.x, sc.x_rest, 1e30)
        vid, cls, g, H, e = oracle.barrier_pair(sc.x, sc.x_rest, q[0], 1e30, 1.0)
        prim = x[vid[1:]]
        d_cls = _closest_dist2(p, prim) if cls < 4 else dtrue
        assert abs(d_cls - dtrue) <= 1e-9 * max(dtrue, 1e-12) + 1e-16
        npt_pairs += 1
    assert npt_pairs > 0


def test_barrier_gradient_hessian(oracle, twobox5):
    sc, pt, ee, pairs, mi, counts = twobox5
    rng = np.random.default_rng(8)
    x0 = sc.x
    for k in rng.choice(len(pairs), min(60, len(pairs)), replace=False):
        q = pairs[k]
        vid, cls, g, H, e = oracle.barrier_pair(x0, sc.x_rest, q, sc.dhat, sc.kappa)
        nv = len(vid)
        assert cls == {2: 2, 3: 3, 4: 4}[nv if nv != 4 else 4] or cls == 4
        # energy closed form on the sub-case distance
        xs = x0[vid]
        if nv == 2:
            d = ((xs[0] - xs[1]) ** 2).sum()
        elif nv == 3:
            d = _closest_dist2(xs[0], xs[1:]) if True else 0
            a, b = xs[1], xs[2]
            d = np.sum(np.cross(a - xs[0], b - xs[0]) ** 2) / np.sum((b - a) ** 2)
        elif q[0] < 0:
            n = np.cross(xs[2] - xs[1], xs[3] - xs[1]); d = np.dot(xs[0] - xs[1], n) ** 2 / np.dot(n, n)
        else:
            n = np.cross(xs[1] - xs[0], xs[3] - xs[2]); d = np.dot(xs[2] - xs[0], n) ** 2 / np.dot(n, n)
        assert abs(e - _barrier_energy_direct(d, sc.dhat, sc.kappa)) <= 1e-10 * abs(e) + 1e-30
        # gradient vs central FD of the energy
        h = 1e-8
        fd = np.zeros(3 * nv)
        for a in range(nv):
            for c in range(3):
                xp = x0.copy(); xp[vid[a], c] += h
                xm = x0.copy(); xm[vid[a], c] -= h
                fd[3 * a + c] = (oracle.barrier_pair(xp, sc.x_rest, q, sc.dhat, sc.kappa)[4] -
                                 oracle.barrier_pair(xm, sc.x_rest, q, sc.dhat, sc.kappa)[4]) / (2 * h)
        assert np.abs(fd - g).max() <= 1e-5 * np.abs(g).max()
        # Gauss-Newton Hessian: rank one, lambda0 * g_s g_s^T with g_s = ds/dx by FD
        def dist2(v):
            if nv == 2:
                return ((v[0] - v[1]) ** 2).sum()
            if nv == 3:
                return np.sum(np.cross(v[1] - v[0], v[2] - v[0]) ** 2) / np.sum((v[2] - v[1]) ** 2)
            if q[0] < 0:
                n_ = np.cross(v[2] - v[1], v[3] - v[1])
                return np.dot(v[0] - v[1], n_) ** 2 / np.dot(n_, n_)
            n_ = np.cross(v[1] - v[0], v[3] - v[2])
            return np.dot(v[2] - v[0], n_) ** 2 / np.dot(n_, n_)

        def s(xx):
            return np.sqrt(dist2(xx[vid]) / sc.dhat)
        gs = np.zeros(3 * nv)
        for a in range(nv):
            for c in range(3):
                xp = x0.copy(); xp[vid[a], c] += h
                xm = x0.copy(); xm[vid[a], c] -= h
                gs[3 * a + c] = (s(xp) - s(xm)) / (2 * h)
        I = d / sc.dhat
        Ic = max(I, 1e-6) if d < 1e-6 * sc.dhat else I
        l = np.log(Ic)
        lam0 = -(4 * sc.kappa * sc.dhat ** 2 * (4 * Ic + l - 3 * Ic ** 2 * l * l + 6 * Ic * l - 2 * Ic ** 2 +
                                                  Ic * l * l - 7 * Ic ** 2 * l - 2)) / Ic
        Href = lam0 * np.outer(gs, gs)
        assert np.abs(H - Href).max() <= 1e-5 * np.abs(Href).max()


def test_mollified_barrier_gradient(oracle, twobox5):
    """Mollified encodings (negative markers): gradient vs FD of e(I1) b(I2),
    Hessian symmetric PSD of rank <= 2."""
    sc, pt, ee, pairs, mi, counts = twobox5
    x = sc.x
    done = 0
    for cand in ee[:400]:
        a0, a1, b0, b1 = cand
        # well-conditioned line-line distance (edges at > ~6 degrees) so that
        # central differences are meaningful; the mollified algebra is the same
        e0 = x[a1] - x[a0]; e1 = x[b1] - x[b0]
        if np.sum(np.cross(e0, e1) ** 2) < 1e-2 * np.dot(e0, e0) * np.dot(e1, e1):
            continue
        for q in ([a0, a1, b0, -b1 - 1], [-a0 - 1, -b0 - 1, -a1 - 1, -b1 - 1], [-a0 - 1, -b0 - 1, b1, -a1 - 1]):
            q = np.array(q, np.int32)
            vid, cls, g, H, e = oracle.barrier_pair(x, sc.x_rest, q, 1e30, 1.0)
            if e == 0.0 or not np.isfinite(e):
                continue
            # use a dhat above the current distance so the pair is active
            vid, cls, g, H, e = oracle.barrier_pair(x, sc.x_rest, q, sc.dhat, sc.kappa)
            if e == 0.0:
                continue
            h = 1e-8
            fd = np.zeros(12)
            for a in range(4):
                for c in range(3):
                    xp = x.copy(); xp[vid[a], c] += h
                    xm = x.copy(); xm[vid[a], c] -= h
                    fd[3 * a + c] = (oracle.barrier_pair(xp, sc.x_rest, q, sc.dhat, sc.kappa)[4] -
                                     oracle.barrier_pair(xm, sc.x_rest, q, sc.dhat, sc.kappa)[4]) / (2 * h)
            assert np.abs(fd - g).max() <= 1e-4 * max(np.abs(g).max(), 1e-30)
            assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
            w = np.linalg.eigvalsh(H)
            assert w.min() >= -1e-9 * max(w.max(), 1e-300) and (w > 1e-9 * max(w.max(), 1e-300)).sum() <= 2
            done += 1
    assert done > 5


def test_barrier_assemble_layout(oracle, twobox5):
    sc, pt, ee, pairs, mi, counts = twobox5
    fb, vals, rows, cols = oracle.barrier_assemble(pairs, mi, counts, sc.x, sc.x_rest, sc.dhat, sc.kappa, sc.btype)
    assert len(rows) == 16 * counts[2] + 9 * counts[1] + 4 * counts[0]
    assert np.all(rows <= cols)
    # each pair's nonzero blocks are the upper blocks of its dense Hessian
    E = oracle.barrier_energy(pairs, sc.x, sc.x_rest, sc.dhat, sc.kappa)
    assert E > 0 and np.isfinite(fb).all()


def test_broad_phase_reference_is_complete(oracle):
    """The brute-force box candidates lose no pair the narrow phase keeps:
    classifying every point-triangle and edge-edge pair of the surfaces
    gives the same contact pairs as classifying only the box candidates."""
    from sgipc import scene
    from broad_phase_ref import candidates
    sc = scene.two_box_scene(4, seed=2)
    faces, edges, surf = scene.surface_primitives(sc.tets)
    pt, ee = candidates(sc.x, faces, edges, surf, sc.dhat)
    assert len(pt) and len(ee)
    f = faces.astype(np.int64); e = edges.astype(np.int64)
    P, F = np.meshgrid(surf.astype(np.int64), np.arange(len(f)), indexing="ij")
    allpt = np.stack([P.ravel(), f[F.ravel(), 0], f[F.ravel(), 1], f[F.ravel(), 2]], 1)
    allpt = allpt[(allpt[:, 0] != allpt[:, 1]) & (allpt[:, 0] != allpt[:, 2]) & (allpt[:, 0] != allpt[:, 3])]
    I, J = np.triu_indices(len(e), 1)
    allee = np.concatenate([e[I], e[J]], 1)
    share = (allee[:, 0] == allee[:, 2]) | (allee[:, 0] == allee[:, 3]) | (allee[:, 1] == allee[:, 2]) | \
        (allee[:, 1] == allee[:, 3])
    allee = allee[~share]
    pa, _, ca = oracle.classify(allpt.astype(np.int32), allee.astype(np.int32), sc.x, sc.x_rest, sc.dhat)
    pb, _, cb = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)
    assert len(pa) > 0 and np.array_equal(ca, cb)
    assert set(map(tuple, _canon(pa, sc.n))) == set(map(tuple, _canon(pb, sc.n)))
    assert len(pt) < len(allpt) // 10                      # and the boxes do prune


def _canon(pairs, n):
    """encoded pairs with negative fourth entries made uniform: the barrier
    reads only their sign, and parallel edge pairs store -(candidate)-2
    there, which depends on the candidate list"""
    q = pairs.copy()
    q[q[:, 3] < 0, 3] = -1
    return q
