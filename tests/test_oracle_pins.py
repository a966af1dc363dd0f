"""Independent pins for three oracle functions whose earlier checks only
compared them with the GPU (which follows the same reading of the
reference):

* MAS contact connectivity (S/MASPreconditioner.cu:1132-1540): the level
  tables with contact pairs equal connected components computed from
  scratch, within each bank, of the mesh graph plus one clique per contact
  pair.
* EE pair classification (S/mlbvh.cu:449-532, encoders :685-1060): the
  distance of the sub-case a candidate is classified into equals the exact
  segment-segment distance, and the fourth index encodes the near-parallel
  test of the reference (-(candidate) - 2 when |eA x eB|^2 < eps_x).
* Mollified EE Hessian (S/GIPC.cu:3560-3619): the 2x2 block in (c, s),
  c = |eA x eB|, s = sqrt(d / dhat), recovered from the oracle's 12x12
  Hessian through J = [dc/dx; ds/dx] (complex-step derivatives of
  independent geometry), equals the eigen-clamped central-difference
  Hessian of the closed-form energy e(c^2) b(s^2).
"""
import numpy as np
import pytest

from conftest import oracle_system
from test_oracle import brute_force_tables, _mesh_edges


# --------------------------------------------------------------------------
# shared geometry, written from the definitions (not from the oracle)

def _decode_vertices(q):
    """Vertex ids of an int4 contact pair (sign encoding of S/mlbvh.cu:535-1060:
    x >= 0 edge-edge; x < 0 point-based with z < 0 point-point, w < 0
    point-edge, else point-triangle; negative entries v hold -v - 1)."""
    x, y, z, w = (int(t) for t in q)
    dec = lambda t: t if t >= 0 else -t - 1
    if x >= 0:
        return [x, y, z, dec(w)]
    if z < 0:
        if y < 0:
            return [dec(x), dec(y), dec(z), dec(w)]
        return [dec(x), y]
    if w < 0:
        if y < 0:
            return [dec(x), dec(y), z, dec(w)]
        return [dec(x), y, z]
    return [dec(x), y, z, w]


def _point_segment_d2(p, a, b):
    e = b - a
    t = np.clip(np.dot(p - a, e) / np.dot(e, e), 0.0, 1.0)
    r = p - (a + t * e)
    return float(np.dot(r, r))


def _segment_segment_d2(p0, p1, q0, q1):
    """Exact squared distance of two segments: the smaller of the four
    endpoint-to-segment distances and, for non-parallel lines, the interior
    stationary point of |P(s) - Q(t)|^2 when it lies in (0, 1)^2."""
    best = min(_point_segment_d2(p0, q0, q1), _point_segment_d2(p1, q0, q1),
               _point_segment_d2(q0, p0, p1), _point_segment_d2(q1, p0, p1))
    u, v, w = p1 - p0, q1 - q0, p0 - q0
    a, b, c, d, e = np.dot(u, u), np.dot(u, v), np.dot(v, v), np.dot(u, w), np.dot(v, w)
    D = a * c - b * b
    if D > 1e-14 * a * c:
        s = (b * e - c * d) / D
        t = (a * e - b * d) / D
        if 0.0 < s < 1.0 and 0.0 < t < 1.0:
            r = w + s * u - t * v
            best = min(best, float(np.dot(r, r)))
    return best


def _subcase_d2(kind, v):
    """Squared distance of a classified sub-case: point-point, point to the
    line of an edge, line-line."""
    if kind == "pp":
        return float(np.sum((v[0] - v[1]) ** 2))
    if kind == "pe":
        e = v[2] - v[1]
        return float(np.sum(np.cross(v[1] - v[0], v[2] - v[0]) ** 2) / np.dot(e, e))
    n = np.cross(v[1] - v[0], v[3] - v[2])
    return float(np.dot(v[2] - v[0], n) ** 2 / np.dot(n, n))


def _subcase_d2_analytic(kind, v):
    """_subcase_d2 with sums of squares instead of norms (no conjugation)"""
    if kind == "pp":
        w = v[0] - v[1]
        return np.sum(w * w)
    if kind == "pe":
        e = v[2] - v[1]
        w = np.cross(v[1] - v[0], v[2] - v[0])
        return np.sum(w * w) / np.sum(e * e)
    n = np.cross(v[1] - v[0], v[3] - v[2])
    return np.sum((v[2] - v[0]) * n) ** 2 / np.sum(n * n)


# --------------------------------------------------------------------------
# MAS contact connectivity

def _contact_clique_edges(pairs, n):
    e = []
    for q in pairs:
        vs = [v for v in _decode_vertices(q) if 0 <= v < n]
        for i in range(len(vs)):
            for j in range(i + 1, len(vs)):
                if vs[i] != vs[j]:
                    e.append((vs[i], vs[j]))
                    e.append((vs[j], vs[i]))
    return e


@pytest.mark.parametrize("cells", [5, 8])
def test_mas_tables_with_contacts_vs_bruteforce(oracle, cells):
    from sgipc import scene
    sc = scene.two_box_scene(cells, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    pairs, mi, counts = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)
    assert len(pairs) > 20
    ref = oracle_system(oracle, sc)
    kr, kc, kv = ref["bcoo"]
    ns, nn, nl = oracle.vertex_neighbors(sc.n, sc.tets)
    pm, rm = oracle.bank_map(sc.part, sc.bank)
    mesh_edges = _mesh_edges(ns, nn, nl)

    with_c = oracle.Mas(sc.n, pm, rm, ns, nn, nl, sc.bank)
    with_c.build(kr, kc, kv, pairs=pairs)
    t = with_c.tables()
    L = with_c.levelnum
    counts_bf, table_bf = brute_force_tables(sc.n, pm, rm, mesh_edges + _contact_clique_edges(pairs, sc.n),
                                             sc.bank, L)
    for l in range(1, L):
        assert t["level_size"][l, 0] == counts_bf[l - 1]
        assert np.array_equal(t["coarse_space"][l - 1], table_bf[l - 1])

    # the contacts change the aggregation (the check above is not vacuous)
    without = oracle.Mas(sc.n, pm, rm, ns, nn, nl, sc.bank)
    without.build(kr, kc, kv)
    t0 = without.tables()
    assert any(not np.array_equal(t0["coarse_space"][l - 1], t["coarse_space"][l - 1]) for l in range(1, L))


# --------------------------------------------------------------------------
# EE classification

def _check_ee_candidates(oracle, x, xr, ee):
    pairs, mi, counts = oracle.classify(np.zeros((0, 4), np.int32), ee, x, xr, 1e30)
    assert len(pairs) == len(ee)              # nothing dropped below dhat = 1e30
    kinds = {}
    for i, (cand, q) in enumerate(zip(ee, pairs)):
        p0, p1, q0, q1 = (x[k] for k in cand)
        dtrue = _segment_segment_d2(p0, p1, q0, q1)
        vs = _decode_vertices(q)
        kind = {2: "pp", 3: "pe", 4: "ee"}[len(vs)]
        assert set(vs) <= set(int(k) for k in cand)
        d = _subcase_d2(kind, x[vs])
        scale = max(dtrue, 1e-300)
        assert abs(d - dtrue) <= 1e-9 * scale + 1e-24, (i, kind, d, dtrue)
        # near-parallel marker: eps_x = 1e-3 |eA0|^2 |eB0|^2 on rest edges
        ea, eb = xr[cand[0]] - xr[cand[1]], xr[cand[2]] - xr[cand[3]]
        eps = 1e-3 * np.dot(ea, ea) * np.dot(eb, eb)
        par = np.sum(np.cross(p0 - p1, q0 - q1) ** 2) < eps
        if kind == "ee":
            assert np.array_equal(q, cand)
        else:
            assert q[3] == (-i - 2 if par else -1)
        kinds[(kind, bool(par))] = kinds.get((kind, bool(par)), 0) + 1
    return kinds


@pytest.mark.parametrize("cells", [5, 6, 8])
def test_ee_classification_is_exact_segment_distance(oracle, cells):
    from sgipc import scene
    sc = scene.two_box_scene(cells, seed=2)
    _, ee = scene.facing_contact_candidates(sc)
    assert len(ee) > 100
    kinds = _check_ee_candidates(oracle, sc.x, sc.x_rest, ee)
    # the facing grid produces endpoint, point-edge and interior cases, and
    # near-parallel edges
    assert kinds.get(("ee", False), 0) > 0
    assert sum(v for (k, p), v in kinds.items() if k == "pe") > 0
    assert sum(v for (k, p), v in kinds.items() if p) > 0


def test_ee_classification_degenerate_configurations(oracle):
    rng = np.random.default_rng(31)
    pts = []
    # exactly parallel offset, collinear overlapping, collinear disjoint,
    # perpendicular crossing, shared-endpoint-like touching, T-junction
    base = [
        ([0, 0, 0], [1, 0, 0], [0.3, 0.2, 0], [1.4, 0.2, 0]),
        ([0, 0, 0], [1, 0, 0], [0.5, 0, 0], [1.5, 0, 0]),
        ([0, 0, 0], [1, 0, 0], [2.0, 0, 0], [3.0, 0, 0]),
        ([0, 0, 0], [1, 0, 0], [0.5, -0.5, 0.1], [0.5, 0.5, 0.1]),
        ([0, 0, 0], [1, 0, 0], [1.0, 0.0, 1e-3], [2.0, 0.5, 1e-3]),
        ([0, 0, 0], [1, 0, 0], [0.5, 0.1, 0], [0.5, 1.0, 0]),
        ([0, 0, 0], [1, 0, 0], [-0.2, 0.05, 0.01], [0.3, 0.05 + 1e-5, 0.01]),
    ]
    for a0, a1, b0, b1 in base:
        pts.append(np.array([a0, a1, b0, b1], float))
    # random near-parallel pairs at small offsets
    for _ in range(200):
        a0 = rng.normal(size=3); d = rng.normal(size=3); d /= np.linalg.norm(d)
        a1 = a0 + d * rng.uniform(0.5, 1.5)
        off = rng.normal(size=3) * 1e-2
        tilt = rng.normal(size=3) * rng.choice([0.0, 1e-4, 1e-2, 0.3])
        b0 = a0 + d * rng.uniform(-1.0, 1.5) + off
        b1 = b0 + (d + tilt) * rng.uniform(0.5, 1.5)
        pts.append(np.array([a0, a1, b0, b1]))
    x = np.concatenate(pts)
    ee = np.arange(x.shape[0], dtype=np.int32).reshape(-1, 4)
    kinds = _check_ee_candidates(oracle, x, x.copy(), ee)
    assert sum(v for (k, p), v in kinds.items() if p) > 20
    assert sum(v for (k, p), v in kinds.items() if k == "pp") > 0


# --------------------------------------------------------------------------
# mollified EE Hessian

def _b(I, kappa, dhat):
    """barrier b(I) = kappa dhat^2 (I - 1)^2 ln^2 I (S/GIPC.cu:725-726)"""
    return kappa * dhat * dhat * (I - 1.0) ** 2 * np.log(I) ** 2


def _mollified_energy(c, s, eps, kappa, dhat):
    """e(c^2) b(s^2), e(I1) = -I1 (I1 - 2 eps) / eps^2"""
    I1 = c * c
    return -I1 * (I1 - 2 * eps) / (eps * eps) * _b(s * s, kappa, dhat)


def _mollified_edges_and_subcase(q):
    """vertex slots of the two mollifier edges and the distance sub-case of a
    mollified encoding (EE: (a0,a1,b0,-b1-1); PP: all negative; PE:
    (-a0-1,-b0-1,b1,-a1-1))."""
    if q[0] >= 0:
        return (0, 1, 2, 3), "ee"
    if q[2] < 0:
        return (0, 2, 1, 3), "pp"
    return (0, 3, 1, 2), "pe"


def _psd2(M):
    w, V = np.linalg.eigh(M)
    return (V * np.maximum(w, 0.0)) @ V.T


def test_mollified_ee_hessian_is_projected_energy_hessian(oracle):
    rng = np.random.default_rng(17)
    kappa, dhat = 1e4, 1e-4                    # dhat squared (reference convention)
    checked, indefinite = 0, 0
    for it in range(400):
        # nearly parallel edges a few dhat apart: c^2 < eps_x and d < dhat
        a0 = rng.normal(size=3); d = rng.normal(size=3); d /= np.linalg.norm(d)
        a1 = a0 + d * rng.uniform(0.8, 1.2)
        nrm = np.cross(d, rng.normal(size=3)); nrm /= np.linalg.norm(nrm)
        gap = np.sqrt(dhat) * rng.uniform(0.1, 0.9)
        tilt = np.cross(d, nrm) * rng.uniform(-0.02, 0.02)
        b0 = a0 + d * rng.uniform(-0.3, 0.3) + nrm * gap
        b1 = b0 + (d + tilt) * rng.uniform(0.8, 1.2)
        x = np.array([a0, a1, b0, b1])
        xr = x + rng.normal(size=x.shape) * 1e-3
        enc = [np.array([0, 1, 2, -3 - 1], np.int32),            # EE (interior)
               np.array([-0 - 1, -2 - 1, -1 - 1, -3 - 1], np.int32),  # PP a0-b0
               np.array([-0 - 1, -2 - 1, 3, -1 - 1], np.int32)]   # PE a0 to edge (b0, b1)
        q = enc[it % 3]
        vid, cls, g, H, e = oracle.barrier_pair(x, xr, q, dhat, kappa)
        if e == 0.0:
            continue                            # outside dhat for this sub-case
        sl, kind = _mollified_edges_and_subcase(q)
        xv = x[vid]
        ra, rb = xr[vid[sl[0]]] - xr[vid[sl[1]]], xr[vid[sl[2]]] - xr[vid[sl[3]]]
        eps = 1e-3 * np.dot(ra, ra) * np.dot(rb, rb)

        # c and s written without conjugation, so that they are analytic in
        # the coordinates and complex-step differentiation applies
        def c_of(v):
            w = np.cross(v[sl[0]] - v[sl[1]], v[sl[2]] - v[sl[3]])
            return np.sqrt(np.sum(w * w))

        def s_of(v):
            return np.sqrt(_subcase_d2_analytic(kind, v) / dhat)
        c0, s0 = float(c_of(xv)), float(s_of(xv))
        assert abs(e - _mollified_energy(c0, s0, eps, kappa, dhat)) <= 1e-10 * abs(e)
        # J = [dc/dx; ds/dx] by complex-step differentiation (exact to round-off)
        hcs = 1e-30
        J = np.zeros((2, 12))
        for a in range(4):
            for k in range(3):
                vp = xv.astype(complex); vp[a, k] += 1j * hcs
                J[0, 3 * a + k] = c_of(vp).imag / hcs
                J[1, 3 * a + k] = s_of(vp).imag / hcs
        # 2x2 Hessian of the closed-form energy in (c, s), central differences
        hc, hs = 1e-4 * c0, 1e-4 * s0
        E = lambda cc, ss: _mollified_energy(cc, ss, eps, kappa, dhat)
        Ecc = (E(c0 + hc, s0) - 2 * E(c0, s0) + E(c0 - hc, s0)) / hc ** 2
        Ess = (E(c0, s0 + hs) - 2 * E(c0, s0) + E(c0, s0 - hs)) / hs ** 2
        Ecs = (E(c0 + hc, s0 + hs) - E(c0 + hc, s0 - hs) - E(c0 - hc, s0 + hs) + E(c0 - hc, s0 - hs)) / (4 * hc * hs)
        M = np.array([[Ecc, Ecs], [Ecs, Ess]])
        P = _psd2(M)
        # the oracle's block, recovered from H = J^T P J
        G = J @ J.T
        Pr = np.linalg.solve(G, np.linalg.solve(G, J @ H @ J.T).T).T
        scale = np.abs(P).max()
        assert np.abs(J.T @ Pr @ J - H).max() <= 1e-6 * np.abs(H).max()
        assert np.abs(Pr - P).max() <= 1e-4 * scale, (kind, Pr, P)
        # gradient: dE/dc dc/dx + dE/ds ds/dx (complex-step dE/dc, dE/ds); the
        # two terms can nearly cancel, so the bound is on their magnitudes
        Ec = E(c0 + 1j * hcs, s0).imag / hcs
        Es = E(c0, s0 + 1j * hcs).imag / hcs
        gref = Ec * J[0] + Es * J[1]
        assert np.abs(g - gref).max() <= 1e-9 * (np.abs(Ec * J[0]) + np.abs(Es * J[1])).max()
        checked += 1
        indefinite += np.linalg.eigvalsh(M)[0] < 0
    assert checked > 100
    assert indefinite > 10                      # the projection is exercised
