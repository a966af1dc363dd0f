"""Synthetic scene generation for the Newton-step path.

Regular tet grids (5 tets per cell, alternating parity), block partitions into
<= bank-vertex parts with part-contiguous vertex numbering (the ordering the
reference's MAS assumes, produced there by MeshProcess/metis_partition/src/
metis_sort.cpp:25-383), rest-state precomputation (S/gl_main.cu:557-617), and
the workload configurations the benchmarks and tests use.

This is host-side input preparation in numpy; it is not on the timed path.
"""
from dataclasses import dataclass, field
import numpy as np

# corner offsets of a cell, index = a + 2 b + 4 c for offset (a, b, c)
_CORNERS = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.int64)


def _cid(a, b, c):
    return a + 2 * b + 4 * c


# even parity: central tet on the even corners {000,110,101,011}
_TETS_EVEN = [
    (_cid(0, 0, 0), _cid(1, 1, 0), _cid(1, 0, 1), _cid(0, 1, 1)),
    (_cid(1, 0, 0), _cid(0, 0, 0), _cid(1, 1, 0), _cid(1, 0, 1)),
    (_cid(0, 1, 0), _cid(0, 0, 0), _cid(1, 1, 0), _cid(0, 1, 1)),
    (_cid(0, 0, 1), _cid(0, 0, 0), _cid(1, 0, 1), _cid(0, 1, 1)),
    (_cid(1, 1, 1), _cid(1, 1, 0), _cid(1, 0, 1), _cid(0, 1, 1)),
]
# odd parity: central tet on the odd corners {100,010,001,111}
_TETS_ODD = [
    (_cid(1, 0, 0), _cid(0, 1, 0), _cid(0, 0, 1), _cid(1, 1, 1)),
    (_cid(0, 0, 0), _cid(1, 0, 0), _cid(0, 1, 0), _cid(0, 0, 1)),
    (_cid(1, 1, 0), _cid(1, 0, 0), _cid(0, 1, 0), _cid(1, 1, 1)),
    (_cid(1, 0, 1), _cid(1, 0, 0), _cid(0, 0, 1), _cid(1, 1, 1)),
    (_cid(0, 1, 1), _cid(0, 1, 0), _cid(0, 0, 1), _cid(1, 1, 1)),
]


def tet_grid(nx, ny, nz, h=1.0, origin=(0.0, 0.0, 0.0)):
    """Vertices (grid order, x fastest) and positively oriented tets."""
    gx, gy, gz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ijk = np.stack([gx.transpose(2, 1, 0).ravel(), gy.transpose(2, 1, 0).ravel(), gz.transpose(2, 1, 0).ravel()], 1)
    verts = ijk.astype(np.float64) * h + np.asarray(origin, dtype=np.float64)

    def vid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).ravel(); cj = cj.transpose(2, 1, 0).ravel(); ck = ck.transpose(2, 1, 0).ravel()
    corner_ids = np.stack([vid(ci + o[0], cj + o[1], ck + o[2]) for o in _CORNERS], 1)  # (cells, 8)
    parity = (ci + cj + ck) % 2
    tets = np.empty((ci.shape[0], 5, 4), dtype=np.int64)
    even = np.asarray(_TETS_EVEN); odd = np.asarray(_TETS_ODD)
    tets[parity == 0] = corner_ids[parity == 0][:, even]
    tets[parity == 1] = corner_ids[parity == 1][:, odd]
    tets = tets.reshape(-1, 4)
    # orient positively
    x = verts
    d = np.einsum("ij,ij->i", np.cross(x[tets[:, 1]] - x[tets[:, 0]], x[tets[:, 2]] - x[tets[:, 0]]),
                  x[tets[:, 3]] - x[tets[:, 0]])
    neg = d < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return verts, tets.astype(np.uint32), ijk


def _morton3(i, j, k):
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v
    return spread(i) | (spread(j) << np.uint64(1)) | (spread(k) << np.uint64(2))


def block_partition(ijk, bank=16, block=None):
    """Part id per vertex: vertex-grid blocks of <= bank vertices (2x2x4 for
    bank 16, 2x4x4 for 32), blocks numbered in Morton order so consecutive
    parts are spatially compact.  Returns part ids (not yet sorted)."""
    if block is None:
        block = (2, 2, 4) if bank == 16 else (2, 4, 4)
    b = np.asarray(block)
    bc = ijk // b
    code = _morton3(bc[:, 0], bc[:, 1], bc[:, 2])
    uniq, part = np.unique(code, return_inverse=True)
    return part.astype(np.int32)


def renumber_by_part(part, *arrays_to_permute):
    """Stable sort of vertices by part id; returns (perm, inv, sorted part)."""
    perm = np.argsort(part, kind="stable")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.shape[0])
    return perm, inv, part[perm]


def order_tets_by_vertex(tets):
    """Stable sort of tets by their smallest vertex id.  With vertices in part
    order this puts the tets around a vertex next to each other, so the
    per-tet Hessian triplets that one block row sums are written close
    together and the convert's merge gathers them from few cache lines
    (in cell order the tets around a vertex are a z-slab of cells apart)."""
    order = np.argsort(tets.min(axis=1), kind="stable")
    return np.ascontiguousarray(tets[order])


@dataclass
class Body:
    """One deformable body of the FEM subsystem (already part-sorted)."""
    x_rest: np.ndarray
    tets: np.ndarray
    part: np.ndarray
    btype: np.ndarray


@dataclass
class Scene:
    name: str
    x_rest: np.ndarray          # (n,3) float64 rest positions
    x: np.ndarray               # (n,3) current positions (Newton iterate)
    x_tilde: np.ndarray         # (n,3) inertial target
    tets: np.ndarray            # (T,4) uint32
    part: np.ndarray            # (n,) int32 ascending part ids
    btype: np.ndarray           # (n,) int32 (nonzero = pinned)
    DmInv: np.ndarray           # (T,9) row-major
    vol: np.ndarray
    mass: np.ndarray
    mu: np.ndarray
    lam: np.ndarray
    dt: float
    pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    dhat: float = 0.0           # squared barrier distance (reference convention)
    kappa: float = 0.0
    bank: int = 16
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.x.shape[0]

    @property
    def T(self):
        return self.tets.shape[0]


def lame_snk(E, nu):
    """(mu', lambda') of the SNK1 energy from (E, nu) (S/gl_main.cu:609-617)."""
    lmu = E / (2.0 * (1.0 + nu))
    llam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return 4.0 * lmu / 3.0, llam + 5.0 * lmu / 6.0


def rest_precompute(x_rest, tets, E, nu, density):
    """DmInv (row-major), volume, lumped mass, mu', lambda' per tet."""
    x = x_rest
    t = tets.astype(np.int64)
    Dm = np.stack([x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]], x[t[:, 3]] - x[t[:, 0]]], axis=2)
    DmInv = np.linalg.inv(Dm).reshape(-1, 9)
    vol = np.abs(np.linalg.det(Dm)) / 6.0
    mass = np.zeros(x.shape[0])
    for a in range(4):
        np.add.at(mass, t[:, a], vol * density / 4.0)
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), vol.shape)
    mu, lam = lame_snk(E, nu)
    return DmInv, vol, mass, np.ascontiguousarray(mu), np.ascontiguousarray(lam)


def make_box(nx, ny, nz, h, origin=(0.0, 0.0, 0.0), bank=16, pin_bottom=False):
    verts, tets, ijk = tet_grid(nx, ny, nz, h, origin)
    part = block_partition(ijk, bank)
    perm, inv, part_sorted = renumber_by_part(part)
    x = verts[perm]
    tets = order_tets_by_vertex(inv[tets.astype(np.int64)].astype(np.uint32))
    ijk = ijk[perm]
    btype = np.zeros(x.shape[0], np.int32)
    if pin_bottom:
        btype[ijk[:, 1] == 0] = 1
    return Body(x, tets, part_sorted.astype(np.int32), btype)


def concat_bodies(bodies):
    xs, ts, ps, bs = [], [], [], []
    voff = 0; poff = 0
    for b in bodies:
        xs.append(b.x_rest); ts.append(b.tets.astype(np.int64) + voff)
        ps.append(b.part.astype(np.int64) + poff); bs.append(b.btype)
        voff += b.x_rest.shape[0]; poff += int(b.part.max()) + 1
    return (np.concatenate(xs), np.concatenate(ts).astype(np.uint32), np.concatenate(ps).astype(np.int32),
            np.concatenate(bs).astype(np.int32))


def cube_scene(cells=10, seed=1, E=1e5, nu=0.45, density=1000.0, dt=0.01, gravity=-9.8, sigma_rel=0.01,
               bank=16, size=1.0):
    """Unit cube, cells^3 grid, 5 tets/cell, bottom face pinned, seeded
    Gaussian perturbation of the positions (the no-contact FP64 workload)."""
    h = size / cells
    body = make_box(cells, cells, cells, h, bank=bank, pin_bottom=True)
    rng = np.random.default_rng(seed)
    x_rest = body.x_rest
    DmInv, vol, mass, mu, lam = rest_precompute(x_rest, body.tets, E, nu, density)
    x = x_rest + rng.normal(0.0, sigma_rel * h, size=x_rest.shape)
    x[body.btype != 0] = x_rest[body.btype != 0]
    v = np.zeros_like(x)
    g = np.array([0.0, gravity, 0.0])
    x_tilde = x + dt * v + dt * dt * g
    x_tilde[body.btype != 0] = x[body.btype != 0]
    return Scene("cube%d" % cells, x_rest, x, x_tilde, body.tets, body.part, body.btype, DmInv, vol, mass,
                 mu, lam, dt, bank=bank, meta=dict(cells=cells, E=E, nu=nu, seed=seed))


def two_box_scene(cells=40, seed=2, E=1e6, nu=0.4, density=1000.0, dt=0.01, gravity=-9.8, size=1.0,
                  rel_dhat=1e-3, kappa=1e4, vel=0.5, sigma_rel=0.01, bank=16):
    """Two soft boxes (cells^3 each, 5 tets/cell) stacked along y with a gap
    of 0.5 dhat, dhat = rel_dhat * box size (a distance); the lower box's
    bottom face is pinned.  Seeded uniform initial velocities in [-vel, vel]
    and a seeded Gaussian perturbation (sigma_rel of the edge) of the
    positions.  Returns the Scene with dhat stored squared (reference
    convention, S/GIPC.cu:8602)."""
    h = size / cells
    dhat = rel_dhat * size
    gap = 0.5 * dhat
    lower = make_box(cells, cells, cells, h, origin=(0.0, 0.0, 0.0), bank=bank, pin_bottom=True)
    upper = make_box(cells, cells, cells, h, origin=(0.0, size + gap, 0.0), bank=bank, pin_bottom=False)
    x_rest, tets, part, btype = concat_bodies([lower, upper])
    rng = np.random.default_rng(seed)
    DmInv, vol, mass, mu, lam = rest_precompute(x_rest, tets, E, nu, density)
    # perturb tangentially only on the facing faces so the gap stays 0.5 dhat
    x = x_rest.copy()
    noise = rng.normal(0.0, sigma_rel * h, size=x.shape)
    facing = (np.abs(x_rest[:, 1] - size) < 1e-12) | (np.abs(x_rest[:, 1] - (size + gap)) < 1e-12)
    noise[facing, 1] = 0.0
    x += noise
    x[btype != 0] = x_rest[btype != 0]
    v = rng.uniform(-vel, vel, size=x.shape)
    v[btype != 0] = 0.0
    g = np.array([0.0, gravity, 0.0])
    x_tilde = x + dt * v + dt * dt * g
    x_tilde[btype != 0] = x[btype != 0]
    return Scene("twobox%d" % cells, x_rest, x, x_tilde, tets, part, btype, DmInv, vol, mass, mu, lam, dt,
                 dhat=dhat * dhat, kappa=kappa, bank=bank,
                 meta=dict(cells=cells, E=E, nu=nu, seed=seed, dhat_dist=dhat, gap=gap,
                           n_lower=lower.x_rest.shape[0]))


def boundary_triangles(tets):
    """Faces that belong to exactly one tet (unoriented vertex triples)."""
    t = tets.astype(np.int64)
    faces = np.concatenate([t[:, [0, 1, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 3]], t[:, [1, 2, 3]]])
    key = np.sort(faces, axis=1)
    uniq, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    return faces[idx[cnt == 1]]


def _grid_cells(pts_xz, h):
    return np.floor(pts_xz / h + 1e-9).astype(np.int64)


def facing_contact_candidates(sc, h=None):
    """Point-triangle and edge-edge candidate quadruples between the two
    facing faces of a two-box scene (the broad-phase stand-in: primitives
    whose grid cells are within one cell of each other).  Returns
    (pt (m,4) int32, ee (k,4) int32)."""
    size = 1.0
    cells = sc.meta["cells"]
    h = size / cells if h is None else h
    gap = sc.meta["gap"]
    n_lower = sc.meta["n_lower"]
    tris = boundary_triangles(sc.tets)
    xr = sc.x_rest
    ylo = np.abs(xr[:, 1] - size) < 1e-9
    yhi = np.abs(xr[:, 1] - (size + gap)) < 1e-9
    faces = []
    for sel in (ylo, yhi):
        fm = sel[tris].all(axis=1)
        faces.append(tris[fm])
    out_pt, out_ee = [], []

    def cell_of(idx):
        return _grid_cells(xr[idx][..., [0, 2]], h)

    # edges of each face
    edges = []
    for f in faces:
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [0, 2]]])
        e = np.unique(np.sort(e, axis=1), axis=0)
        edges.append(e)
    for a, b in ((0, 1), (1, 0)):
        # points of face a vs triangles of face b
        pts = np.unique(faces[a].reshape(-1))
        tri = faces[b]
        tcell = cell_of(tri).min(axis=1)
        buckets = {}
        for i, c in enumerate(map(tuple, tcell)):
            buckets.setdefault(c, []).append(i)
        pcell = cell_of(pts)
        for p, c in zip(pts, map(tuple, pcell)):
            for di in (-1, 0):
                for dk in (-1, 0):
                    for ti in buckets.get((c[0] + di, c[1] + dk), ()):
                        out_pt.append((p, tri[ti, 0], tri[ti, 1], tri[ti, 2]))
    ea, eb = edges
    bcell = cell_of(eb).min(axis=1)
    buckets = {}
    for i, c in enumerate(map(tuple, bcell)):
        buckets.setdefault(c, []).append(i)
    acell = cell_of(ea).min(axis=1)
    for i, c in enumerate(map(tuple, acell)):
        for di in (-1, 0, 1):
            for dk in (-1, 0, 1):
                for j in buckets.get((c[0] + di, c[1] + dk), ()):
                    out_ee.append((ea[i, 0], ea[i, 1], eb[j, 0], eb[j, 1]))
    pt = np.asarray(out_pt, np.int32).reshape(-1, 4)
    ee = np.asarray(out_ee, np.int32).reshape(-1, 4)
    return pt, ee


def lame_coefficients(nu):
    """(c_mu, c_lam) with (mu', lambda') = E * (c_mu, c_lam): the SNK1
    parameters are linear in Young's modulus at fixed Poisson ratio."""
    mu, lam = lame_snk(1.0, nu)
    return float(mu), float(lam)


def cantilever_scene(nx=60, ny=12, nz=12, seed=4, length=1.0, E_mean=1e5, E_logstd=0.2, nu=0.4, density=1000.0,
                     dt=0.01, gravity=-9.8, clearance=0.05, rel_dhat=1e-2, kappa=1e6, vel=0.1, target_std=0.01,
                     bank=16):
    """Cantilever bar of nx x ny x nz cells (edge length/nx, 5 tets/cell)
    clamped at its x = 0 face, its lower face `clearance` above the ground
    plane y = 0.  Per-tet ln E ~ N(ln E_mean, E_logstd^2), seeded uniform
    initial velocities in [-vel, vel] on the free vertices, and a seeded
    Gaussian target displacement field (std target_std) for the loss.  The
    ground barrier uses dhat = rel_dhat * length (a distance; stored
    squared, as the reference's dHat) and stiffness kappa.  Returns the
    Scene at rest (x = x_rest); meta carries log_E, v0, target, the surface
    vertex ids and the ground plane."""
    h = length / nx
    verts, tets, ijk = tet_grid(nx, ny, nz, h, (0.0, clearance, 0.0))
    part = block_partition(ijk, bank)
    perm, inv, part_sorted = renumber_by_part(part)
    x_rest = verts[perm]
    tets = inv[tets.astype(np.int64)].astype(np.uint32)
    ijk = ijk[perm]
    btype = (ijk[:, 0] == 0).astype(np.int32)
    rng = np.random.default_rng(seed)
    log_E = rng.normal(np.log(E_mean), E_logstd, size=tets.shape[0])
    DmInv, vol, mass, mu, lam = rest_precompute(x_rest, tets, np.exp(log_E), nu, density)
    v0 = rng.uniform(-vel, vel, size=x_rest.shape)
    v0[btype != 0] = 0.0
    target = x_rest + rng.normal(0.0, target_std, size=x_rest.shape)
    surf = np.unique(boundary_triangles(tets).reshape(-1)).astype(np.uint32)
    g = np.array([0.0, gravity, 0.0])
    x_tilde = x_rest + dt * v0 + dt * dt * g
    x_tilde[btype != 0] = x_rest[btype != 0]
    dhat = rel_dhat * length
    return Scene("cantilever%dx%dx%d" % (nx, ny, nz), x_rest, x_rest.copy(), x_tilde, tets,
                 part_sorted.astype(np.int32), btype, DmInv, vol, mass, mu, lam, dt, dhat=dhat * dhat, kappa=kappa,
                 bank=bank,
                 meta=dict(cells=(nx, ny, nz), nu=nu, density=density, gravity=gravity, seed=seed, log_E=log_E,
                           v0=v0, target=target, surf=surf, ground_normal=np.array([0.0, 1.0, 0.0]),
                           ground_offset=0.0, dhat_dist=dhat))


def ground_box_scene(n_target, aspect, E, nu, seed, dt=0.01, density=1000.0, gravity=-9.8, size=0.5,
                     rel_dhat=1e-3, kappa=1e4, bank=16, origin_xz=(0.0, 0.0)):
    """One free box of about n_target vertices with side ratios `aspect`
    (longest side `size`), resting 0.5 dhat above the ground plane y = 0
    (dhat = rel_dhat * size, stored squared), seeded Gaussian position
    perturbation (1% of the edge), zero velocity, gravity."""
    a = np.asarray(aspect, dtype=np.float64)
    k = (n_target / np.prod(a)) ** (1.0 / 3.0)
    cells = np.maximum(1, np.round(k * a - 1.0)).astype(int)
    h = size / cells.max()
    dhat = rel_dhat * size
    body = make_box(int(cells[0]), int(cells[1]), int(cells[2]), h, origin=(origin_xz[0], 0.5 * dhat, origin_xz[1]),
                    bank=bank)
    rng = np.random.default_rng(seed)
    x_rest = body.x_rest
    DmInv, vol, mass, mu, lam = rest_precompute(x_rest, body.tets, E, nu, density)
    x = x_rest + rng.normal(0.0, 0.01 * h, size=x_rest.shape)
    low = x_rest[:, 1] < 0.5 * dhat + 1e-12
    x[low, 1] = x_rest[low, 1]                      # the bottom face keeps its 0.5 dhat gap
    g = np.array([0.0, gravity, 0.0])
    x_tilde = x + dt * dt * g
    surf = np.unique(boundary_triangles(body.tets).reshape(-1)).astype(np.uint32)
    return Scene("box%d" % x.shape[0], x_rest, x, x_tilde, body.tets, body.part, body.btype, DmInv, vol, mass, mu,
                 lam, dt, dhat=dhat * dhat, kappa=kappa, bank=bank,
                 meta=dict(cells=tuple(int(c) for c in cells), E=E, nu=nu, seed=seed, surf=surf,
                           ground_normal=np.array([0.0, 1.0, 0.0]), ground_offset=0.0, gravity=gravity,
                           dhat_dist=dhat))


def batch_scenes(n_scenes=256, seed=5, n_min=500, n_max=20000, bank=16):
    """Independent scenes of uneven size: vertex counts log-uniform in
    [n_min, n_max], side ratios uniform in [1, 4], E log-uniform in
    [1e4, 1e7], nu uniform in [0.3, 0.48], each a free box on the ground
    plane; scene i is seeded with seed + i."""
    out = []
    for i in range(n_scenes):
        rng = np.random.default_rng(seed + i)
        n_t = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        aspect = rng.uniform(1.0, 4.0, size=3)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.3, 0.48))
        out.append(ground_box_scene(n_t, aspect, E, nu, seed + i, bank=bank, origin_xz=(0.6 * (i % 16), 0.6 * (i // 16))))
    return out


def union_scene(scenes):
    """Concatenate independent scenes (vertex, tet and part ranges in
    order); returns (Scene, row_start) with row_start[s] the first vertex of
    scene s.  The scenes share dt, dhat, kappa and the ground plane."""
    offs = np.cumsum([0] + [s.n for s in scenes])
    poff = np.cumsum([0] + [int(s.part.max()) + 1 for s in scenes])
    cat = lambda f: np.concatenate([f(s) for s in scenes])
    tets = np.concatenate([s.tets.astype(np.int64) + o for s, o in zip(scenes, offs)]).astype(np.uint32)
    part = np.concatenate([s.part.astype(np.int64) + p for s, p in zip(scenes, poff)]).astype(np.int32)
    surf = np.concatenate([s.meta["surf"].astype(np.int64) + o for s, o in zip(scenes, offs)]).astype(np.uint32)
    s0 = scenes[0]
    sc = Scene("batch%d" % len(scenes), cat(lambda s: s.x_rest), cat(lambda s: s.x), cat(lambda s: s.x_tilde), tets,
               part, cat(lambda s: s.btype).astype(np.int32), cat(lambda s: s.DmInv), cat(lambda s: s.vol),
               cat(lambda s: s.mass), cat(lambda s: s.mu), cat(lambda s: s.lam), s0.dt, dhat=s0.dhat,
               kappa=s0.kappa, bank=s0.bank,
               meta=dict(surf=surf, ground_normal=s0.meta["ground_normal"], ground_offset=s0.meta["ground_offset"],
                         gravity=s0.meta["gravity"], nu=s0.meta["nu"], scenes=len(scenes)))
    return sc, offs.astype(np.int32)


def surface_primitives(tets):
    """Surface triangles (boundary faces, as stored in the tet), surface
    edges (unique, sorted (lo, hi), lexicographic order) and surface
    vertices (ascending) of a tet mesh, as uint32 arrays -- the reference's
    _faces / _edges / _surfVerts inputs of the broad phase."""
    faces = np.ascontiguousarray(boundary_triangles(tets), np.uint32)
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [0, 2]]]).astype(np.int64)
    e = np.unique(np.sort(e, axis=1), axis=0)
    verts = np.unique(faces.reshape(-1))
    return faces, np.ascontiguousarray(e, np.uint32), np.ascontiguousarray(verts, np.uint32)


def mesh_scene(x_rest, tets, part, E=1e5, nu=0.45, density=1000.0, dt=0.01, gravity=-9.8, pinned=None, bank=16):
    """Scene at rest from a loaded tet mesh (e.g. sgipc.io.read_msh of the
    reference's *_sorted.16.msh with its .part ids, vertices already in part
    order): lumped masses, SNK1 parameters, optional pinned-vertex mask."""
    x_rest = np.ascontiguousarray(x_rest, np.float64).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    # positive orientation, as the precomputation expects
    t = tets.astype(np.int64)
    d = np.einsum("ij,ij->i", np.cross(x_rest[t[:, 1]] - x_rest[t[:, 0]], x_rest[t[:, 2]] - x_rest[t[:, 0]]),
                  x_rest[t[:, 3]] - x_rest[t[:, 0]])
    neg = d < 0
    tets = tets.copy()
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    DmInv, vol, mass, mu, lam = rest_precompute(x_rest, tets, E, nu, density)
    btype = np.zeros(x_rest.shape[0], np.int32) if pinned is None else np.asarray(pinned, np.int32)
    g = np.array([0.0, gravity, 0.0])
    x_tilde = x_rest + dt * dt * g
    x_tilde[btype != 0] = x_rest[btype != 0]
    return Scene("mesh%d" % x_rest.shape[0], x_rest, x_rest.copy(), x_tilde, tets,
                 np.ascontiguousarray(part, np.int32), btype, DmInv, vol, mass, mu, lam, dt, bank=bank,
                 meta=dict(E=E, nu=nu, density=density, gravity=gravity))


@dataclass
class HybridScene:
    """Affine bodies (stiff, 12 DOF each) over a soft FEM body.  Vertex ids:
    the nav body vertices first (grouped by body), then the FEM vertices of
    `fem`; DOF block rows: 4 per body, then one per FEM vertex."""
    fem: Scene                  # soft part (its own vertex numbering)
    vstart: np.ndarray          # (nb + 1,) body vertex ranges
    xbar: np.ndarray            # (nav, 3) body-frame rest coordinates (about the mass centre)
    body_tets: np.ndarray       # (Tb, 4) body tets in body-vertex ids
    q: np.ndarray               # (nb, 12) current q = (p, a1, a2, a3)
    q_tilde: np.ndarray         # (nb, 12) inertial target
    mass12: np.ndarray          # (nb, 144) row-major sum_i m_i J_i^T J_i
    volume: np.ndarray          # (nb,)
    fixed: np.ndarray           # (nb,) int32
    kappa_abd: float
    faces: np.ndarray           # union surface triangles (vertex space)
    edges: np.ndarray
    surf: np.ndarray
    body_id: np.ndarray         # (nav + nf,) int32: body of a body vertex, -1 for FEM vertices
    meta: dict = field(default_factory=dict)

    @property
    def nb(self):
        return self.q.shape[0]

    @property
    def nav(self):
        return self.xbar.shape[0]


def _rot_y(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def hybrid_scene(cells=(100, 20, 100), lattice=4, seed=3, E=1e5, nu=0.45, density=1000.0, dt=0.005,
                 gravity=-9.8, size=1.0, cube=0.1, rel_dhat=1e-3, kappa=1e4, kappa_abd=1e11, bank=16,
                 jitter=0.05, max_turn=0.3):
    """Hybrid stiff/soft workload: lattice^3 affine cubes (one cell, 8
    vertices and 5 tets each) over a soft SNK1 slab of cells = (nx, ny, nz)
    cells (y up, bottom face pinned), slab width `size`.  The cubes sit on a
    seeded jittered lattice (x/z offsets up to `jitter` * size, turns about y
    up to `max_turn` rad) in columns of `lattice` layers, each layer resting
    0.5 dhat above the one below and the bottom layer 0.5 dhat above the
    slab, so body-slab and body-body contact pairs are active.  dhat =
    rel_dhat * size (a distance; stored squared).  q starts at (mass centre,
    rotation rows); xbar about the lumped mass centre (ABDSystem::_setup_J and
    _calculate_body_mass_center, StiffGIPC/abd_system/abd_system.cu:172-251);
    masses lumped per tet vertex (_setup_unique_point_mass, :141-170)."""
    nx, ny, nz = cells
    h = size / nx
    thick = ny * h
    dhat = rel_dhat * size
    gap = 0.5 * dhat
    slab = make_box(nx, ny, nz, h, bank=bank, pin_bottom=True)
    DmInv, vol, mass, mu, lam = rest_precompute(slab.x_rest, slab.tets, E, nu, density)
    g = np.array([0.0, gravity, 0.0])
    x_tilde = slab.x_rest + dt * dt * g
    x_tilde[slab.btype != 0] = slab.x_rest[slab.btype != 0]
    fem = Scene("slab%dx%dx%d" % (nx, ny, nz), slab.x_rest, slab.x_rest.copy(), x_tilde, slab.tets, slab.part,
                slab.btype, DmInv, vol, mass, mu, lam, dt, dhat=dhat * dhat, kappa=kappa, bank=bank,
                meta=dict(E=E, nu=nu, gravity=gravity))
    # one cube: a single grid cell, lumped masses, mass centre
    cv, ct, _ = tet_grid(1, 1, 1, cube)
    _, cvol, cmass, _, _ = rest_precompute(cv, ct, E, nu, density)
    centre = (cmass[:, None] * cv).sum(0) / cmass.sum()
    xb = cv - centre
    J = np.zeros((8, 3, 12))
    J[:, :, :3] = np.eye(3)
    for k in range(3):
        J[:, k, 3 + 3 * k:6 + 3 * k] = xb
    M1 = np.einsum("i,iak,ial->kl", cmass, J, J)
    rng = np.random.default_rng(seed)
    nb = lattice ** 3
    q = np.zeros((nb, 12))
    pos = np.zeros((nb * 8, 3))
    spacing = size / lattice
    b = 0
    for layer in range(lattice):
        y0 = thick + gap + layer * (cube + gap)      # bottom face of this layer
        for k in range(lattice):
            for i in range(lattice):
                R = _rot_y(rng.uniform(-max_turn, max_turn))
                cx = (i + 0.5) * spacing + rng.uniform(-jitter, jitter) * size
                cz = (k + 0.5) * spacing + rng.uniform(-jitter, jitter) * size
                p = np.array([cx, y0 + centre[1], cz])
                q[b, :3] = p
                q[b, 3:] = R.reshape(9)               # rows a1, a2, a3
                pos[8 * b:8 * b + 8] = p + xb @ R.T
                b += 1
    vstart = (8 * np.arange(nb + 1)).astype(np.int32)
    xbar = np.tile(xb, (nb, 1))
    body_tets = (ct.astype(np.int64)[None, :, :] + 8 * np.arange(nb)[:, None, None]).reshape(-1, 4)
    nav = 8 * nb
    all_tets = np.concatenate([body_tets, slab.tets.astype(np.int64) + nav]).astype(np.uint32)
    faces, edges, surf = surface_primitives(all_tets)
    qg = np.zeros(12)
    qg[:3] = g                                        # mass-centred xbar: gravity moves p only
    q_tilde = q + dt * dt * qg
    body_id = np.concatenate([np.repeat(np.arange(nb), 8), -np.ones(slab.x_rest.shape[0], np.int64)]).astype(np.int32)
    return HybridScene(fem, vstart, xbar, body_tets.astype(np.uint32), q, q_tilde,
                       np.tile(M1.reshape(1, 144), (nb, 1)), np.full(nb, float(cvol.sum())), np.zeros(nb, np.int32),
                       float(kappa_abd), faces, edges, surf, body_id,
                       meta=dict(cells=cells, lattice=lattice, seed=seed, dhat_dist=dhat, gap=gap, cube=cube,
                                 positions=pos, gravity=gravity))
