"""Brute-force CPU reference of the contact broad phase (all pairs).

TEST INFRASTRUCTURE ONLY: the checker for the library's LBVH broad phase.
Semantics of S/mlbvh.cu: point box (the vertex) against triangle boxes, and
edge boxes against edge boxes, kept when closer than sqrt(dhat) on every
axis (`overlap`, StiffGIPC/mlbvh.cu:33-42), with the query filters of
_selfQuery_vf (:1306-1406: point not in the face) and _selfQuery_ee
(:1509-1632: obj > self, no shared vertex), and for both: not all four
vertices of btype >= 2, and different bodies or body id -1.
"""
import numpy as np


def _boxes(x, prim, x1=None):
    p = x[prim.astype(np.int64)]
    if x1 is not None:  # swept from x to x1
        p = np.concatenate([p, x1[prim.astype(np.int64)]], axis=1)
    return p.min(axis=1), p.max(axis=1)


def _near(alo, ahi, blo, bhi, gap):
    ok = np.ones((alo.shape[0], blo.shape[0]), bool)
    for k in range(3):
        ok &= ~((blo[None, :, k] - ahi[:, None, k] >= gap) | (alo[:, None, k] - bhi[None, :, k] >= gap))
    return ok


def candidates(x, faces, edges, surf, dhat, btype=None, body=None, dir=None, alpha=0.0):
    """(pt (m,4) int32, ee (k,4) int32) in lexicographic order.  With dir,
    every box is swept from x to x - alpha dir (the CCD candidates,
    StiffGIPC/mlbvh.cu:1172-1204, :1408-1760)."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    x1 = None if dir is None else x - alpha * np.asarray(dir, np.float64).reshape(-1, 3)
    gap = np.sqrt(dhat)
    f = faces.astype(np.int64); e = edges.astype(np.int64); sv = surf.astype(np.int64)
    bt = np.zeros(x.shape[0], np.int64) if btype is None else np.asarray(btype, np.int64)
    bd = -np.ones(x.shape[0], np.int64) if body is None else np.asarray(body, np.int64)
    flo, fhi = _boxes(x, f, x1)
    plo, phi = _boxes(x, sv[:, None], x1)
    ok = _near(plo, phi, flo, fhi, gap)
    ok &= (sv[:, None] != f[None, :, 0]) & (sv[:, None] != f[None, :, 1]) & (sv[:, None] != f[None, :, 2])
    ok &= (bd[sv][:, None] == -1) | (bd[sv][:, None] != bd[f[:, 0]][None, :])
    ok &= ~((bt[sv][:, None] >= 2) & (bt[f[:, 0]] >= 2)[None, :] & (bt[f[:, 1]] >= 2)[None, :] &
            (bt[f[:, 2]] >= 2)[None, :])
    i, j = np.nonzero(ok)
    pt = np.stack([sv[i], f[j, 0], f[j, 1], f[j, 2]], 1)
    elo, ehi = _boxes(x, e, x1)
    ok = _near(elo, ehi, elo, ehi, gap)
    ok &= np.triu(np.ones_like(ok), 1).astype(bool)
    for a in range(2):
        for b in range(2):
            ok &= e[:, None, a] != e[None, :, b]
    ok &= (bd[e[:, 0]][:, None] == -1) | (bd[e[:, 0]][:, None] != bd[e[:, 0]][None, :])
    ok &= ~((bt[e[:, 0]] >= 2)[:, None] & (bt[e[:, 1]] >= 2)[:, None] & (bt[e[:, 0]] >= 2)[None, :] &
            (bt[e[:, 1]] >= 2)[None, :])
    i, j = np.nonzero(ok)
    ee = np.stack([e[i, 0], e[i, 1], e[j, 0], e[j, 1]], 1)
    return pt.astype(np.int32), ee.astype(np.int32)
