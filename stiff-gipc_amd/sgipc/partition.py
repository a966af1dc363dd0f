"""Vertex partition into MAS subdomains through the library's host
partitioner (sgipc_partition_mesh; the reference runs METIS in metis_sort,
MeshProcess/metis_partition/src/metis_sort.cpp:25-383)."""
import ctypes as C

import numpy as np

from .capi import SgipcError, lib


def _check(rc):
    if rc:
        raise SgipcError(rc, (lib().sgipc_last_error(None) or b"").decode())


def partition_mesh(n, tets, bank=16):
    """(part (n,) int32, order (n,) int32, n_parts, offset): part[v] = part of
    vertex v (at most `bank` vertices each); order = vertex ids sorted by part
    (new vertex i is old vertex order[i])."""
    t = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    part = np.zeros(max(n, 1), np.int32)
    order = np.zeros(max(n, 1), np.int32)
    k, off = C.c_int(), C.c_int()
    _check(lib().sgipc_partition_mesh(int(n), t.shape[0], t.ctypes.data, int(bank), part.ctypes.data,
                                      order.ctypes.data, C.byref(k), C.byref(off)))
    return part[:n], order[:n], k.value, off.value


def edge_cut(n, tets, part):
    """Weight of the vertex-graph edges (tets sharing the edge) cut by part."""
    t = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    p = np.ascontiguousarray(part, np.int32)
    cut = C.c_longlong()
    _check(lib().sgipc_partition_edge_cut(int(n), t.shape[0], t.ctypes.data, p.ctypes.data, C.byref(cut)))
    return cut.value


def renumber(x, tets, part, order):
    """Vertices in part order: (x', tets', part' ascending)."""
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order), dtype=order.dtype)
    return (np.ascontiguousarray(np.asarray(x)[order]), inv[np.asarray(tets, np.int64)].astype(np.uint32),
            np.ascontiguousarray(np.asarray(part)[order]))
