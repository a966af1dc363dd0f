"""The reference's partition step, with METIS built from the reference's
vendored sources (oracle/_ref/libmetis.so, `make -C oracle _ref/libmetis.so`).

TEST INFRASTRUCTURE ONLY: the yardstick for the library's own partitioner
(sgipc_partition_mesh).  Restates metis_sort's tet branch
(MeshProcess/metis_partition/src/metis_sort.cpp:217-290): the vertex graph
with edge weight = tets sharing the edge (TetMesh::_build_adj,
mesh.cpp:486-549, neighbours ascending), METIS_PartGraphKway with default
options (NodeEdgeModel::k_way_partition, node_edge_model.cpp:5-28) into
nParts = ceil(n / (16 - off)) parts, off raised until no part exceeds 16.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_ref", "libmetis.so")
_lib = None


def available():
    if os.path.exists(_LIB):
        return True
    if os.path.isdir("/root/reference/MeshProcess/External/METIS"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "_ref/libmetis.so"])
        return os.path.exists(_LIB)
    return False


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libmetis.so is not built (needs the reference tree)")
        L = C.CDLL(_LIB)
        L.METIS_PartGraphKway.argtypes = [C.c_void_p] * 13
        L.METIS_PartGraphKway.restype = C.c_int
        _lib = L
    return _lib


def vertex_graph(n, tets):
    """CSR (xadj, adjncy, adjwgt) int32 as TetMesh::_build_adj: for each
    tet, each ordered vertex pair adds 1; neighbours ascending."""
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    a = np.concatenate([t[:, i] for i in range(4) for j in range(4) if i != j])
    b = np.concatenate([t[:, j] for i in range(4) for j in range(4) if i != j])
    key = a * n + b
    uk, cnt = np.unique(key, return_counts=True)
    src = (uk // n).astype(np.int32)
    adj = (uk % n).astype(np.int32)
    xadj = np.zeros(n + 1, np.int32)
    np.add.at(xadj, src + 1, 1)
    return np.cumsum(xadj).astype(np.int32), adj, cnt.astype(np.int32)


def part_graph_kway(xadj, adjncy, adjwgt, nparts):
    n = C.c_int32(len(xadj) - 1)
    ncon = C.c_int32(1)
    np_ = C.c_int32(nparts)
    obj = C.c_int32(0)
    part = np.zeros(len(xadj) - 1, np.int32)
    xadj, adjncy, adjwgt = (np.ascontiguousarray(a, np.int32) for a in (xadj, adjncy, adjwgt))
    ptr = lambda a: C.cast(a.ctypes.data, C.c_void_p)
    rc = lib().METIS_PartGraphKway(C.addressof(n), C.addressof(ncon), ptr(xadj), ptr(adjncy), None, None,
                                   ptr(adjwgt), C.addressof(np_), None, None, None, C.addressof(obj), ptr(part))
    if rc != 1:  # METIS_OK
        raise RuntimeError("METIS_PartGraphKway returned %d" % rc)
    return part, obj.value


def metis_sort_parts(n, tets, block=16):
    """(part, n_parts, offset, edge cut) by the reference's loop."""
    xadj, adj, w = vertex_graph(n, tets)
    for off in range(block):
        k = (n + block - off - 1) // (block - off)
        if k <= 1:
            return np.zeros(n, np.int32), 1, off, 0
        part, cut = part_graph_kway(xadj, adj, w, k)
        if np.bincount(part, minlength=k).max() <= block:
            return part, k, off, cut
    raise RuntimeError("no offset gave parts within the block size")
