"""CPU tests of the C ABI library: it loads without a GPU, exports every
symbol declared in include/sgipc.h, and its host-side topology functions
(no device work) match the oracle bit for bit; error paths return codes."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "sgipc.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sgipc_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from sgipc import capi
    L = capi.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(names) <= set(capi.SIGNATURES), set(names) - set(capi.SIGNATURES)


def test_version_and_levels():
    from sgipc import capi
    L = capi.lib()
    assert L.sgipc_version() >= 1
    # computeNumLevels (MASPreconditioner.cu:1724-1743): 1331 verts -> 4 levels
    assert L.sgipc_mas_num_levels(1424, 16) == 4
    assert L.sgipc_mas_num_levels(147072, 16) == 6


def test_host_neighbors_and_bank_map_match_oracle(oracle, cube10):
    from sgipc.system import host_neighbors, host_bank_map
    sc = cube10
    ns, nn, nl = host_neighbors(sc.n, sc.tets)
    ons, onn, onl = oracle.vertex_neighbors(sc.n, sc.tets)
    assert np.array_equal(ns, ons) and np.array_equal(nn, onn) and np.array_equal(nl, onl)
    pm, rm = host_bank_map(sc.part, 16)
    opm, orm = oracle.bank_map(sc.part, 16)
    assert np.array_equal(pm, opm) and np.array_equal(rm, orm)


def test_bank_map_on_reference_partition(oracle):
    from sgipc.system import host_bank_map, host_neighbors
    d = np.load(os.path.join(ROOT, "tests", "golden", "high_mat_sorted16.npz"))
    pm, rm = host_bank_map(d["part"], 16)
    opm, orm = oracle.bank_map(d["part"], 16)
    assert np.array_equal(pm, opm) and np.array_equal(rm, orm)
    ns, nn, nl = host_neighbors(d["part"].shape[0], d["tets"])
    ons, onn, onl = oracle.vertex_neighbors(d["part"].shape[0], d["tets"].astype(np.uint32))
    assert np.array_equal(nl, onl) and np.array_equal(nn, onn)


def test_error_codes_without_device():
    from sgipc import capi
    L = capi.lib()
    # unsorted part ids -> invalid argument, message set
    part = np.array([0, 1, 0], np.int32)
    pm = np.zeros(32, np.int32); rm = np.zeros(3, np.int32); M = C.c_int()
    rc = L.sgipc_build_bank_map(3, part.ctypes.data, 16, pm.ctypes.data, 32, C.byref(M), rm.ctypes.data)
    assert rc == 1 and b"ascending" in L.sgipc_last_error(None)
    # a part larger than the bank
    part = np.zeros(17, np.int32)
    rc = L.sgipc_build_bank_map(17, part.ctypes.data, 16, pm.ctypes.data, 32, C.byref(M), rm.ctypes.data)
    assert rc == 1
    # bank size other than 16/32
    rc = L.sgipc_build_bank_map(3, np.zeros(3, np.int32).ctypes.data, 8, pm.ctypes.data, 32, C.byref(M),
                                rm.ctypes.data)
    assert rc == 1
    # capacity too small for the neighbour list
    tets = np.array([[0, 1, 2, 3]], np.uint32)
    st = np.zeros(4, np.uint32); nu = np.zeros(4, np.uint32); lst = np.zeros(4, np.uint32); ln = C.c_int()
    rc = L.sgipc_build_vertex_neighbors(4, 1, tets.ctypes.data, st.ctypes.data, nu.ctypes.data, lst.ctypes.data, 4,
                                        C.byref(ln))
    assert rc == 3
    # null context
    assert L.sgipc_convert(None, 0, 1, None) == 1
