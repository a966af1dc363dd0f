"""The reference's file formats through the library's host readers:
gmsh 2.2 tet meshes, .part files and parameterSetting.txt (see
include/sgipc.h).  Arrays come back as numpy arrays."""
import ctypes as C

import numpy as np

from .capi import SgipcError, lib

PARAMETER_NAMES = ("density", "poisson_rate", "friction_rate", "gd_friction_rate", "cloth_thickness",
                   "cloth_youngs_modulus", "bend_youngs_modulus", "cloth_density", "strain_rate", "soft_motion_rate",
                   "collision_detection_buff_scale", "motion_rate", "dt", "pcg_threshold", "newton_threshold",
                   "relative_dhat")


class Parameters(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARAMETER_NAMES]


def _check(rc):
    if rc:
        raise SgipcError(rc, (lib().sgipc_io_last_error() or b"").decode())


def read_msh(path):
    """(x (n,3) float64, tets (T,4) uint32 0-based)"""
    n, T = C.c_int(), C.c_int()
    _check(lib().sgipc_msh_counts(str(path).encode(), C.byref(n), C.byref(T)))
    x = np.zeros((max(n.value, 1), 3)); t = np.zeros((max(T.value, 1), 4), np.uint32)
    _check(lib().sgipc_msh_read(str(path).encode(), x.ctypes.data, n.value, t.ctypes.data, T.value))
    return x[:n.value], t[:T.value]


def write_msh(path, x, tets):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    _check(lib().sgipc_msh_write(str(path).encode(), x.shape[0], x.ctypes.data, t.shape[0], t.ctypes.data))


def read_part(path):
    n = C.c_int()
    rc = lib().sgipc_part_read(str(path).encode(), None, 0, C.byref(n))
    if rc not in (0, 3):  # 3: capacity (the first pass only counts)
        _check(rc)
    p = np.zeros(max(n.value, 1), np.int32)
    _check(lib().sgipc_part_read(str(path).encode(), p.ctypes.data, n.value, C.byref(n)))
    return p[:n.value]


def read_parameters(path):
    p = Parameters()
    _check(lib().sgipc_parameters_read(str(path).encode(), C.byref(p)))
    return {n: getattr(p, n) for n in PARAMETER_NAMES}
