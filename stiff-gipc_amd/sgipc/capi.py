"""ctypes binding of libsgipc.so (the C ABI declared in include/sgipc.h).

Device memory is passed as raw pointers (torch tensors' data_ptr()).  The
library must exist: nothing here falls back to a CPU implementation.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libsgipc.so")

OK = 0
ERR = {1: "invalid argument", 2: "HIP error", 3: "capacity exceeded", 4: "invalid state", 5: "internal error"}
PRECOND_NONE, PRECOND_DIAG, PRECOND_MAS, PRECOND_MAS_FP32 = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1
STOP_REF_RZ, STOP_TRUE_REL = 0, 1
BANK_INVERSE_VALU64, BANK_INVERSE_MFMA64, BANK_INVERSE_MFMA32_REFINED = 0, 1, 2
BANK_STORE_F32, BANK_STORE_F16 = 0, 1
SORT_OWN, SORT_ROCPRIM = 0, 1
PROF_SPMV, PROF_MAS_SOLVE, PROF_FEM, PROF_CONVERT, PROF_PRECOND_BUILD, PROF_SOLVE, PROF_BARRIER = range(7)


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("stop_rule", C.c_int), ("tol", C.c_double),
                ("max_iter_ratio", C.c_double), ("max_iter", C.c_int), ("check_every", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("refinements", C.c_int),
                ("rel_residual", C.c_double), ("rz0", C.c_double)]


class SgipcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sgipc error %d (%s): %s" % (code, ERR.get(code, "?"), msg))
        self.code = code


_lib = None

# exported symbols and their signatures (must match include/sgipc.h)
P, I, LL, D, SZ = C.c_void_p, C.c_int, C.c_longlong, C.c_double, C.c_size_t
SIGNATURES = {
    "sgipc_version": ([], I),
    "sgipc_msh_counts": ([C.c_char_p, C.POINTER(I), C.POINTER(I)], I),
    "sgipc_msh_read": ([C.c_char_p, P, I, P, I], I),
    "sgipc_msh_write": ([C.c_char_p, I, P, I, P], I),
    "sgipc_part_read": ([C.c_char_p, P, I, C.POINTER(I)], I),
    "sgipc_parameters_read": ([C.c_char_p, P], I),
    "sgipc_io_last_error": ([], C.c_char_p),
    "sgipc_ctx_create": ([I, C.POINTER(P)], I),
    "sgipc_ctx_destroy": ([P], I),
    "sgipc_ctx_set_stream": ([P, P], I),
    "sgipc_ctx_synchronize": ([P], I),
    "sgipc_last_error": ([P], C.c_char_p),
    "sgipc_build_vertex_neighbors": ([I, I, P, P, P, P, I, C.POINTER(I)], I),
    "sgipc_build_bank_map": ([I, P, I, P, I, C.POINTER(I), P], I),
    "sgipc_mas_num_levels": ([I, I], I),
    "sgipc_partition_mesh": ([I, I, P, I, P, P, C.POINTER(I), C.POINTER(I)], I),
    "sgipc_partition_edge_cut": ([I, I, P, P, C.POINTER(LL)], I),
    "sgipc_set_fem_mesh": ([P, I, I, P], I),
    "sgipc_set_mas_topology": ([P, I, I, P, P, P, I, P, I, P, I], I),
    "sgipc_triplets_view": ([P, P, P, P, LL], I),
    "sgipc_kinetic_gradient": ([P, I, P, P, P, P], I),
    "sgipc_assemble_fem": ([P, I, P, P, P, P, P, P, D, P, I, LL, P], I),
    "sgipc_mass_triplets": ([P, I, P, I, LL], I),
    "sgipc_fem_rhs": ([P, I, P, P, P, P], I),
    "sgipc_ground_detect": ([P, I, P, P, P, D, D, P, C.POINTER(I)], I),
    "sgipc_assemble_ground": ([P, I, P, P, P, D, D, D, P, I, LL, P], I),
    "sgipc_ground_energy": ([P, I, P, P, P, D, D, D, C.POINTER(D)], I),
    "sgipc_ground_step_bound": ([P, I, P, P, P, P, D, D, C.POINTER(D)], I),
    "sgipc_set_fem_projection": ([P, I], I),
    "sgipc_abd_setup": ([P, I, I, P, P, P, P, P], I),
    "sgipc_abd_positions": ([P, P, P], I),
    "sgipc_abd_assemble": ([P, P, P, D, D, LL, P], I),
    "sgipc_abd_contact_gradient": ([P, P, P], I),
    "sgipc_abd_map_contacts": ([P, LL, LL, LL, C.POINTER(LL)], I),
    "sgipc_abd_energy": ([P, P, P, D, D, C.POINTER(D)], I),
    "sgipc_set_deterministic": ([P, I], I),
    "sgipc_broad_phase_ccd": ([P, I, P, I, P, I, P, P, P, D, P, P, D, P, I, C.POINTER(I), P, I, C.POINTER(I)], I),
    "sgipc_ccd_step_bound": ([P, I, P, I, P, P, P, D, C.POINTER(D)], I),
    "sgipc_broad_phase": ([P, I, P, I, P, I, P, P, P, P, D, P, I, C.POINTER(I), P, I, C.POINTER(I)], I),
    "sgipc_solve_batched": ([P, I, P, P, P, P, P], I),
    "sgipc_autotune": ([P, P, I, P, I, C.POINTER(I), C.POINTER(I)], I),
    "sgipc_set_schedule": ([P, I, I], I),
    "sgipc_set_batch_mode": ([P, I, I, I], I),
    "sgipc_set_sort": ([P, I], I),
    "sgipc_set_inner_tol": ([P, D], I),
    "sgipc_step_forward": ([P, I, P, P, P, P, D, I], I),
    "sgipc_cfl_largest_speed": ([P, I, P, P, C.POINTER(D)], I),
    "sgipc_sort_pairs": ([P, P, P, P, P, LL, I, I, I], I),
    "sgipc_get_batch_mode": ([P, C.POINTER(I), C.POINTER(I)], I),
    "sgipc_fem_param_adjoint": ([P, I, P, P, P, P, D, P, P, P, P], I),
    "sgipc_classify_contacts": ([P, I, P, I, P, P, P, D, P, P, I, C.POINTER(I), P], I),
    "sgipc_assemble_barrier": ([P, I, P, P, P, P, P, D, D, P, I, LL, P], I),
    "sgipc_convert": ([P, LL, I, C.POINTER(I)], I),
    "sgipc_spmv": ([P, D, P, D, P], I),
    "sgipc_precond_build": ([P, I, I, I, P, I], I),
    "sgipc_precond_apply": ([P, P, P], I),
    "sgipc_solve": ([P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)], I),
    "sgipc_kinetic_energy": ([P, I, P, P, P, C.POINTER(D)], I),
    "sgipc_fem_energy": ([P, I, P, P, P, P, P, P, D, C.POINTER(D)], I),
    "sgipc_fem_energy_noalloc": ([P, I, P, P, P, P, P, P, D, P], I),
    "sgipc_barrier_energy": ([P, I, P, P, P, D, D, C.POINTER(D)], I),
    "sgipc_profile_enable": ([P, I], I),
    "sgipc_profile_read": ([P, I, C.POINTER(LL), C.POINTER(D), I], I),
    "sgipc_mas_info": ([P, C.POINTER(I), C.POINTER(I), P], I),
    "sgipc_mas_tables": ([P, P, P, P, P, P], I),
    "sgipc_mas_bank_inverse": ([P, I, P], I),
    "sgipc_mas_bank_inverse_f32": ([P, I, P], I),
    "sgipc_set_bank_inverse": ([P, I], I),
    "sgipc_mas_bank_inverse_f16": ([P, I, P], I),
    "sgipc_set_bank_storage": ([P, I], I),
    "sgipc_malloc": ([P, C.POINTER(P), SZ], I),
    "sgipc_free": ([P, P], I),
    "sgipc_memcpy_h2d": ([P, P, P, SZ], I),
    "sgipc_memcpy_d2h": ([P, P, P, SZ], I),
    "sgipc_memcpy_d2d": ([P, P, P, SZ], I),
}


def lib():
    """Load libsgipc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsgipc.so not found at %s: run __graft_entry__.build() or "
                               "make -C stiff-gipc_amd" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != OK:
        msg = lib().sgipc_last_error(ctx)
        raise SgipcError(rc, msg.decode() if msg else "")
    return rc
