"""ctypes binding of ``libpetgem_amd.so`` (the C ABI declared in ``include/petgem_amd.h``).

The library is built in-tree by ``petgem_amd/build.py``.  There is no
fallback: if the library is missing, every entry point raises
``NativeLibraryMissing``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpetgem_amd.so")

PG_OK = 0
PG_ERR_ARG, PG_ERR_HIP, PG_ERR_UNSUPPORTED, PG_ERR_NOT_FOUND, PG_ERR_BREAKDOWN = 1, 2, 3, 4, 5
PG_KSP_COCG, PG_KSP_GMRES = 0, 1
PG_PC_NONE, PG_PC_JACOBI, PG_PC_BJACOBI = 0, 1, 2
PG_PREC_FP64, PG_PREC_MIXED = 0, 1

# (name, restype, argtypes) for every exported entry point
_P = ctypes.c_void_p
_I32, _I64, _D, _SZ = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t


class PlanDev(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int32), ("n_edges", _I64), ("n_faces", _I64), ("n_elems", _I64),
                ("ndof", _I64), ("nnz", _I64), ("npairs", _I64),
                ("rowptr", _P), ("cols", _P), ("pair_slot0", _P), ("pair_rowlen", _P),
                ("pair_cbeg", _P), ("pair_ab", _P), ("contrib", _P), ("entity_pair", _P),
                ("entity_row_max", ctypes.c_int32)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("ksp", _I32), ("pc", _I32), ("precision", _I32), ("restart", _I32),
                ("maxit", _I32), ("check_every", _I32), ("rtol", _D), ("inner_rtol", _D),
                ("block_ptr", _P), ("nblocks", _I64), ("spmv", _I32), ("stall_iters", _I32),
                ("entity_order", _P), ("eop", _P), ("like", _P)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", _I32), ("refinements", _I32), ("status", _I32),
                ("fp64_fallback", _I32),
                ("relres", _D), ("spmv_ms", _D), ("spmv_samples", _I32), ("pad2", _I32),
                ("spmv_first_ms", _D)]


SIGNATURES = {
    "pg_abi_version": (ctypes.c_int, []),
    "pg_last_error": (ctypes.c_char_p, []),
    "pg_tables_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "pg_tables_destroy": (ctypes.c_int, [_P]),
    "pg_tables_host": (ctypes.c_int, [ctypes.c_int, _P, _I64]),
    "pg_element_workspace_bytes": (_SZ, [_I64]),
    "pg_element_geometry": (ctypes.c_int, [ctypes.c_int, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pg_element_orientation": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P, _P]),
    "pg_element_matrices": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    "pg_plan_build": (ctypes.c_int, [ctypes.c_int, _I64, _I64, _I64, _P, _P, ctypes.POINTER(_P)]),
    "pg_plan_info": (ctypes.c_int, [_P, _P]),
    "pg_plan_export": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "pg_plan_entity_pairs": (ctypes.c_int, [_P, _P, _P]),
    "pg_plan_destroy": (ctypes.c_int, [_P]),
    "pg_assemble": (ctypes.c_int, [_P, ctypes.POINTER(PlanDev), _P, _D, _P, _P, _P]),
    "pg_assemble_km": (ctypes.c_int, [_P, ctypes.POINTER(PlanDev), _P, _P, _P, _P, _P]),
    "pg_apply_dirichlet": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P]),
    "pg_locate_points": (ctypes.c_int, [_I64, _P, _I64, _P, _D, _P, _P]),
    "pg_source_vector": (ctypes.c_int, [ctypes.c_int, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _D,
                                        _P, _P]),
    "pg_receiver_fields": (ctypes.c_int, [ctypes.c_int, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                          _D, _P, _P]),
    "pg_adjoint_rhs": (ctypes.c_int, [ctypes.c_int, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _P]),
    "pg_adjoint_gradient": (ctypes.c_int, [_P, _I64, _P, _P, ctypes.c_int, _I64, _P, _P, _P, _P,
                                           _P]),
    "pg_adjoint_gradient_rows": (ctypes.c_int, [_P, _I64, _P, _P, ctypes.c_int, _I64, _P, _P, _P,
                                                _P, _P]),
    "pg_adjoint_gradient_noalloc": (ctypes.c_int, [_P, _I64, _P, _P, ctypes.c_int, _I64, _P, _P,
                                                   _P, _P, _P]),
    "pg_eop_create": (ctypes.c_int, [ctypes.c_int, _I64, _P, _P, _I64, _P, _I64, _P, _P, _D,
                                     ctypes.POINTER(_P), _P]),
    "pg_eop_create_multi": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P,
                                           _P, _P, _P, ctypes.POINTER(_P), _P]),
    "pg_eop_refresh": (ctypes.c_int, [_P, _P, _D, _P]),
    "pg_eop_apply_z": (ctypes.c_int, [_P, _P, _P, _P]),
    "pg_eop_apply_c": (ctypes.c_int, [_P, _P, _P, _P]),
    "pg_eop_destroy": (ctypes.c_int, [_P]),
    "pg_spmv_z": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P]),
    "pg_spmv_c": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P]),
    "pg_demote_z2c": (ctypes.c_int, [_I64, _P, _P, _P]),
    "pg_solver_create": (ctypes.c_int, [_I64, _I64, _P, _P, _P, ctypes.POINTER(SolveOpts),
                                        ctypes.POINTER(_P), _P]),
    "pg_solver_solve": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(SolveInfo), _P]),
    "pg_solver_destroy": (ctypes.c_int, [_P]),
    "pg_solver_refresh": (ctypes.c_int, [_P, _P]),
    "pg_batch_create": (ctypes.c_int, [_I64, _I64, _P, _P, _P, _P, ctypes.c_int, _P, _D,
                                       ctypes.c_int, ctypes.POINTER(_P), _P]),
    "pg_batch_solve": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "pg_batch_refresh": (ctypes.c_int, [_P, _P]),
    "pg_batch_set_mixed": (ctypes.c_int, [_P, _D, ctypes.c_int, _P]),
    "pg_batch_set_block_jacobi": (ctypes.c_int, [_P, _P]),
    "pg_batch_set_entities": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "pg_batch_set_element_operator": (ctypes.c_int, [_P, _P, _P, _P]),
    "pg_batch_destroy": (ctypes.c_int, [_P]),
    "pg_solver_profile": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P, _P]),
    "pg_multi_create": (ctypes.c_int, [ctypes.c_int, _P, _P, _P, _P, _P, _D, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P), _P]),
    "pg_multi_solve": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "pg_multi_set_element_operator": (ctypes.c_int, [_P, _P, _P]),
    "pg_multi_set_mixed": (ctypes.c_int, [_P, _D, ctypes.c_int, _P]),
    "pg_multi_destroy": (ctypes.c_int, [_P]),
}


class NativeLibraryMissing(RuntimeError):
    pass


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"petgem_amd error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the native library (once) and declare every entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -m petgem_amd.build` "
                "(there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    if rc != PG_OK:
        raise PgError(rc, lib().pg_last_error().decode(errors="replace"))
    return rc


def ptr(t):
    """Device/host pointer of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
