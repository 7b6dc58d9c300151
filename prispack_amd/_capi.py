"""ctypes declarations of include/prispack_amd.h.

The shared library is libprispack_amd.so next to this file (built in-tree by
prispack_amd/build.py).  There is no fallback: a missing library raises.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprispack_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "prispack_amd.h")

PP_OK, PP_ERR_INVALID, PP_ERR_HIP, PP_ERR_NOT_SPD, PP_ERR_ZERO_PIVOT, PP_ERR_ZERO_ROW = 0, 1, 2, 3, 4, 5
PP_ERR_NOT_READY, PP_ERR_BREAKDOWN, PP_ERR_NOT_FINITE, PP_ERR_NOMEM, PP_ERR_UNSUPPORTED = 6, 7, 8, 9, 10
PP_F64, PP_F32 = 0, 1
PP_PC_NONE, PP_PC_ICT, PP_PC_ILUT, PP_PC_ILDLT, PP_PC_IC0, PP_PC_ILU0, PP_PC_AMSED, PP_PC_GMSLR = range(8)
PP_TRSV_AUTO, PP_TRSV_LEVELS, PP_TRSV_SYNCFREE, PP_TRSV_BLOCK, PP_TRSV_PIPE = range(5)
PP_SOLVER_CG, PP_SOLVER_GMRES = 0, 1
PP_FP64, PP_MIXED = 0, 1
PP_LAUNCH_GRAPH, PP_LAUNCH_STREAM = 0, 1
PP_CG_AUTO, PP_CG_REFERENCE, PP_CG_SINGLE_REDUCTION, PP_CG_REDUCED = 0, 1, 2, 3
PP_ORDER_NATURAL, PP_ORDER_MULTICOLOR, PP_ORDER_RCM = 0, 1, 2
PP_LOWRANK_AUTO, PP_LOWRANK_REFERENCE, PP_LOWRANK_DEFLATION, PP_LOWRANK_SUBSPACE = 0, 1, 2, 3

STATUS_NAMES = {0: "PP_OK", 1: "PP_ERR_INVALID", 2: "PP_ERR_HIP", 3: "PP_ERR_NOT_SPD", 4: "PP_ERR_ZERO_PIVOT",
                5: "PP_ERR_ZERO_ROW", 6: "PP_ERR_NOT_READY", 7: "PP_ERR_BREAKDOWN", 8: "PP_ERR_NOT_FINITE",
                9: "PP_ERR_NOMEM", 10: "PP_ERR_UNSUPPORTED"}


class PrisPackError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class PrecondOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("tau", C.c_double), ("lfil", C.c_uint32), ("trsv", C.c_int),
                ("make_f32", C.c_int), ("ordering", C.c_int), ("levels", C.c_uint32), ("ndom", C.c_uint32),
                ("rank", C.c_uint32), ("eps", C.c_double), ("lanczos_steps", C.c_uint32), ("lowrank", C.c_int32),
                ("reserved", C.c_uint32 * 6)]


class PrecondInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("ready", C.c_int), ("n", C.c_uint32), ("annz", C.c_uint32),
                ("mnnz", C.c_uint32), ("fill_ratio", C.c_double), ("levels_lower", C.c_uint32),
                ("levels_upper", C.c_uint32), ("setup_ms", C.c_double), ("fail_row", C.c_uint32),
                ("sweep_variant", C.c_int32), ("fused_levels", C.c_int32), ("trsv", C.c_int32),
                ("phase_ms", C.c_float * 4), ("device_refactors", C.c_uint32), ("reduced_form", C.c_int32),
                ("reduced_form_ms", C.c_float), ("setup_on_device", C.c_int32)]


class SolveOpts(C.Structure):
    _fields_ = [("solver", C.c_int), ("flexible", C.c_int), ("precision", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_uint32), ("restart", C.c_uint32), ("reference_quirks", C.c_int),
                ("check_every", C.c_uint32), ("inner_tol", C.c_double), ("max_outer", C.c_uint32),
                ("record_history", C.c_int), ("launch", C.c_int), ("cg_variant", C.c_int),
                ("reserved", C.c_uint32 * 6)]


class SolveReport(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("final_error", C.c_double),
                ("true_relres", C.c_double), ("outer_iters", C.c_int32), ("inner_iters", C.c_int32),
                ("solve_ms", C.c_double), ("status", C.c_int32), ("nhist", C.c_int32),
                ("cg_variant", C.c_int32), ("inner_fp64_steps", C.c_int32),
                ("reserved", C.c_uint32 * 6)]


vp = C.c_void_p
u32 = C.c_uint32
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)

# name -> (restype, argtypes)
SIGS = {
    "pp_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "pp_ctx_set_stream": (C.c_int, [vp, vp]),
    "pp_ctx_stream": (vp, [vp]),
    "pp_ctx_sync": (C.c_int, [vp]),
    "pp_ctx_destroy": (None, [vp]),
    "pp_last_error": (C.c_char_p, [vp]),
    "pp_global_error": (C.c_char_p, []),
    "pp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pp_malloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    "pp_free": (C.c_int, [vp, vp]),
    "pp_memcpy_h2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "pp_memcpy_d2h": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "pp_memcpy_d2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "pp_memset_zero": (C.c_int, [vp, vp, C.c_size_t]),
    "pp_copy_stream": (C.c_int, [vp, vp, vp, C.c_size_t, C.c_uint32]),
    "pp_csr_create": (C.c_int, [vp, u32, u32, u32p, u32p, f64p, C.POINTER(vp)]),
    "pp_csr_set_values": (C.c_int, [vp, vp, vp, C.c_int]),
    "pp_csr_dims": (C.c_int, [vp, u32p, u32p, u32p]),
    "pp_csr_device_ptrs": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "pp_csr_destroy": (None, [vp]),
    "pp_spmv": (C.c_int, [vp, vp, C.c_int, C.c_double, vp, C.c_double, vp]),
    "pp_spmv_t": (C.c_int, [vp, vp, C.c_int, C.c_double, vp, C.c_double, vp]),
    "pp_dot": (C.c_int, [vp, C.c_int, u32, vp, vp, f64p]),
    "pp_nrm2": (C.c_int, [vp, C.c_int, u32, vp, f64p]),
    "pp_axpby": (C.c_int, [vp, C.c_int, u32, C.c_double, vp, C.c_double, vp, vp]),
    "pp_scale": (C.c_int, [vp, C.c_int, u32, C.c_double, vp]),
    "pp_convert": (C.c_int, [vp, u32, C.c_int, vp, C.c_int, vp]),
    "pp_precond_opts_default": (None, [C.POINTER(PrecondOpts)]),
    "pp_precond_create": (C.c_int, [vp, vp, C.POINTER(PrecondOpts), C.POINTER(vp)]),
    "pp_precond_setup": (C.c_int, [vp, vp]),
    "pp_precond_refactor": (C.c_int, [vp, vp]),
    "pp_precond_apply": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "pp_precond_apply_part": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
    "pp_precond_info_get": (C.c_int, [vp, C.POINTER(PrecondInfo)]),
    "pp_precond_set_sweep_variant": (C.c_int, [vp, C.c_int]),
    "pp_precond_factor": (C.c_int, [vp, C.c_int, u32p, u32p, u32p, f64p, f64p]),
    "pp_precond_destroy": (None, [vp]),
    "pp_factor_host": (C.c_int, [C.POINTER(PrecondOpts), u32, u32p, u32p, f64p, C.POINTER(vp)]),
    "pp_hfactor_get": (C.c_int, [vp, C.c_int, u32p, u32p, u32p, f64p, f64p]),
    "pp_hfactor_info": (C.c_int, [vp, u32p, u32p, u32p, u32p]),
    "pp_hfactor_destroy": (None, [vp]),
    "pp_ml_info": (C.c_int, [vp, u32p, u32p, u32p, u32]),
    "pp_ml_permutation": (C.c_int, [vp, u32p]),
    "pp_gemm_tall_f64": (C.c_int, [vp, u32, u32, u32, vp, vp, vp]),
    "pp_solve_opts_default": (None, [C.POINTER(SolveOpts), C.c_int]),
    "pp_solve": (C.c_int, [vp, vp, vp, C.POINTER(SolveOpts), vp, vp, C.POINTER(SolveReport)]),
    "pp_solve_host": (C.c_int, [vp, vp, vp, C.POINTER(SolveOpts), f64p, f64p, C.POINTER(SolveReport)]),
    "pp_solve_noalloc": (C.c_int, [vp, vp, vp, C.POINTER(SolveOpts), vp, vp, C.POINTER(SolveReport)]),
    "pp_mtx_read": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
    "pp_mtx_dims": (C.c_int, [vp, u32p, u32p, u32p]),
    "pp_mtx_csr": (C.c_int, [vp, u32p, u32p, f64p]),
    "pp_mtx_destroy": (None, [vp]),
    "pp_mtx_read_dense": (C.c_int, [C.c_char_p, u32p, u32p, f64p]),
    "pp_reorder_rcm": (C.c_int, [C.c_uint32, u32p, u32p, u32p]),
    "pp_solve_history": (C.c_int, [vp, f64p, u32, u32p]),
    "pp_adjoint_solve": (C.c_int, [vp, vp, C.POINTER(PrecondOpts), C.POINTER(SolveOpts), vp, vp,
                                   C.POINTER(SolveReport)]),
    "pp_csr_outer_on_pattern": (C.c_int, [vp, vp, C.c_double, vp, vp, vp]),
    "pp_batch_create": (C.c_int, [vp, u32, u32p, u32p, u32p, f64p, C.POINTER(vp)]),
    "pp_batch_set_values": (C.c_int, [vp, vp, vp, C.c_int]),
    "pp_csr_spmv_bytes": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint64)]),
    "pp_csr_spmv_variant": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    "pp_csr_set_spmv_variant": (C.c_int, [vp, C.c_int, C.c_int]),
    "pp_precond_apply_bytes": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_uint64)]),
    "pp_batch_set_ordering": (C.c_int, [vp, vp, C.c_int]),
    "pp_batch_solve": (C.c_int, [vp, vp, C.POINTER(SolveOpts), vp, vp, i32p, C.POINTER(SolveReport)]),
    "pp_batch_solve_t": (C.c_int, [vp, vp, C.POINTER(SolveOpts), vp, vp, i32p, C.POINTER(SolveReport)]),
    "pp_batch_tune": (C.c_int, [vp, vp, C.POINTER(C.c_int)]),
    "pp_batch_outer_on_pattern": (C.c_int, [vp, vp, C.c_double, vp, vp, vp]),
    "pp_batch_destroy": (None, [vp]),
    "pp_generate_laplacian": (C.c_int, [C.c_int] * 3 + [C.c_double] * 4 + [u32p, u32p, u32p, u32p, f64p]),
    "pp_generate_upwind": (C.c_int, [C.c_int] * 3 + [C.c_double] * 3 + [u32p, u32p, u32p, u32p, f64p]),
    "pp_generate_vardiff2d": (C.c_int, [C.c_int, f64p, u32p, u32p, u32p, u32p, f64p]),
    "pp_fill_random": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, f64p]),
}

_lib = None


def header_symbols():
    """Entry points declared in include/prispack_amd.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m prispack_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != PP_OK:
        L = lib()
        msg = (L.pp_last_error(ctx) if ctx else L.pp_global_error()) or b""
        raise PrisPackError(status, msg.decode(errors="replace"))
