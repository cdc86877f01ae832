"""ctypes binding of libmgx.so (include/mgx.h).

The library must have been built in-tree (compressiontests_amd/libmgx.so); there is no
fallback: importing a solver without it raises, so nothing silently runs elsewhere.

Note on HIP runtimes: the torch ROCm wheel ships its own libamdhip64.so.  If torch is
imported *after* libmgx.so is loaded, a second HIP runtime enters the process.  This
package never imports torch; a caller that needs both must import torch first (libmgx then
binds to the runtime that is already loaded, whose soname is the same).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgx.so")

MGX_OK = 0
MGX_ERR_INVALID_ARGUMENT = 1
MGX_ERR_INVALID_OPERATION = 2
MGX_ERR_NOT_IMPLEMENTED = 3
MGX_ERR_DEVICE = 4
MGX_ERR_OUT_OF_MEMORY = 5
MGX_ERR_NOT_SPD = 6

METHOD_GMG, METHOD_PCG_JACOBI, METHOD_MGPCG = 0, 1, 2
FORMAT_CSR, FORMAT_DUVI = 0, 1
SMOOTHER_JACOBI, SMOOTHER_GS = 0, 1
STOP_MAXNORM, STOP_REL_L2 = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1


class MgxError(RuntimeError):
    """Device or runtime failure (the reference's OpenCLException)."""


class MgxInvalidOperation(MgxError):
    """Solve before Initialize (the reference's InvalidOperationException)."""


class MgxNotImplemented(MgxError, NotImplementedError):
    """A configuration the reference also rejects (NotImplementedException)."""


class MgxInvalidArgument(MgxError, ValueError):
    """Bad sizes or arrays."""


class MgxNotSPD(MgxError, ValueError):
    """The coarsest matrix could not be factorised."""


_ERRORS = {
    MGX_ERR_INVALID_ARGUMENT: MgxInvalidArgument,
    MGX_ERR_INVALID_OPERATION: MgxInvalidOperation,
    MGX_ERR_NOT_IMPLEMENTED: MgxNotImplemented,
    MGX_ERR_DEVICE: MgxError,
    MGX_ERR_OUT_OF_MEMORY: MemoryError,
    MGX_ERR_NOT_SPD: MgxNotSPD,
}


class Options(C.Structure):
    _fields_ = [("method", C.c_int32), ("format", C.c_int32), ("smoother", C.c_int32),
                ("coarse_relaxation", C.c_int32), ("omega", C.c_double), ("max_iterations", C.c_int32),
                ("tolerance", C.c_double), ("stop", C.c_int32), ("depth", C.c_int32),
                ("level_iterations", C.c_int32), ("precision", C.c_int32), ("nrhs", C.c_int32),
                ("duvi_tolerance", C.c_double), ("duvi_mode", C.c_int32), ("inner_iterations", C.c_int32),
                ("inner_tolerance", C.c_double), ("profile", C.c_int32), ("check_every", C.c_int32),
                ("use_graph", C.c_int32), ("block_width", C.c_int32), ("strip_window", C.c_double),
                ("gs_pairs", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual_norm", C.c_double),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double), ("outer_iterations", C.c_int32),
                ("spmv_launches", C.c_int32), ("spmv_ms", C.c_double), ("block_width", C.c_int32),
                ("gs_launches", C.c_int32), ("gs_ms", C.c_double)]


class CsrHost(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("nnz", C.c_int64),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class CsrView(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("nnz", C.c_int64),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class Duvi(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("vi_bytes", C.c_int32), ("nnz", C.c_int64), ("nblocks", C.c_int64),
                ("nuniq", C.c_int64), ("row_blocks", C.c_void_p), ("row_dist", C.c_void_p),
                ("block_col", C.c_void_p), ("dist", C.c_void_p), ("vi", C.c_void_p), ("values", C.c_void_p)]


class HierarchySummary(C.Structure):
    _fields_ = [("version", C.c_int32), ("nlevels", C.c_int32), ("has_grid", C.c_int32), ("has_duvi", C.c_int32),
                ("payload_bytes", C.c_int64)]


class CsrOwned(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class DuviHost(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("vi_bytes", C.c_int32), ("nnz", C.c_int64), ("nblocks", C.c_int64),
                ("nuniq", C.c_int64), ("row_blocks", C.POINTER(C.c_int32)), ("row_dist", C.POINTER(C.c_int32)),
                ("block_col", C.POINTER(C.c_int32)), ("dist", C.POINTER(C.c_uint8)), ("vi", C.c_void_p),
                ("values", C.POINTER(C.c_double))]


_lib = None

# name -> (restype, argtypes)
_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_D = C.c_double
_SIGS = {
    "mgx_last_error": (C.c_char_p, []),
    "mgx_version": (_I32, []),
    "mgx_device_count": (_I32, []),
    "mgx_csr_spmv": (C.c_int, [_I32, _P, _P, _P, _P, _P, _I32, _P]),
    "mgx_csr_spmv_f32": (C.c_int, [_I32, _P, _P, _P, _P, _P, _I32, _P]),
    "mgx_csr_spmm": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P, _P, C.c_int32, _P]),
    "mgx_csr_spmm_f32": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P, _P, C.c_int32, _P]),
    "mgx_time_spmm": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, C.c_int32, C.c_int32,
                                C.POINTER(C.c_double)]),
    "mgx_csr_residual": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P]),
    "mgx_csr_residual_check": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _D, _P, _P]),
    "mgx_jacobi_initial": (C.c_int, [_I32, _P, _P, _P, _P]),
    "mgx_csr_jacobi": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mgx_csr_gs_color": (C.c_int, [_I32, _P, _P, _P, _P, _P, _P, _I32, _P]),
    "mgx_dot": (C.c_int, [_I32, _P, _P, _P, _P]),
    "mgx_dot_noalloc": (C.c_int, [_I32, _P, _P, _P, _P]),
    "mgx_duvi_spmv": (C.c_int, [C.POINTER(Duvi), _P, _P, _I32, _P]),
    "mgx_duvi_residual": (C.c_int, [C.POINTER(Duvi), _P, _P, _P, _P]),
    "mgx_duvi_residual_check": (C.c_int, [C.POINTER(Duvi), _P, _P, _P, _D, _P, _P]),
    "mgx_duvi_jacobi": (C.c_int, [C.POINTER(Duvi), _P, _P, _P, _P, _P]),
    "mgx_duvi_gs_color": (C.c_int, [C.POINTER(Duvi), _I32, _P, _P, _P, _I32, _P]),
    "mgx_duvi_spmv_f32": (C.c_int, [C.POINTER(Duvi), _P, _P, _P, _I32, _P]),
    "mgx_duvi_encode": (C.c_int, [_I32, _P, _P, _P, _D, _I32, C.POINTER(DuviHost)]),
    "mgx_duvi_host_free": (None, [C.POINTER(DuviHost)]),
    "mgx_nodeblock_from_csr": (C.c_int, [_I32, _P, _P, _P, _I32, C.POINTER(_P)]),
    "mgx_nodeblock_spmm": (C.c_int, [_P, _I32, _P, _P, _I32, _P]),
    "mgx_nodeblock_destroy": (C.c_int, [_P]),
    "mgx_nodetransfer_from_csr": (C.c_int, [_I32, _I32, _P, _P, _P, _I32, C.POINTER(_P)]),
    "mgx_nodetransfer_apply": (C.c_int, [_P, _I32, _P, _P, _I32, _P]),
    "mgx_nodetransfer_destroy": (C.c_int, [_P]),
    "mgx_simp_gradient": (C.c_int, [_I32, _I32, _I32, _D, _P, _D, _D, _D, _P, _P, _P]),
    "mgx_matrix_cantilever": (C.c_int, [_I32, _I32, _I32, _D, _D, _P, C.POINTER(_P)]),
    "mgx_matrix_transfer_reference": (C.c_int, [_I32, _I32, _I32, C.POINTER(_P), C.POINTER(_P)]),
    "mgx_matrix_transfer_trilinear": (C.c_int, [_I32, _I32, _I32, C.POINTER(_P), C.POINTER(_P)]),
    "mgx_plate_free_dofs": (C.c_int, [_I32, _I32, _D, _D, C.POINTER(C.c_int32), _P]),
    "mgx_matrix_plate": (C.c_int, [_I32, _I32, _D, _D, _D, _D, _D, _D, C.POINTER(_P), _P]),
    "mgx_matrix_transfer_plate": (C.c_int, [_I32, _I32, _D, _D, C.POINTER(_P), C.POINTER(_P)]),
    "mgx_matrix_cantilever2d": (C.c_int, [_I32, _I32, _D, _D, _P, C.POINTER(_P)]),
    "mgx_matrix_transfer_cantilever2d": (C.c_int, [_I32, _I32, C.POINTER(_P), C.POINTER(_P)]),
    "mgx_matrix_galerkin_hierarchy": (C.c_int, [_I32, _I32, _I32, _D, _D, _P, _I32, C.POINTER(_P)]),
    "mgx_matrix_from_host": (C.c_int, [_I32, _I32, _P, _P, _P, C.POINTER(_P)]),
    "mgx_matrix_view": (C.c_int, [_P, C.POINTER(CsrView)]),
    "mgx_matrix_download": (C.c_int, [_P, _P, _P, _P]),
    "mgx_matrix_destroy": (C.c_int, [_P]),
    "mgx_options_default": (None, [C.POINTER(Options)]),
    "mgx_solver_create": (C.c_int, [C.POINTER(Options), C.POINTER(_P)]),
    "mgx_solver_set_level": (C.c_int, [_P, _I32, C.POINTER(CsrHost), C.POINTER(CsrHost), C.POINTER(CsrHost)]),
    "mgx_solver_set_level_device": (C.c_int, [_P, _I32, _P, _P, _P]),
    "mgx_solver_set_rhs": (C.c_int, [_P, _P, _I32, _I32]),
    "mgx_solver_initialize": (C.c_int, [_P]),
    "mgx_solver_solve": (C.c_int, [_P, _P, _I32, C.POINTER(Stats)]),
    "mgx_solver_solve_device": (C.c_int, [_P, _P, _P, _I32, C.POINTER(Stats)]),
    "mgx_solver_block_width": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "mgx_solver_level_layout": (C.c_int, [_P, _I32, C.POINTER(C.c_int32)]),
    "mgx_solver_transfer_layout": (C.c_int, [_P, _I32, C.POINTER(C.c_int32)]),
    "mgx_solver_level_colors": (C.c_int, [_P, _I32, C.POINTER(C.c_int32)]),
    "mgx_solver_level_groups": (C.c_int, [_P, _I32, C.POINTER(C.c_int32), _P]),
    "mgx_solver_gs_order": (C.c_int, [_P, _I32, _P]),
    "mgx_solver_gs_pairs": (C.c_int, [_P, _I32, C.POINTER(C.c_int32)]),
    "mgx_solver_strip_window": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mgx_solver_release": (C.c_int, [_P]),
    "mgx_device_malloc": (C.c_int, [C.POINTER(_P), _U64]),
    "mgx_device_free": (C.c_int, [_P]),
    "mgx_memcpy_htod": (C.c_int, [_P, _P, _U64]),
    "mgx_memcpy_dtoh": (C.c_int, [_P, _P, _U64]),
    "mgx_memcpy_dtod": (C.c_int, [_P, _P, _U64]),
    "mgx_memset": (C.c_int, [_P, _I32, _U64]),
    "mgx_device_synchronize": (C.c_int, []),
    "mgx_time_csr_kernel": (C.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, C.POINTER(_D)]),
    "mgx_time_duvi_spmv": (C.c_int, [C.POINTER(Duvi), _P, _P, _I32, _I32, _I32, C.POINTER(_D)]),
    "mgx_time_copy_kernel": (C.c_int, [_P, _P, _U64, _I32, _I32, C.POINTER(_D)]),
    "mgx_dense_block_apply": (C.c_int, [_I32, _P, _I32, _P, _P, _I32, _P]),
    "mgx_time_read_kernel": (C.c_int, [_P, _U64, _P, _I32, _I32, C.POINTER(_D)]),
    "mgx_hierarchy_write": (C.c_int, [C.c_char_p, _I32, C.POINTER(CsrHost), C.POINTER(CsrHost), C.POINTER(CsrHost),
                                      _P, _I32, _D]),
    "mgx_hierarchy_info": (C.c_int, [C.c_char_p, C.POINTER(HierarchySummary)]),
    "mgx_hierarchy_read_csr": (C.c_int, [C.c_char_p, _I32, _I32, C.POINTER(CsrOwned)]),
    "mgx_csr_owned_free": (None, [C.POINTER(CsrOwned)]),
    "mgx_plate_system_host": (C.c_int, [_I32, _I32, _D, _D, _D, _D, _D, _D, C.POINTER(CsrOwned), _P]),
    "mgx_plate_transfer_host": (C.c_int, [_I32, _I32, _D, _D, C.POINTER(CsrOwned), C.POINTER(CsrOwned)]),
    "mgx_cantilever2d_system_host": (C.c_int, [_I32, _I32, _D, _D, _P, C.POINTER(CsrOwned)]),
    "mgx_cantilever2d_transfer_host": (C.c_int, [_I32, _I32, C.POINTER(CsrOwned), C.POINTER(CsrOwned)]),
    "mgx_hierarchy_read_duvi": (C.c_int, [C.c_char_p, _I32, _I32, C.POINTER(DuviHost)]),
    "mgx_hierarchy_read_grid": (C.c_int, [C.c_char_p, _I32, _P]),
    "mgx_hierarchy_save": (C.c_int, [C.c_char_p, _I32, _P, _P, _P, _I32, _D]),
    "mgx_hierarchy_load": (C.c_int, [C.c_char_p, _I32, C.POINTER(C.c_int32), _P, _P, _P]),
    "mgx_matrix_grid": (C.c_int, [_P, _P]),
    "mgx_matrix_set_grid": (C.c_int, [_P, _P]),
    "mgx_rng_uniform": (C.c_int, [_U64, _I64, _P]),
    "mgx_rng_normal": (C.c_int, [_U64, _I64, _P]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libmgx.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MgxError("libmgx.so is not built: run `python -m compressiontests_amd.build` "
                           "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != MGX_OK:
        msg = lib().mgx_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, MgxError)(msg)
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def default_options(**kw):
    o = Options()
    lib().mgx_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def device_count():
    return lib().mgx_device_count()
