"""ctypes binding of libripley_hip.so (the C ABI in include/ripley_hip.h).

The product path has no CPU fallback: if the library is missing this module
raises ImportError-like RuntimeError on first use.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libripley_hip.so")

RG_OK = 0
RG_MAX_ITER_REACHED = 1
RG_INPUT_ERROR = 2
RG_MEMORY_ERROR = 3
RG_BREAKDOWN = 4
RG_NEGATIVE_NORM = 5
RG_DIVERGENCE = 6
RG_VALUE_ERROR = 20
RG_STATE_ERROR = 21
RG_DEVICE_ERROR = 22

RG_MATRIX_SYMMETRIC = 1
RG_COEF_EMPTY, RG_COEF_ELEMENT, RG_COEF_EXPANDED = 0, 1, 2
(RG_METHOD_PCG, RG_METHOD_PCG_PLAIN, RG_METHOD_PCG_MIXED, RG_METHOD_FCG_LEARNED, RG_METHOD_BICGSTAB,
 RG_METHOD_MINRES) = 0, 1, 2, 3, 4, 5
RG_CRITERION_PASO, RG_CRITERION_TRUE_RESIDUAL = 0, 1
RG_PRECONDITIONER_JACOBI, RG_PRECONDITIONER_GAUSS_SEIDEL, RG_PRECONDITIONER_ILU0, RG_PRECONDITIONER_NONE = 0, 1, 2, 3
RG_BYTEORDER_LITTLE_ENDIAN, RG_BYTEORDER_BIG_ENDIAN = 1234, 4321
RG_BYTEORDER_NATIVE = RG_BYTEORDER_LITTLE_ENDIAN
RG_DATATYPE_INT32, RG_DATATYPE_FLOAT32, RG_DATATYPE_FLOAT64 = 1, 2, 3
RG_SCRATCH_BYTES = 16 << 20

# every function the header declares (checked by the CPU test suite)
EXPORTS = [
    "rg_last_error", "rg_version",
    "rg_context_create", "rg_context_destroy", "rg_context_set_stream", "rg_context_synchronize",
    "rg_matrix_create", "rg_matrix_destroy", "rg_matrix_get_info", "rg_matrix_set_values",
    "rg_matrix_reset_values", "rg_matrix_add_element_blocks", "rg_matrix_assemble",
    "rg_matrix_vector", "rg_matrix_ypAx", "rg_matrix_ypAx_host",
    "rg_matrix_nullify_rows_and_cols", "rg_matrix_nullify_rows_and_cols_host",
    "rg_matrix_export_csr", "rg_matrix_save_mm", "rg_options_default", "rg_solve", "rg_solve_host",
    "rg_lpc_num_params", "rg_lpc_create", "rg_lpc_destroy", "rg_lpc_set_params", "rg_lpc_get_params",
    "rg_lpc_apply", "rg_lpc_gradients", "rg_lpc_train_step", "rg_matrix_set_learned_preconditioner",
    "rg_solve_multi", "rg_matrix_vector_multi", "rg_multi_autotune", "rg_multi_set_shape",
    "rg_batch_create", "rg_batch_destroy", "rg_matrix_assemble_boundary", "rg_matrix_add_dirac",
    "rg_batch_offsets", "rg_batch_vector", "rg_batch_set_tile", "rg_batch_autotune", "rg_batch_solve",
    "rg_matrix_solve_preconditioner", "rg_matrix_coloring", "rg_read_binary_grid", "rg_write_binary_grid",
    "rg_matrix_set_multilevel", "rg_matrix_coarse_operator", "rg_lpc_set_loss_operator",
    "rg_matrix_smooth", "rg_matrix_add_element_blocks_noalloc", "rg_matrix_vcycle",
]


class RgCoef(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("mode", ctypes.c_int32), ("elem_stride", ctypes.c_int32),
                ("isotropic", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RgPdeCoefs(ctypes.Structure):
    _fields_ = [(k, RgCoef) for k in "ABCDXY"]


class RgOptions(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("inner_tolerance", ctypes.c_double),
                ("iter_max", ctypes.c_int32), ("sweeps", ctypes.c_int32), ("method", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("accept_failed_convergence", ctypes.c_int32),
                ("check_interval", ctypes.c_int32), ("criterion", ctypes.c_int32),
                ("preconditioner", ctypes.c_int32)]


class RgDiagnostics(ctypes.Structure):
    _fields_ = [("num_iter", ctypes.c_int32), ("converged", ctypes.c_int32), ("status", ctypes.c_int32),
                ("outer_steps", ctypes.c_int32), ("residual_norm", ctypes.c_double),
                ("set_up_time", ctypes.c_double), ("net_time", ctypes.c_double), ("time", ctypes.c_double),
                ("fallbacks", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RipleyHipError(RuntimeError):
    """Error raised by the library; ``code`` is the C status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class PasoException(RipleyHipError):
    """Solver failures, named after the reference's exception class."""


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libripley_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "rg_last_error": (ctypes.c_char_p, []),
        "rg_version": (ctypes.c_char_p, []),
        "rg_context_create": (I, [I, P, PP]),
        "rg_context_destroy": (I, [P]),
        "rg_context_set_stream": (I, [P, P]),
        "rg_context_synchronize": (I, [P]),
        "rg_matrix_create": (I, [P, I, P, P, I, I, PP]),
        "rg_matrix_destroy": (I, [P]),
        "rg_matrix_get_info": (I, [P, P, P, P, P]),
        "rg_matrix_set_values": (I, [P, D]),
        "rg_matrix_reset_values": (I, [P, I]),
        "rg_matrix_add_element_blocks": (I, [P, I64, I, P, I, P]),
        "rg_matrix_add_element_blocks_noalloc": (I, [P, I64, I, P, I, P]),
        "rg_matrix_assemble": (I, [P, ctypes.POINTER(RgPdeCoefs), I, P]),
        "rg_matrix_vector": (I, [P, D, P, D, P]),
        "rg_matrix_ypAx": (I, [P, P, P]),
        "rg_matrix_ypAx_host": (I, [P, P, P]),
        "rg_matrix_nullify_rows_and_cols": (I, [P, P, P, D]),
        "rg_matrix_nullify_rows_and_cols_host": (I, [P, P, P, D]),
        "rg_matrix_export_csr": (I, [P, P, P, P]),
        "rg_matrix_save_mm": (I, [P, ctypes.c_char_p]),
        "rg_options_default": (None, [ctypes.POINTER(RgOptions)]),
        "rg_solve": (I, [P, P, P, ctypes.POINTER(RgOptions), ctypes.POINTER(RgDiagnostics)]),
        "rg_solve_host": (I, [P, P, P, ctypes.POINTER(RgOptions), ctypes.POINTER(RgDiagnostics)]),
        "rg_matrix_solve_preconditioner": (I, [P, ctypes.POINTER(RgOptions), P, P]),
        "rg_matrix_coloring": (I, [P, P, ctypes.POINTER(ctypes.c_int)]),
        "rg_read_binary_grid": (I, [ctypes.c_char_p, I, P, I, I, P, P, P, P, I, I, P]),
        "rg_write_binary_grid": (I, [ctypes.c_char_p, I, P, P, I, I]),
        "rg_lpc_num_params": (I64, []),
        "rg_lpc_create": (I, [P, P, I, PP]),
        "rg_lpc_destroy": (I, [P]),
        "rg_lpc_set_params": (I, [P, P]),
        "rg_lpc_get_params": (I, [P, P]),
        "rg_lpc_apply": (I, [P, P, P, P, P]),
        "rg_lpc_gradients": (I, [P, P, P, P, P, P, P]),
        "rg_lpc_train_step": (I, [P, P, P, P, P, ctypes.c_float, P]),
        "rg_matrix_set_learned_preconditioner": (I, [P, P, P]),
        "rg_matrix_set_multilevel": (I, [P, I, I, I]),
        "rg_matrix_coarse_operator": (I, [P, I, PP]),
        "rg_lpc_set_loss_operator": (I, [P, P]),
        "rg_matrix_smooth": (I, [P, P, P, I]),
        "rg_matrix_vcycle": (I, [P, P, P]),
        "rg_solve_multi": (I, [P, P, P, I, ctypes.POINTER(RgOptions), P]),
        "rg_matrix_vector_multi": (I, [P, P, P, I]),
        "rg_multi_autotune": (I, [P, I, I, I, P, P]),
        "rg_multi_set_shape": (I, [P, I]),
        "rg_matrix_assemble_boundary": (I, [P, ctypes.POINTER(RgCoef), ctypes.POINTER(RgCoef), I, P]),
        "rg_matrix_add_dirac": (I, [P, I64, P, P, P, P]),
        "rg_batch_create": (I, [P, I, P, PP]),
        "rg_batch_destroy": (I, [P]),
        "rg_batch_offsets": (I, [P, P]),
        "rg_batch_vector": (I, [P, P, P]),
        "rg_batch_set_tile": (I, [P, I]),
        "rg_batch_autotune": (I, [P, I, I, P, P]),
        "rg_batch_solve": (I, [P, P, P, ctypes.POINTER(RgOptions), P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc == RG_OK:
        return
    msg = load().rg_last_error().decode(errors="replace")
    if rc <= RG_DIVERGENCE:
        raise PasoException(rc, msg)
    if rc == RG_VALUE_ERROR:
        raise RipleyHipError(rc, msg)
    raise RipleyHipError(rc, msg)
