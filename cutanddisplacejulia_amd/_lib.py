"""ctypes binding of libcdjhip.so (include/cdj.h).

There is no CPU fallback: if the library is missing or no HIP device is present, calls raise.
"""
import atexit
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcdjhip.so")

_d = ctypes.c_double
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_p = ctypes.c_void_p

CDJ_OK = 0
CDJ_ERR_DIMS = 1
CDJ_ERR_HALFSPACE = 2
CDJ_ERR_DEGENERATE = 3
CDJ_ERR_DEVICE = 4
CDJ_ERR_NOMEM = 5
CDJ_ERR_NOT_CONVERGED = 6
CDJ_ERR_ARG = 7
CDJ_ERR_SINGULAR = 8

CDJ_PREC_FP64 = 0
CDJ_PREC_MIXED32 = 1
CDJ_PREC_LEARNED = 2
CDJ_PREC_MIXED16 = 3
CDJ_PC_NONE = 0
CDJ_PC_JACOBI = 1
CDJ_PC_LEARNED = 2
CDJ_PC_LU32 = 3
CDJ_PC_LEARNED_BF16 = 4
CDJ_ASM_FP64 = 0
CDJ_ASM_MIXED = 1

# every symbol declared in include/cdj.h
EXPORTS = ["cdj_create", "cdj_destroy", "cdj_last_error", "cdj_last_kernel_ms", "cdj_td_influence", "cdj_td_field",
           "cdj_traction_influence", "cdj_slip_calculator", "cdj_solve", "cdj_assemble_dev",
           "cdj_td_influence_dev", "cdj_lu_factor_dev", "cdj_lu_factor_batched_dev", "cdj_lu_solve_dev",
           "cdj_mlp_create", "cdj_mlp_destroy",
           "cdj_mlp_nparams", "cdj_mlp_get_params", "cdj_mlp_set_params", "cdj_mlp_train_dev", "cdj_mlp_loss_grad",
           "cdj_precond_solve_dev", "cdj_last_near_entries", "cdj_batched_slip", "cdj_batch_tuning",
           "cdj_traction_rows", "cdj_lstsq", "cdj_crack_tip_elements", "cdj_stress_intensity",
           "cdj_stress_intensity_dev", "cdj_slip_friction", "cdj_slip_fluid_volume", "cdj_mlp_forward",
           "cdj_assemble_dev_noalloc", "cdj_last_edge_stats", "cdj_remesh_isotropic", "cdj_remesh_take"]


class CdjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("precision", _i32), ("tile", _i32), ("tol", _d), ("max_iter", _i32), ("assembly", _i32),
                ("near_factor", _d)]


class Report(ctypes.Structure):
    _fields_ = [("path", _i32), ("iterations", _i32), ("residual", _d), ("assembly_ms", _d), ("solve_ms", _d),
                ("near_entries", _i64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None
_CTX = None
_LOCK = threading.Lock()


def load(path=LIB_PATH):
    """Load the shared library and declare every entry point's signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise CdjError(CDJ_ERR_DEVICE, "libcdjhip.so is not built (%s); run cutanddisplacejulia_amd.build" % path)
    # torch ships its own HIP runtime (same soname, libamdhip64.so.7, but opened by the unversioned name).  Loading
    # it first makes libcdjhip.so bind to that copy, so the process holds one HIP runtime; loading ours first would
    # make torch open a second one, which then sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    L.cdj_create.argtypes = [ctypes.c_int, ctypes.POINTER(_p)]
    L.cdj_destroy.argtypes = [_p]
    L.cdj_last_error.argtypes = [_p]
    L.cdj_last_error.restype = ctypes.c_char_p
    L.cdj_last_kernel_ms.argtypes = [_p]
    L.cdj_last_near_entries.argtypes = [_p]
    L.cdj_last_edge_stats.argtypes = [_p, _p, _p]
    L.cdj_batched_slip.argtypes = [_p, _p, _i64, _d, _d, _d, _i32, _p, _p, _p]
    L.cdj_batch_tuning.argtypes = [_p, _p, _p]
    L.cdj_traction_rows.argtypes = [_p, _p, _p, _i64, _p, _p, _p, _i64, _d, _d, _d, _i32, _p, _p]
    L.cdj_lstsq.argtypes = [_p, _p, _i64, _i64, _p, _i64, _p, _p]
    L.cdj_mlp_forward.argtypes = [_p, _p, _p, _p, _p, _i64, _i32, _p]
    L.cdj_slip_friction.argtypes = [_p, _p, _p, _p, _i64, _d, _d, _d, _p, _p, _i32, _p, _p, _p, _p, _p, _p, _p]
    L.cdj_slip_fluid_volume.argtypes = [_p, _p, _p, _p, _i64, _d, _d, _d, _p, _p, _i32, _p, _p, _d, _p, _p, _p, _p,
                                        _p]
    L.cdj_crack_tip_elements.argtypes = [_p, _p, _p, _p, _p, _p, _i64, _p, _p]
    L.cdj_stress_intensity.argtypes = [_p, _p, _p, _p, _i64, _d, _d, _p, _p, _p, _p]
    L.cdj_stress_intensity_dev.argtypes = [_p, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _d, _d, _p, _p, _p, _p]
    td_args = [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _d, _d, _i32, _i32, _i32, _p, _p]
    L.cdj_td_influence.argtypes = td_args
    L.cdj_td_field.argtypes = td_args
    L.cdj_td_influence_dev.argtypes = td_args + [_p]
    L.cdj_traction_influence.argtypes = [_p, _p, _p, _p, _p, _p, _i64, _d, _d, _d, _i32, _p, _p, _p]
    L.cdj_slip_calculator.argtypes = [_p, _p, _p, _p, _i64, _d, _d, _d, _p, _p, _i32, _p, _i64, _p, _p, _p, _p, _p]
    L.cdj_solve.argtypes = [_p, _p, _p, _i64, _i64, _p, _p, _p]
    L.cdj_assemble_dev.argtypes = [_p, _p, _p, _p, _p, _p, _i64, _d, _d, _d, _i32, _p, _i64, _p, _p]
    L.cdj_assemble_dev_noalloc.argtypes = L.cdj_assemble_dev.argtypes
    L.cdj_lu_factor_dev.argtypes = [_p, _p, _i64, _i64, _p, _p]
    L.cdj_lu_factor_batched_dev.argtypes = [_p, _p, _p, _p, _i64, _p, _p, _p]
    L.cdj_lu_solve_dev.argtypes = [_p, _p, _i64, _i64, _p, _p, _i64, _i64, _p]
    L.cdj_remesh_isotropic.argtypes = [_p, _i64, _p, _i64, _d, _i32, _p, _p, _p]
    L.cdj_remesh_take.argtypes = [_p, _p, _p]
    for name in EXPORTS:
        f = getattr(L, name)
        if name == "cdj_last_kernel_ms":
            f.restype = ctypes.c_double
        elif name in ("cdj_mlp_nparams", "cdj_last_near_entries"):
            f.restype = ctypes.c_int64
        elif name != "cdj_last_error":
            f.restype = ctypes.c_int
    _LIB = L
    return L


def last_error():
    """The context's last message (an error, or why a mixed-precision solve fell back to FP64)."""
    msg = load().cdj_last_error(ctx())
    return msg.decode() if msg else ""


def ctx():
    """Process-wide context on device 0 (created on first use)."""
    global _CTX
    with _LOCK:
        if _CTX is None:
            L = load()
            h = _p()
            rc = L.cdj_create(0, ctypes.byref(h))
            if rc != CDJ_OK:
                raise CdjError(rc, "cdj_create failed: " + L.cdj_last_error(None).decode())
            _CTX = h
            # release the context (its stream, events, scratch, captured graphs) while the HIP runtime and any
            # profiler attached to it are still up, rather than leaving it to process teardown
            atexit.register(_destroy)
        return _CTX


def _destroy():
    global _CTX
    with _LOCK:
        if _CTX is not None and _LIB is not None:
            _LIB.cdj_destroy(_CTX)
        _CTX = None


def check(rc):
    if rc != CDJ_OK:
        raise CdjError(rc, load().cdj_last_error(ctx()).decode())


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def ptr_array(arrs):
    return (_p * len(arrs))(*[a.ctypes.data for a in arrs])
