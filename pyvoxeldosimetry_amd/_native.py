"""ctypes binding of ``libpvd.so`` (the C ABI declared in ``include/pvd.h``).

The product path has no CPU fallback: if the library is missing or no HIP
device is present, every compute call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpvd.so")

PVD_OK = 0
PVD_EINVAL = -1
PVD_ENOMEM = -2
PVD_EHIP = -3
PVD_ENOTSUP = -4

PVD_MODE_REFERENCE = 0
PVD_MODE_LINEAR = 1
PVD_ALGO_AUTO = 0
PVD_ALGO_FFT = 1
PVD_ALGO_DIRECT = 2
PVD_FLAG_PROFILE = 1
PVD_FLAG_GENERAL_SPECTRUM = 2
PVD_FLAG_STOCKHAM = 4
PVD_FLAG_SMOOTH_LENGTHS = 8
PVD_PATH_STOCKHAM = 0
PVD_PATH_SQ256 = 1
PVD_PATH_TWO_STEP = 2
PATHS = {PVD_PATH_STOCKHAM: "stockham", PVD_PATH_SQ256: "sq256", PVD_PATH_TWO_STEP: "two_step"}
PVD_MAX_TIMEPOINTS = 16
PVD_MAX_FFT_LEN = 2048
PVD_NOALLOC_MAX_TAPS = 4096
PVD_MC_MAX_LINES = 4

PVD_DPK_BETA = 0
PVD_DPK_PHOTON = 1
PVD_DPK_NAN_CENTRE = 1

MODES = {"reference": PVD_MODE_REFERENCE, "linear": PVD_MODE_LINEAR}
ALGOS = {"auto": PVD_ALGO_AUTO, "fft": PVD_ALGO_FFT, "direct": PVD_ALGO_DIRECT}


class PlanParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("k", ctypes.c_int * 3), ("mode", ctypes.c_int),
                ("algo", ctypes.c_int), ("fft_n", ctypes.c_int * 3), ("flags", ctypes.c_int),
                ("tile", ctypes.c_int * 3), ("max_radix", ctypes.c_int)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int), ("symmetric_spectrum", ctypes.c_int), ("fft_n", ctypes.c_int * 3),
                ("kernel_used", ctypes.c_int * 3), ("n_passes", ctypes.c_int), ("workspace_bytes", ctypes.c_size_t),
                ("tile", ctypes.c_int * 3), ("max_radix", ctypes.c_int), ("fft_path", ctypes.c_int * 3),
                ("block", ctypes.c_int * 5), ("grid", ctypes.c_int * 5)]


class DpkTerm(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("flags", ctypes.c_int), ("amp", ctypes.c_double),
                ("param", ctypes.c_double)]


# every symbol include/pvd.h declares, with (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
SIGNATURES = {
    "pvd_last_error": (ctypes.c_char_p, []),
    "pvd_version": (_I, []),
    "pvd_dpk_generate": (_I, [_D, _I, _I, _I, ctypes.POINTER(DpkTerm), _I, _D, _P, _P]),
    "pvd_plan_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(PlanParams)]),
    "pvd_plan_info_get": (_I, [_P, ctypes.POINTER(PlanInfo)]),
    "pvd_plan_describe": (_I, [ctypes.POINTER(PlanParams), ctypes.POINTER(PlanInfo)]),
    "pvd_plan_destroy": (_I, [_P]),
    "pvd_plan_set_kernel": (_I, [_P, _P, _P]),
    "pvd_conv_dose": (_I, [_P, _P, _P, _P, _D, _P]),
    "pvd_conv_dose_weighted": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_D), _I, _P, _P, _D, _P]),
    "pvd_conv_dose_timeseries": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_D), _I, _P, _P, _D, _P]),
    "pvd_weighted_sum": (_I, [ctypes.POINTER(_P), ctypes.POINTER(_D), _I, ctypes.c_size_t, _P, _P]),
    "pvd_batch_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(_P), _I]),
    "pvd_batch_set_volume": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_D), _I, _P, _P, _D]),
    "pvd_batch_run": (_I, [_P, _P]),
    "pvd_batch_run_graph": (_I, [_P, _P]),
    "pvd_batch_destroy": (_I, [_P]),
    "pvd_plan_timings": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I]),
    "pvd_conv_dose_direct_noalloc": (_I, [ctypes.POINTER(_I), ctypes.POINTER(_I), _I, _P, _P, _P, _P, _D, _P]),
    "pvd_pass_name": (ctypes.c_char_p, [_P, _I]),
    # include/pvd_learned.h
    "pvd_mc_dose": (_I, [ctypes.c_void_p, _P, _P, _P, _P, _P]),
    "pvd_cnn_conv_fwd": (_I, [_P, _I, _P, _P, _I, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P]),
    "pvd_cnn_region_block_count": (ctypes.c_size_t, [_I, _I, _I]),
    "pvd_cnn_region_blocks": (_I, [_P, _I, _I, _I, _P, ctypes.c_float, _P, _P, _P, _P]),
    "pvd_cnn_flip_weights": (_I, [_P, _P, _P]),
    "pvd_cnn_relu_grad": (_I, [_P, _P, _P, _P, ctypes.c_longlong, _P]),
    "pvd_cnn_wgrad_work_floats": (ctypes.c_size_t, [_I, _I, _I]),
    "pvd_cnn_wgrad": (_I, [_P, _I, _P, _P, _P, _I, _I, _I, _P]),
    "pvd_cnn_loss": (_I, [_P, _P, _P, _P, ctypes.c_longlong, ctypes.c_float, _P, _P, _P]),
    "pvd_cnn_adam": (_I, [_P, _P, _P, _P, _P, ctypes.c_longlong, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, _I, _P]),
    "pvd_reduce_rows": (_I, [_P, _I, _I, _P, _I, _P]),
    "pvd_cnn_pack_inputs": (_I, [_P, _P, _P, _P, _P, _P, ctypes.c_longlong, _P]),
    "pvd_cnn_apply_correction": (_I, [_P, _P, ctypes.c_float, _P, ctypes.c_float, ctypes.c_float, _P, _P,
                                      ctypes.c_longlong, _P]),
    "pvd_cnn_apply_fallback": (_I, [_P, _P, _D, _P, ctypes.c_longlong, _P]),
    "pvd_cnn_input_maxima": (_I, [_P, _P, ctypes.c_longlong, _P, _P]),
    # include/pvd_tissue.h
    "pvd_tissue_map": (_I, [_P, _P, _I, _I, _I, _P, _P, _P, _P]),
    # include/pvd_resample.h
    "pvd_resample": (_I, [_P, _I, _I, _I, ctypes.POINTER(_D), _P, _I, _I, _I, ctypes.POINTER(_D), _I, _D, _P]),
    "pvd_bspline_prefilter": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    # include/pvd_tac.h
    "pvd_tac_n_params": (_I, [_I]),
    "pvd_tac_fit": (_I, [_P, ctypes.POINTER(_P), ctypes.c_longlong, _P, _P, _P, _P]),
}

PVD_TAC_MAX_T = 16
PVD_TAC_MAX_P = 5


class TacParams(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("n_times", ctypes.c_int), ("t", ctypes.c_double * PVD_TAC_MAX_T),
                ("w", ctypes.c_double * PVD_TAC_MAX_T), ("lambda_phys", ctypes.c_double),
                ("p0", ctypes.c_double * PVD_TAC_MAX_P), ("p0_from_data", ctypes.c_int),
                ("force_zero_at_t0", ctypes.c_int), ("max_iter", ctypes.c_int), ("tol", ctypes.c_double),
                ("fail_params", ctypes.c_double * PVD_TAC_MAX_P), ("auc_mode", ctypes.c_int),
                ("t_int", ctypes.c_double)]

PVD_TISSUE_MAX_CLASSES = 8
PVD_TISSUE_MAX_MATERIALS = 32
PVD_TISSUE_MAX_RADIUS = 16


class TissueParams(ctypes.Structure):
    _fields_ = [("n_classes", ctypes.c_int), ("class_lo", ctypes.c_float * PVD_TISSUE_MAX_CLASSES),
                ("class_hi", ctypes.c_float * PVD_TISSUE_MAX_CLASSES), ("n_materials", ctypes.c_int),
                ("mat_lo", ctypes.c_float * PVD_TISSUE_MAX_MATERIALS),
                ("mat_density", ctypes.c_float * PVD_TISSUE_MAX_MATERIALS), ("handle_artifacts", ctypes.c_int),
                ("metal_hu", ctypes.c_float), ("sigma", ctypes.c_double), ("truncate", ctypes.c_double)]


class McParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("voxel_mm", ctypes.c_double), ("histories", ctypes.c_int),
                ("seed", ctypes.c_ulonglong), ("e_of_u", ctypes.c_void_p), ("n_u", ctypes.c_int),
                ("e_of_r", ctypes.c_void_p), ("n_r", ctypes.c_int), ("r_max_mm", ctypes.c_double),
                ("katz_a", ctypes.c_double), ("katz_b", ctypes.c_double), ("katz_c", ctypes.c_double),
                ("step_mm", ctypes.c_double), ("scatter", ctypes.c_double), ("rho_min", ctypes.c_double),
                ("active", ctypes.c_void_p), ("n_active", ctypes.c_longlong), ("n_lines", ctypes.c_int),
                ("line_energy", ctypes.c_double * 4), ("line_yield", ctypes.c_double * 4), ("brems_z", ctypes.c_double),
                ("photon_every", ctypes.c_int), ("electrons", ctypes.c_int), ("mu_en", ctypes.c_void_p),
                ("n_mu", ctypes.c_int), ("mu_log_e_lo", ctypes.c_double), ("mu_log_e_hi", ctypes.c_double)]


PVD_CNN_EPI_RELU = 0
PVD_CNN_EPI_RESID_RELU = 1
PVD_CNN_EPI_LINEAR_F32 = 2
PVD_CNN_EPI_ACCUM_F32 = 3
PVD_CNN_MAXIMA_FLOATS = 2 + 2 * 1024

_lib = None
_lock = threading.Lock()


class NativeLibraryMissing(RuntimeError):
    pass


def load(path: str | None = None):
    """Load and type the library (does not touch the GPU)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise NativeLibraryMissing(
                f"{p} is missing: build it with `python -m pyvoxeldosimetry_amd.build` "
                "(there is no CPU fallback)")
        # torch first, so the HIP runtime it ships is the one libpvd binds to
        import torch  # noqa: F401
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(rc: int, what: str = "pvd call"):
    """Map a PVD status code to the reference's Python exception types."""
    if rc >= 0:
        return rc
    msg = (load().pvd_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}"
    if rc == PVD_EINVAL:
        raise ValueError(text)
    if rc == PVD_ENOMEM:
        raise MemoryError(text)
    if rc == PVD_ENOTSUP:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def require_device():
    """The product path runs only on a HIP device."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyvoxeldosimetry_amd needs a ROCm GPU (HIP device); none is available")


def current_stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
