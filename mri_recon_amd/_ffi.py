"""ctypes binding of libmrn.so (declarations in include/mrn.h).

The library is the only compute path: importing this module on a host where
libmrn.so is missing or cannot be loaded raises, it never falls back to a
CPU implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmrn.so")

MRN_OK, MRN_EINVAL, MRN_ENOMEM, MRN_EHIP, MRN_EUNSUPPORTED, MRN_ESTATE = range(6)
KERNELS = {"triangle": 0, "kaiser": 1, "sinc": 2, "poly": 3}
KMODES = {"table": 0, "exact_kb": 1}
FORMS = {"batched": 0, "plain": 1}

# every symbol include/mrn.h declares
EXPORTS = [
    "mrn_params_default", "mrn_parse_args", "mrn_op_create", "mrn_op_destroy",
    "mrn_get_geometry", "mrn_get_table", "mrn_get_window", "mrn_set_traj", "mrn_set_smaps",
    "mrn_image_to_kspace", "mrn_kspace_to_image", "mrn_normal", "mrn_coil_backward",
    "mrn_coil_forward", "mrn_set_kdata", "mrn_cg", "mrn_cg_init", "mrn_cg_iterate",
    "mrn_power_method", "mrn_fista", "mrn_traj_grad", "mrn_autotune", "mrn_get_tuning", "mrn_stage_times",
    "mrn_coil_forward_sos", "mrn_set_k_rad", "mrn_wavelet", "mrn_l1_wavelet", "mrn_set_fista_l1",
    "mrn_get_fista_thresholds", "mrn_set_fista_step", "mrn_get_fista_steps", "mrn_dcf_params_default", "mrn_dcf_parse_args", "mrn_dcf_grid_size",
    "mrn_dcf_calc", "mrn_dcf_last_atomics", "mrn_llr_params_default", "mrn_llr_parse_args", "mrn_llr", "mrn_llr_noalloc", "mrn_set_fista_llr",
    "mrn_get_fista_llr_smax", "mrn_gate_params_default", "mrn_gate_parse_args", "mrn_gate_scale_times",
    "mrn_pregate",
    "mrn_get_frame_stats", "mrn_get_support", "mrn_get_grid", "mrn_grid_zero", "mrn_chop_grid_forward", "mrn_chop_grid_backward",
    "mrn_sync", "mrn_status_string", "mrn_last_error",
]


class MrnTuning(ctypes.Structure):
    """mrn_tuning (include/mrn.h)."""
    _fields_ = [
        ("bin_edge", ctypes.c_int32),
        ("max_visits", ctypes.c_int32),
        ("xcd_run", ctypes.c_int32),
        ("xcd_run_gather", ctypes.c_int32),
        ("gather_coils", ctypes.c_int32),
        ("xcd_run_fft", ctypes.c_int32),
        ("fft_stockham", ctypes.c_int32),
        ("scatter_valu", ctypes.c_int32),
        ("tile", ctypes.c_int32 * 3),
        ("from_cache", ctypes.c_int32),
        ("normal_ms", ctypes.c_float),
        ("gather_tiles", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class MrnFrameStats(ctypes.Structure):
    """mrn_frame_stats (include/mrn.h)."""
    _fields_ = [
        ("samples", ctypes.c_int64),
        ("valid", ctypes.c_int64),
        ("tile_visits", ctypes.c_int64),
        ("scatter_atomics", ctypes.c_int64),
        ("scatter_items", ctypes.c_int32),
        ("split_items", ctypes.c_int32),
        ("split_tiles", ctypes.c_int32),
        ("wide", ctypes.c_int32),
        ("gather_items", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]


class GateParams(ctypes.Structure):
    """mrn_gate_params (include/mrn.h)."""
    _fields_ = [
        ("gate_type", ctypes.c_int32),
        ("vs_type", ctypes.c_int32),
        ("frames", ctypes.c_int32),
        ("tornado_shape", ctypes.c_int32),
        ("wdth_low", ctypes.c_int32),
        ("wdth_high", ctypes.c_int32),
        ("kmax", ctypes.c_float),
        ("min_quantile", ctypes.c_float),
        ("max_quantile", ctypes.c_float),
        ("ecg_outlier_reject", ctypes.c_int32),
        ("reset_dens", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class LlrParams(ctypes.Structure):
    """mrn_llr_params (include/mrn.h)."""
    _fields_ = [
        ("block", ctypes.c_int32 * 3),
        ("block_iter", ctypes.c_int32),
        ("alpha_time", ctypes.c_float),
        ("alpha_encode", ctypes.c_float),
        ("amp", ctypes.c_float),
        ("t2", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("reserved", ctypes.c_int32 * 4),
    ]


class DcfParams(ctypes.Structure):
    """mrn_dcf_params (include/mrn.h)."""
    _fields_ = [
        ("N", ctypes.c_int32 * 3),
        ("iters", ctypes.c_int32),
        ("dwin", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("overgrid", ctypes.c_float),
        ("acc", ctypes.c_float),
        ("kernel", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class MrnParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32 * 3),
        ("overgrid", ctypes.c_float),
        ("grid", ctypes.c_float * 3),
        ("dwin", ctypes.c_float * 3),
        ("kernel", ctypes.c_int32),
        ("grid_in", ctypes.c_int32 * 3),
        ("fft_in", ctypes.c_int32 * 3),
        ("grid_scale", ctypes.c_float * 3),
        ("k_rad", ctypes.c_float),
        ("ncoils", ctypes.c_int32),
        ("nframes", ctypes.c_int32),
        ("kernel_mode", ctypes.c_int32),
        ("form", ctypes.c_int32),
        ("coil_chunk", ctypes.c_int32),
        ("tile", ctypes.c_int32 * 3),
        ("max_visits", ctypes.c_int32),
        ("bin_edge", ctypes.c_int32),
        ("xcd_run", ctypes.c_int32),
        ("xcd_run_gather", ctypes.c_int32),
        ("gather_tiles", ctypes.c_int32),
        ("scatter_valu", ctypes.c_int32),
        ("fft_stockham", ctypes.c_int32),
        ("xcd_run_fft", ctypes.c_int32),
        ("gather_coils", ctypes.c_int32),
        ("fuse_frames", ctypes.c_int32),
    ]


class MrnError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("mrn status %d (%s): %s" % (status, status_string(status), msg))
        self.status = status


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmrn.so not built (run `python -m mri_recon_amd.build`): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        P = ctypes.POINTER(MrnParams)
        sig = {
            "mrn_params_default": (None, [P]),
            "mrn_parse_args": (i32, [P, i32, ctypes.POINTER(ctypes.c_char_p)]),
            "mrn_op_create": (i32, [P, ctypes.POINTER(vp)]),
            "mrn_op_destroy": (i32, [vp]),
            "mrn_get_geometry": (i32, [vp, vp, vp, vp]),
            "mrn_get_table": (i32, [vp, i32, vp, i32, ctypes.POINTER(ctypes.c_int)]),
            "mrn_get_window": (i32, [vp, i32, vp, i32, ctypes.POINTER(ctypes.c_int)]),
            "mrn_set_traj": (i32, [vp, i32, i64, vp, vp, vp, vp, vp]),
            "mrn_set_smaps": (i32, [vp, vp, vp]),
            "mrn_image_to_kspace": (i32, [vp, i32, vp, vp, vp, vp]),
            "mrn_kspace_to_image": (i32, [vp, i32, vp, vp, i32, vp]),
            "mrn_normal": (i32, [vp, i32, vp, vp, vp]),
            "mrn_coil_backward": (i32, [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]),
            "mrn_coil_forward": (i32, [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp]),
            "mrn_coil_forward_sos": (i32, [vp, vp, vp, i64, vp, vp, vp, vp, vp]),
            "mrn_set_k_rad": (i32, [vp, ctypes.c_float, vp]),
            "mrn_wavelet": (i32, [vp, vp, i32, vp, vp, i32, vp]),
            "mrn_l1_wavelet": (i32, [vp, vp, i32, ctypes.c_double, vp, i32, vp, vp, vp]),
            "mrn_set_fista_l1": (i32, [vp, ctypes.c_double, vp, i32, ctypes.c_uint64]),
            "mrn_get_fista_thresholds": (i32, [vp, vp, i32, vp]),
            "mrn_set_fista_step": (i32, [vp, i32, i32]),
            "mrn_get_fista_steps": (i32, [vp, vp, i32, vp]),
            "mrn_set_kdata": (i32, [vp, i32, vp, vp]),
            "mrn_cg": (i32, [vp, vp, i32, i32, vp, vp]),
            "mrn_cg_init": (i32, [vp, vp, vp]),
            "mrn_cg_iterate": (i32, [vp, vp, i32, i32, vp, vp]),
            "mrn_stage_times": (i32, [vp, i32, vp, i32, vp, vp]),
            "mrn_power_method": (i32, [vp, i32, ctypes.POINTER(dbl), vp]),
            "mrn_fista": (i32, [vp, vp, i32, dbl, i32, vp, vp]),
            "mrn_traj_grad": (i32, [vp, i32, vp, vp, vp, vp]),
            "mrn_autotune": (i32, [vp, i32, ctypes.c_char_p, vp]),
            "mrn_get_tuning": (i32, [vp, ctypes.POINTER(MrnTuning)]),
            "mrn_dcf_params_default": (None, [ctypes.POINTER(DcfParams)]),
            "mrn_dcf_parse_args": (i32, [ctypes.POINTER(DcfParams), i32, ctypes.POINTER(ctypes.c_char_p)]),
            "mrn_dcf_grid_size": (i32, [ctypes.POINTER(DcfParams), vp]),
            "mrn_dcf_calc": (i32, [ctypes.POINTER(DcfParams), i64, vp, vp, vp, vp, vp, vp]),
            "mrn_dcf_last_atomics": (i32, [ctypes.POINTER(i64)]),
            "mrn_llr_params_default": (None, [ctypes.POINTER(LlrParams)]),
            "mrn_llr_parse_args": (i32, [ctypes.POINTER(LlrParams), i32, ctypes.POINTER(ctypes.c_char_p)]),
            "mrn_llr": (i32, [ctypes.POINTER(LlrParams), vp, i32, i32, vp, i32, i32, i64, vp, vp]),
            "mrn_llr_noalloc": (i32, [ctypes.POINTER(LlrParams), vp, i32, i32, vp, i32, i32, i64, vp, vp]),
            "mrn_set_fista_llr": (i32, [vp, ctypes.POINTER(LlrParams)]),
            "mrn_get_fista_llr_smax": (i32, [vp, vp, vp]),
            "mrn_gate_params_default": (None, [ctypes.POINTER(GateParams)]),
            "mrn_gate_parse_args": (i32, [ctypes.POINTER(GateParams), i32, ctypes.POINTER(ctypes.c_char_p)]),
            "mrn_gate_scale_times": (i32, [ctypes.POINTER(GateParams), i64, vp, vp]),
            "mrn_pregate": (i32, [ctypes.POINTER(GateParams), i64, i64, vp, vp, vp, vp, vp, vp, i32, vp, vp,
                                  vp, vp, vp, vp, vp, vp]),
            "mrn_get_frame_stats": (i32, [vp, i32, ctypes.POINTER(MrnFrameStats)]),
            "mrn_get_support": (i32, [vp, i32, vp, vp]),
            "mrn_get_grid": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(i64)]),
            "mrn_grid_zero": (i32, [vp, vp]),
            "mrn_chop_grid_forward": (i32, [vp, i32, vp, vp]),
            "mrn_chop_grid_backward": (i32, [vp, i32, vp, vp, vp]),
            "mrn_sync": (i32, [vp, vp]),
            "mrn_status_string": (ctypes.c_char_p, [i32]),
            "mrn_last_error": (ctypes.c_char_p, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(s):
    try:
        return lib().mrn_status_string(int(s)).decode()
    except Exception:
        return "?"


def check(status, handle=None):
    if status != MRN_OK:
        msg = (lib().mrn_last_error(handle) or b"").decode()
        raise MrnError(status, msg)
