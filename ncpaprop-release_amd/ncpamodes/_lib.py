"""ctypes binding of libncpamodes.so (the C ABI declared in include/ncpamodes.h).

The library is loaded from the package directory.  There is no fallback: if the HIP
library is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libncpamodes.so")

NCPA_OK = 0
NCPA_ERR_ARG = -1
NCPA_ERR_HIP = -2
NCPA_ERR_ALLOC = -3
NCPA_ERR_MAX_MODES = -4
NCPA_ERR_STATE = -5
NCPA_MAX_MODES = 4000
NCPA_PULSE_SCRATCH_BYTES = 32 << 20

NCPA_PREC_FP64 = 0
NCPA_PREC_FP32 = 1
NCPA_PREC_BF16X3 = 2
NCPA_PREC_BF16X6 = 3
NCPA_PREC_FP16X3 = 4
NCPA_PREC_FP16X4 = 5
NCPA_PREC_BF16X6F = 6
NCPA_PREC_FP16X4F = 7

NCPA_F_WKB = 1
NCPA_F_WVNUM_FILTER = 2

NCPA_SRC = {"impulse": 0, "pulse1": 1, "pulse2": 2, "spectrum": 3, "waveform": 4}

# Every symbol include/ncpamodes.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "ncpa_ctx_create", "ncpa_ctx_destroy", "ncpa_ctx_error", "ncpa_ctx_synchronize",
    "ncpa_ctx_stream", "ncpa_ctx_set_timing", "ncpa_ctx_kernel_time", "ncpa_ctx_reset_timing", "ncpa_ctx_memory",
    "ncpa_ctx_set_modes_form",
    "ncpa_event_record", "ncpa_event_wait", "ncpa_event_synchronize", "ncpa_event_destroy",
    "ncpa_dev_alloc", "ncpa_dev_free", "ncpa_dev_upload", "ncpa_dev_download", "ncpa_dev_copy",
    "ncpa_profile_upload", "ncpa_profile_set_attenuation", "ncpa_profile_clear",
    "ncpa_profile_read", "ncpa_profile_check", "ncpa_profile_nodes", "ncpa_profile_admittance",
    "ncpa_grid_make", "ncpa_profile_sample",
    "ncpa_ess_solve", "ncpa_ess_count", "ncpa_modes_free", "ncpa_modes_counts", "ncpa_modes_interval",
    "ncpa_modes_download", "ncpa_modes_iterations", "ncpa_modes_sweep_rows", "ncpa_modes_device", "ncpa_modes_ceff", "ncpa_modes_rows", "ncpa_tl1d", "ncpa_field2d",
    "ncpa_field2d_stage", "ncpa_field2d_contract", "ncpa_field_stage_free",
    "ncpa_field_variants", "ncpa_field_variant_info", "ncpa_field_set_variant", "ncpa_field_get_variant", "ncpa_field_tune",
    "ncpa_field_compare", "ncpa_ess_eigen",
    "ncpa_wmod_solve", "ncpa_wmod_eigen",
    "ncpa_dispersion_write", "ncpa_dispersion_read", "ncpa_dispersion_info", "ncpa_dispersion_arrays",
    "ncpa_dispersion_free",
    "ncpa_fft_batch", "ncpa_kpert", "ncpa_bb_transfer", "ncpa_bb_pulse", "ncpa_bb_pulse_noalloc",
    "ncpa_bb_source",
    "ncpa_bb_plan_create", "ncpa_bb_plan_set_problems", "ncpa_bb_plan_run", "ncpa_bb_plan_result",
    "ncpa_bb_plan_tuning", "ncpa_bb_plan_destroy",
]


class Grid(C.Structure):
    _fields_ = [("Nz", C.c_int), ("z_min", C.c_double), ("dz", C.c_double),
                ("maxheight", C.c_double)]


class Problem(C.Structure):
    _fields_ = [("profile", C.c_int), ("flags", C.c_int), ("freq", C.c_double),
                ("azimuth", C.c_double), ("alpha_freq", C.c_double), ("zs", C.c_double),
                ("zr", C.c_double), ("admittance", C.c_double), ("c_min", C.c_double),
                ("c_max", C.c_double)]


class BbJob(C.Structure):
    _fields_ = [("n_sets", C.c_int), ("n_freq", C.c_int), ("f_vec", C.POINTER(C.c_double)),
                ("spectrum", C.POINTER(C.c_double)), ("ranges_m", C.POINTER(C.c_double)),
                ("n_ranges", C.c_int), ("rho_ratio", C.POINTER(C.c_double)), ("max_celerity", C.c_double),
                ("nfft", C.c_int), ("zero_attn", C.c_int), ("max_modes_total", C.c_int), ("tune", C.c_int)]


class ProfileInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("has_z0", C.c_int), ("z0", C.c_double),
                ("z_bottom", C.c_double), ("z_top", C.c_double)]


_lib = None


class NcpaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ncpamodes error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libncpamodes.so (raises OSError if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libncpamodes.so not built (run __graft_entry__.build()): " + LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp, ip, vp = P(C.c_double), P(C.c_int), C.c_void_p
    sig = {
        "ncpa_ctx_create": (C.c_int, [C.c_int, vp, P(vp)]),
        "ncpa_ctx_destroy": (None, [vp]),
        "ncpa_ctx_error": (C.c_char_p, [vp]),
        "ncpa_ctx_synchronize": (C.c_int, [vp]),
        "ncpa_ctx_stream": (vp, [vp]),
        "ncpa_ctx_set_timing": (C.c_int, [vp, C.c_int]),
        "ncpa_ctx_set_modes_form": (C.c_int, [vp, C.c_int]),
        "ncpa_ctx_kernel_time": (C.c_int, [vp, C.c_char_p, dp, ip]),
        "ncpa_ctx_reset_timing": (C.c_int, [vp]),
        "ncpa_ctx_memory": (C.c_int, [vp, C.c_int, P(C.c_size_t), P(C.c_size_t), P(C.c_size_t)]),
        "ncpa_event_record": (C.c_int, [vp, P(vp)]),
        "ncpa_event_wait": (C.c_int, [vp, vp]),
        "ncpa_event_synchronize": (C.c_int, [vp]),
        "ncpa_event_destroy": (None, [vp]),
        "ncpa_dev_alloc": (C.c_int, [vp, C.c_size_t, P(vp)]),
        "ncpa_dev_free": (C.c_int, [vp, vp]),
        "ncpa_dev_upload": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "ncpa_dev_download": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "ncpa_dev_copy": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "ncpa_profile_upload": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, dp, dp, ip]),
        "ncpa_profile_set_attenuation": (C.c_int, [vp, C.c_int, dp]),
        "ncpa_profile_clear": (C.c_int, [vp]),
        "ncpa_profile_read": (C.c_int, [vp, C.c_char_p, C.c_char_p, ip, P(ProfileInfo)]),
        "ncpa_profile_check": (C.c_int, [C.c_char_p, C.c_char_p, P(ProfileInfo), C.c_char_p, C.c_int]),
        "ncpa_profile_nodes": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "ncpa_profile_admittance": (C.c_int, [vp, C.c_int, C.c_double, dp]),
        "ncpa_grid_make": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, P(Grid)]),
        "ncpa_profile_sample": (C.c_int, [vp, C.c_int, P(Grid), dp, dp, dp, dp]),
        "ncpa_ess_solve": (C.c_int, [vp, P(Grid), P(Problem), C.c_int, P(vp)]),
        "ncpa_ess_count": (C.c_int, [vp, P(Grid), P(Problem), C.c_int, ip]),
        "ncpa_modes_free": (None, [vp]),
        "ncpa_modes_counts": (C.c_int, [vp, ip, ip]),
        "ncpa_modes_interval": (C.c_int, [vp, dp, dp]),
        "ncpa_modes_download": (C.c_int, [vp, dp, dp, dp, dp, dp]),
        "ncpa_modes_iterations": (C.c_int, [vp, ip]),
        "ncpa_modes_sweep_rows": (C.c_int, [vp, ip]),
        "ncpa_modes_rows": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, vp]),
        "ncpa_modes_device": (C.c_int, [vp, P(vp), P(vp), P(vp), P(vp)]),
        "ncpa_modes_ceff": (C.c_int, [vp, C.c_int, dp]),
        "ncpa_tl1d": (C.c_int, [vp, vp, C.c_double, C.c_int, C.c_int, vp, vp, vp]),
        "ncpa_field2d": (C.c_int, [vp, vp, vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, vp]),
        "ncpa_field2d_stage": (C.c_int, [vp, vp, vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, P(vp)]),
        "ncpa_field2d_contract": (C.c_int, [vp, vp, vp]),
        "ncpa_field_stage_free": (None, [vp]),
        "ncpa_field_variants": (C.c_int, [C.c_int, ip]),
        "ncpa_field_variant_info": (C.c_int, [C.c_int, C.c_int, P(C.c_char_p), ip]),
        "ncpa_field_set_variant": (C.c_int, [vp, C.c_int, C.c_int]),
        "ncpa_field_get_variant": (C.c_int, [vp, C.c_int, ip]),
        "ncpa_field_tune": (C.c_int, [vp, vp, vp, C.c_int, ip, dp, C.c_int]),
        "ncpa_field_compare": (C.c_int, [vp, vp, vp, C.c_int, C.c_size_t, C.c_double, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_size_t)]),
        "ncpa_ess_eigen": (C.c_int, [vp, dp, C.c_int, C.c_double, C.c_double, dp, dp, C.c_int,
                                     ip, dp, dp, C.c_int]),
        "ncpa_wmod_solve": (C.c_int, [vp, P(Grid), P(Problem), C.c_int, P(vp)]),
        "ncpa_dispersion_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, ip, dp, dp, dp, dp, dp, dp,
                                            C.c_int]),
        "ncpa_dispersion_read": (C.c_int, [C.c_char_p, C.c_int, P(vp)]),
        "ncpa_dispersion_info": (C.c_int, [vp, ip, P(C.c_size_t)]),
        "ncpa_dispersion_arrays": (C.c_int, [vp, dp, ip, dp, dp, dp, dp, dp, dp]),
        "ncpa_dispersion_free": (None, [vp]),
        "ncpa_wmod_eigen": (C.c_int, [vp, dp, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int, ip, dp, dp, C.c_int]),
        "ncpa_fft_batch": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_int]),
        "ncpa_kpert": (C.c_int, [vp, vp, vp, C.c_int, vp, vp]),
        "ncpa_bb_transfer": (C.c_int, [vp, C.c_int, C.c_int, ip, ip, vp, vp, vp, C.c_int, dp, C.c_int,
                                       dp, vp]),
        "ncpa_bb_pulse": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, dp, dp, dp, C.c_double, C.c_int,
                                    vp]),
        "ncpa_bb_pulse_noalloc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, dp, dp, dp, C.c_double, C.c_int,
                                            vp]),
        "ncpa_bb_source": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, C.c_double, dp, C.c_int, dp,
                                     dp]),
        "ncpa_bb_plan_create": (C.c_int, [vp, P(Grid), P(Problem), P(BbJob), P(vp)]),
        "ncpa_bb_plan_set_problems": (C.c_int, [vp, P(Problem)]),
        "ncpa_bb_plan_run": (C.c_int, [vp]),
        "ncpa_bb_plan_result": (C.c_int, [vp, P(vp), P(vp), P(vp), P(vp), P(vp), ip]),
        "ncpa_bb_plan_tuning": (C.c_int, [vp, ip, ip, dp, dp, ip]),
        "ncpa_bb_plan_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
