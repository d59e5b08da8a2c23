"""ctypes binding of libpiv_amd.so (include/piv.h).

The library is built in-tree by ``make -C piv_amd`` (``__graft_entry__.build()``).  There is
no fallback: if the HIP library is missing, importing the bindings raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpiv_amd.so")

MAX_PASSES = 16

OK, ERR_CONFIG, ERR_SHAPE_MISMATCH, ERR_IMAGE_TOO_SMALL, ERR_DEVICE, ERR_OOM, ERR_ADMISSION, \
    ERR_CANCELLED, ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED = range(10)

PIXEL_U8, PIXEL_U16, PIXEL_F32 = 0, 1, 2


class piv_config(C.Structure):
    _fields_ = [("windows", C.c_uint32 * MAX_PASSES), ("num_windows", C.c_uint32),
                ("overlap", C.c_float), ("final_iterations", C.c_uint32),
                ("min_mean", C.c_float), ("min_std", C.c_float),
                ("ppr_threshold", C.c_float), ("median_threshold", C.c_float),
                ("median_eps", C.c_float), ("replace", C.c_int32),
                ("smooth_sigma", C.c_float), ("normalize_radius", C.c_uint32),
                ("normalize_eps", C.c_float), ("deformation", C.c_uint32)]

DEFORM = {"none": 0, "bilinear": 1, "bicubic": 2}


class piv_grid(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("window", "step", "nx", "ny", "x0", "y0")]


class piv_summary(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in
                ("no_signal", "low_ppr", "median_outliers", "out_of_range", "replaced")]


class piv_field_out(C.Structure):
    _fields_ = [("dx", C.c_void_p), ("dy", C.c_void_p), ("peak_ratio", C.c_void_p),
                ("flags", C.c_void_p), ("capacity", C.c_size_t), ("grid", piv_grid),
                ("summary", piv_summary)]


class piv_pass_trace(C.Structure):
    _fields_ = [("capacity", C.c_size_t), ("grid", piv_grid), ("origins", C.c_void_p),
                ("raw_dx", C.c_void_p), ("raw_dy", C.c_void_p), ("raw_ppr", C.c_void_p),
                ("raw_flags", C.c_void_p), ("peak_index", C.c_void_p), ("dx", C.c_void_p),
                ("dy", C.c_void_p), ("ppr", C.c_void_p), ("flags", C.c_void_p),
                ("summary", piv_summary)]


class piv_kernel_stat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("window", C.c_uint32), ("launches", C.c_uint64),
                ("units", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_double)]


class piv_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_batch_pairs", C.c_uint32),
                ("max_inflight", C.c_uint32)]


class piv_result(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("stream_id", C.c_uint32), ("status", C.c_int),
                ("grid", piv_grid), ("dx", C.POINTER(C.c_float)), ("dy", C.POINTER(C.c_float)),
                ("peak_ratio", C.POINTER(C.c_float)), ("flags", C.POINTER(C.c_uint8)),
                ("summary", piv_summary), ("submit_ms", C.c_double), ("done_ms", C.c_double),
                ("batch_pairs", C.c_uint32)]


class piv_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("batches", C.c_uint64), ("batched_pairs", C.c_uint64),
                ("hbm_allocated", C.c_size_t), ("hbm_peak", C.c_size_t), ("queued", C.c_size_t)]


class piv_tune_entry(C.Structure):
    _fields_ = [("window", C.c_uint32), ("fmt", C.c_int32), ("chosen", C.c_uint32),
                ("num_variants", C.c_uint32), ("occupancy", C.c_uint32 * 4), ("ms", C.c_float * 4)]


class piv_flow_stats(C.Structure):
    _fields_ = [("vorticity_mean", C.c_float), ("vorticity_abs_mean", C.c_float),
                ("ke_mean", C.c_float), ("ke_sum", C.c_float), ("coverage", C.c_float),
                ("noise_px", C.c_float), ("n_usable", C.c_uint32), ("n_vorticity", C.c_uint32)]


# every entry point declared in include/piv.h, with its signature
_SIGS = {
    "piv_last_error": ([], C.c_char_p),
    "piv_abi_version": ([], C.c_int),
    "piv_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "piv_config_default": ([C.POINTER(piv_config)], None),
    "piv_grid_new": ([C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(piv_grid)], C.c_int),
    "piv_analyzer_create": ([C.POINTER(piv_config), C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "piv_analyzer_destroy": ([C.c_void_p], None),
    "piv_analyzer_num_passes": ([C.c_void_p], C.c_uint32),
    "piv_output_grid": ([C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(piv_grid)], C.c_int),
    "piv_frame_prepare": ([C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_size_t,
                           C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "piv_frame_prepare_device_batch": ([C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_size_t, C.c_size_t, C.c_void_p,
                                        C.c_void_p], C.c_int),
    "piv_frame_prepare_device": ([C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32,
                                  C.c_size_t, C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "piv_frame_release": ([C.c_void_p], None),
    "piv_frame_download": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "piv_frame_tables_built": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "piv_analyze_pair": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(piv_field_out)], C.c_int),
    "piv_analyze_traced": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(piv_field_out),
                            C.POINTER(piv_pass_trace), C.c_uint32], C.c_int),
    "piv_analyze_batch_device": ([C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                  C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "piv_analyze_batch_noalloc": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_size_t, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                                  C.c_int),
    "piv_place_windows": ([C.c_void_p, C.POINTER(piv_grid), C.c_uint32, C.c_uint32,
                           C.POINTER(piv_grid), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "piv_correlate_windows": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                               C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p], C.c_int),
    "piv_validate_field": ([C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p, C.c_float, C.POINTER(piv_summary)], C.c_int),
    "piv_background_create": ([C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)], C.c_int),
    "piv_background_add": ([C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_int], C.c_int),
    "piv_background_finish": ([C.c_void_p, C.c_void_p], C.c_int),
    "piv_background_device": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "piv_background_destroy": ([C.c_void_p], None),
    "piv_device_alloc": ([C.c_int, C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "piv_device_free": ([C.c_int, C.c_void_p], C.c_int),
    "piv_copy_to_device": ([C.c_int, C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "piv_copy_to_host": ([C.c_int, C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "piv_analyzer_synchronize": ([C.c_void_p], C.c_int),
    "piv_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "piv_profile_read": ([C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "piv_server_create": ([C.c_int, C.POINTER(piv_server_config), C.POINTER(C.c_void_p)], C.c_int),
    "piv_server_destroy": ([C.c_void_p], None),
    "piv_server_add_stream": ([C.c_void_p, C.POINTER(piv_config), C.c_uint32, C.c_uint32, C.c_int,
                               C.c_void_p, C.POINTER(C.c_uint32)], C.c_int),
    "piv_submit": ([C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                    C.POINTER(C.c_uint64)], C.c_int),
    "piv_submit_device": ([C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                           C.POINTER(C.c_uint64)], C.c_int),
    "piv_submit_device_many": ([C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p),
                                C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint32)], C.c_int),
    "piv_poll": ([C.c_void_p, C.POINTER(piv_result), C.c_uint32, C.POINTER(C.c_uint32), C.c_int],
                 C.c_int),
    "piv_result_release": ([C.c_void_p, C.c_uint64], C.c_int),
    "piv_server_pause": ([C.c_void_p, C.c_int], C.c_int),
    "piv_server_stats_read": ([C.c_void_p, C.POINTER(piv_server_stats)], C.c_int),
    "piv_analyzer_set_precision": ([C.c_void_p, C.c_int], C.c_int),
    "piv_kernel_variant_count": ([C.c_uint32], C.c_uint32),
    "piv_kernel_variant_set": ([C.c_void_p, C.c_uint32, C.c_int, C.c_uint32], C.c_int),
    "piv_batch_lanes_set": ([C.c_void_p, C.c_uint32], C.c_int),
    "piv_batch_lanes_get": ([C.c_void_p, C.POINTER(C.c_uint32)], C.c_int),
    "piv_autotune": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                      C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "piv_write_pair_npz": ([C.c_char_p, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double,
                            C.c_uint32, C.c_uint32, C.POINTER(piv_grid), C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p], C.c_int),
    "piv_normalize_pixels": ([C.c_int, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p], C.c_int),
    "piv_flow_stats_device": ([C.c_void_p, C.POINTER(piv_grid), C.c_uint32, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_double, C.POINTER(piv_flow_stats)],
                              C.c_int),
    "piv_extract_plane_device": ([C.c_int, C.c_void_p, C.c_int, C.c_size_t, C.c_uint32,
                                  C.c_uint32, C.c_void_p], C.c_int),
    "piv_write_stats_npz": ([C.c_char_p, C.c_uint32, C.c_void_p, C.c_void_p], C.c_int),
    "piv_smooth_temporal_device": ([C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint32,
                                    C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_void_p], C.c_int),
    "piv_smooth_spatial_device": ([C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_int32,
                                   C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "piv_mask_device": ([C.c_int, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "piv_cross_validation_errors_device": ([C.c_int, C.c_size_t, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_uint64)], C.c_int),
    "piv_median_error_device": ([C.c_int, C.c_void_p, C.c_size_t, C.c_void_p,
                                 C.POINTER(C.c_float), C.POINTER(C.c_uint64)], C.c_int),
    "piv_pair_to_pair_change_device": ([C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                        C.c_void_p, C.POINTER(C.c_float)], C.c_int),
    "piv_spatial_roughness_device": ([C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_float)], C.c_int),
}

_lib = None


def load():
    """Load libpiv_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not found: build it with `make -C piv_amd` "
                          "(or __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if hasattr(a, "device_ptr"):  # piv_amd.DeviceArray
        return C.c_void_p(a.device_ptr)
    if hasattr(a, "data_ptr"):  # torch tensor on the device
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(int(a))
