"""Loads ``libndfgpu.so`` (the HIP library) and declares its C ABI.

There is no fallback: if the library is missing the import of the native
layer fails loudly.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libndfgpu.so")

OK = 0
ERR_EMPTY_AXIS = 1
ERR_INVALID = 2
ERR_LENGTH_MISMATCH = 3
ERR_MEMORY = 4
ERR_OVERFLOW = 5
ERR_DEVICE = 6
ERR_KIND = 7
ERR_MODE = 8

MODES = ("reflect", "constant", "nearest", "mirror", "wrap")
KINDS = {
    "median": 0,
    "median_network": 1,
    "median_histogram": 2,
    "rank": 3,
    "rank_histogram": 4,
    "min": 5,
    "max": 6,
}
# Robust-filter kinds of the reference (src/pyapi/robust.rs:110-140); only
# the percentile is implemented here.
ROBUST_KINDS = {"percentile": 0}
ROBUST_OTHER_KINDS = ("mad", "mode", "biweight_location", "biweight_scale", "sigma_mean", "sigma_median",
                      "sigma_mode")
ENGINES = {"auto": 0, "generic": 1, "slide": 2, "network": 3, "extrema": 4, "stack": 5}
REPAIR_PATHS = {"auto": 0, "sparse": 1, "dense": 2}

_size_p = ctypes.POINTER(ctypes.c_size_t)
_ssize_p = ctypes.POINTER(ctypes.c_ssize_t)
_i32_p = ctypes.POINTER(ctypes.c_int32)
_err = [ctypes.c_char_p, ctypes.c_size_t]


class NativeLibraryMissing(ImportError):
    """libndfgpu.so has not been built."""


def _declare(lib: ctypes.CDLL) -> None:
    plan_p = ctypes.c_void_p
    lib.ndf_window_filter.argtypes = [
        ctypes.c_int32,
        ctypes.c_void_p,
        ctypes.c_void_p,
        ctypes.c_size_t,
        _size_p,
        _size_p,
        ctypes.c_void_p,
        _ssize_p,
        _i32_p,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_size_t,
        ctypes.c_size_t,
        *_err,
    ]
    lib.ndf_plan_create.argtypes = [
        ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_int32,
        ctypes.c_size_t,
        _size_p,
        _size_p,
        ctypes.c_void_p,
        _ssize_p,
        _i32_p,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_size_t,
        ctypes.c_size_t,
        *_err,
    ]
    lib.ndf_robust_plan_create.argtypes = [
        ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_int32,
        ctypes.c_size_t,
        _size_p,
        _size_p,
        ctypes.c_void_p,
        _ssize_p,
        _i32_p,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_size_t,
        *_err,
    ]
    lib.ndf_robust_filter.argtypes = [
        ctypes.c_int32,
        ctypes.c_void_p,
        ctypes.c_void_p,
        ctypes.c_size_t,
        _size_p,
        _size_p,
        ctypes.c_void_p,
        _ssize_p,
        _i32_p,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_size_t,
        *_err,
    ]
    lib.ndf_mesh_median_filter.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _size_p, _size_p,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_double, *_err,
    ]
    lib.ndf_plan_apply.argtypes = [plan_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, *_err]
    lib.ndf_plan_upload.argtypes = [plan_p, ctypes.c_void_p, ctypes.c_size_t, *_err]
    lib.ndf_plan_compute_download.argtypes = [plan_p, ctypes.c_void_p, ctypes.c_size_t, *_err]
    lib.ndf_plan_apply_device.argtypes = [
        plan_p,
        ctypes.c_void_p,
        ctypes.c_void_p,
        ctypes.c_size_t,
        ctypes.c_void_p,
        *_err,
    ]
    lib.ndf_plan_apply_device_static.argtypes = [
        plan_p,
        ctypes.c_void_p,
        ctypes.c_void_p,
        ctypes.c_size_t,
        ctypes.c_void_p,
        *_err,
    ]
    lib.ndf_plan_set_engine.argtypes = [plan_p, ctypes.c_int32, *_err]
    lib.ndf_plan_describe.argtypes = [plan_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.ndf_plan_destroy.argtypes = [plan_p]
    lib.ndf_plan_destroy.restype = None
    lib.ndf_get_num_threads.argtypes = []
    lib.ndf_get_num_threads.restype = ctypes.c_size_t
    lib.ndf_device_count.argtypes = [_i32_p]
    lib.ndf_device_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.ndf_device_free.argtypes = [ctypes.c_void_p]
    lib.ndf_memcpy_to_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.ndf_memcpy_to_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.ndf_memcpy_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.ndf_device_synchronize.argtypes = []
    lib.ndf_device_fill.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_uint64]
    lib.ndf_plan_last_error.argtypes = [plan_p, *_err]
    lib.ndf_plan_set_output_fill.argtypes = [plan_p, ctypes.c_int32, ctypes.c_uint64]
    lib.ndf_plan_time_device.argtypes = [
        plan_p,
        ctypes.c_void_p,
        ctypes.c_void_p,
        ctypes.c_size_t,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.POINTER(ctypes.c_double),
        ctypes.POINTER(ctypes.c_double),
        *_err,
    ]
    lib.ndf_plan_set_bracket.argtypes = [
        plan_p,
        ctypes.POINTER(ctypes.c_float),
        ctypes.c_size_t,
        ctypes.c_int32,
        *_err,
    ]
    lib.ndf_plan_bracket_report.argtypes = [
        plan_p,
        ctypes.POINTER(ctypes.c_int64),
        ctypes.POINTER(ctypes.c_int64),
    ]
    lib.ndf_plan_bracket_tiles.argtypes = [plan_p]
    lib.ndf_plan_bracket_tiles.restype = ctypes.c_size_t
    _f32_p = ctypes.POINTER(ctypes.c_float)
    lib.ndf_plan_bracket_training_data.argtypes = [
        plan_p,
        ctypes.c_void_p,
        ctypes.c_void_p,
        _f32_p,
        _f32_p,
        _f32_p,
        *_err,
    ]
    lib.ndf_mlp_train.argtypes = [
        _f32_p,
        ctypes.c_size_t,
        _f32_p,
        _f32_p,
        _f32_p,
        ctypes.c_size_t,
        ctypes.c_int32,
        ctypes.c_float,
        ctypes.c_float,
        ctypes.c_float,
        _f32_p,
        *_err,
    ]
    batch_p = ctypes.c_void_p
    lib.ndf_batch_create.argtypes = [
        ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_int32,
        ctypes.c_size_t,
        _size_p,
        _size_p,
        _i32_p,
        ctypes.c_double,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_int32,
        ctypes.c_size_t,
        ctypes.c_size_t,
        *_err,
    ]
    lib.ndf_batch_length.argtypes = [batch_p]
    lib.ndf_batch_length.restype = ctypes.c_size_t
    lib.ndf_batch_configure.argtypes = [batch_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, *_err]
    lib.ndf_batch_autotune.argtypes = [batch_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32,
                                       *_err]
    lib.ndf_batch_apply_device.argtypes = [batch_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_void_p, *_err]
    lib.ndf_batch_time_device.argtypes = [batch_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), *_err]
    lib.ndf_batch_describe.argtypes = [batch_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.ndf_batch_destroy.argtypes = [batch_p]
    lib.ndf_batch_destroy.restype = None
    lib.ndf_fixpix_median.argtypes = [
        ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
        _size_p, _size_p, ctypes.c_void_p, _ssize_p, _i32_p, ctypes.c_double, ctypes.c_int32, ctypes.c_size_t, *_err,
    ]
    lib.ndf_fixpix_upload.argtypes = [plan_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, *_err]
    lib.ndf_fixpix_compute_download.argtypes = [plan_p, ctypes.c_void_p, ctypes.c_size_t, *_err]
    lib.ndf_plan_repair_device.argtypes = [
        plan_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_size_t, ctypes.c_int32, ctypes.c_void_p, *_err,
    ]
    lib.ndf_plan_time_repair_device.argtypes = [
        plan_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_size_t, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), *_err,
    ]
    for name in (
        "ndf_fixpix_median",
        "ndf_fixpix_upload",
        "ndf_fixpix_compute_download",
        "ndf_plan_repair_device",
        "ndf_plan_time_repair_device",
        "ndf_batch_create",
        "ndf_batch_configure",
        "ndf_batch_autotune",
        "ndf_batch_apply_device",
        "ndf_batch_time_device",
        "ndf_batch_describe",
        "ndf_plan_set_bracket",
        "ndf_plan_bracket_report",
        "ndf_plan_bracket_training_data",
        "ndf_mlp_train",
        "ndf_window_filter",
        "ndf_plan_create",
        "ndf_robust_filter",
        "ndf_robust_plan_create",
        "ndf_mesh_median_filter",
        "ndf_plan_apply",
        "ndf_plan_upload",
        "ndf_plan_compute_download",
        "ndf_plan_apply_device",
        "ndf_plan_apply_device_static",
        "ndf_plan_set_engine",
        "ndf_plan_describe",
        "ndf_device_count",
        "ndf_device_malloc",
        "ndf_device_free",
        "ndf_memcpy_to_device",
        "ndf_memcpy_to_host",
        "ndf_memcpy_device",
        "ndf_device_synchronize",
        "ndf_device_fill",
        "ndf_plan_last_error",
        "ndf_plan_set_output_fill",
        "ndf_plan_time_device",
    ):
        getattr(lib, name).restype = ctypes.c_int32


def _open(loader: type) -> ctypes.CDLL:
    if not os.path.exists(LIBRARY):
        raise NativeLibraryMissing(
            f"{LIBRARY} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = loader(LIBRARY)
    _declare(lib)
    return lib


# CDLL releases the GIL during calls (compute); PyDLL keeps it (snapshot).
lib = _open(ctypes.CDLL)
lib_gil = _open(ctypes.PyDLL)
