"""ctypes wrapper of the C oracle in ``oracle/ndf_oracle.c``.

Test infrastructure only: the test suite, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it as the checker. The product path
(``ndfiltrs_amd``) never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from collections.abc import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "_build", "libndf_oracle.so")

MODES = ("reflect", "constant", "nearest", "mirror", "wrap")
STAT_MEDIAN = 0
STAT_RANK = 1
STAT_PERCENTILE = 2
STAT_MIN = 3
STAT_MAX = 4
STATISTICS = {"median": STAT_MEDIAN, "rank": STAT_RANK, "percentile": STAT_PERCENTILE, "min": STAT_MIN,
              "max": STAT_MAX}
METHOD_DIRECT = 0
METHOD_HISTOGRAM = 1

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile and return the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBRARY


def _load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            build()
        lib = ctypes.CDLL(LIBRARY)
        lib.ndfo_map_index.argtypes = [
            ctypes.c_int64,
            ctypes.c_uint64,
            ctypes.c_int32,
            ctypes.POINTER(ctypes.c_int64),
        ]
        lib.ndfo_map_index.restype = ctypes.c_int
        lib.ndfo_midpoint.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.ndfo_midpoint.restype = ctypes.c_double
        lib.ndfo_key32.argtypes = [ctypes.c_uint32]
        lib.ndfo_key32.restype = ctypes.c_uint32
        lib.ndfo_key64.argtypes = [ctypes.c_uint64]
        lib.ndfo_key64.restype = ctypes.c_uint64
        lib.ndfo_unkey32.argtypes = [ctypes.c_uint32]
        lib.ndfo_unkey32.restype = ctypes.c_uint32
        lib.ndfo_unkey64.argtypes = [ctypes.c_uint64]
        lib.ndfo_unkey64.restype = ctypes.c_uint64
        lib.ndfo_filter.argtypes = [
            ctypes.c_int,
            ctypes.c_void_p,
            ctypes.c_void_p,
            ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_size_t),
            ctypes.POINTER(ctypes.c_size_t),
            ctypes.c_void_p,
            ctypes.POINTER(ctypes.c_ssize_t),
            ctypes.POINTER(ctypes.c_int32),
            ctypes.c_double,
            ctypes.c_int32,
            ctypes.c_size_t,
            ctypes.c_int32,
            ctypes.c_int32,
            ctypes.c_size_t,
            ctypes.c_int32,
            ctypes.c_int32,
            ctypes.c_char_p,
            ctypes.c_size_t,
        ]
        lib.ndfo_filter.restype = ctypes.c_int
        lib.ndfo_filter_q.argtypes = lib.ndfo_filter.argtypes[:12] + [ctypes.c_double] + lib.ndfo_filter.argtypes[12:]
        lib.ndfo_filter_q.restype = ctypes.c_int
        lib.ndfo_interpolate.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.ndfo_interpolate.restype = ctypes.c_double
        lib.ndfo_extrema.argtypes = [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
            ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ctypes.c_int32),
            ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_size_t, ctypes.c_char_p,
            ctypes.c_size_t,
        ]
        lib.ndfo_extrema.restype = ctypes.c_int
        _lib = lib
    return _lib


class OracleError(Exception):
    """Error status from the oracle, with the reference's message text."""

    def __init__(self, status: int, message: str) -> None:
        super().__init__(message)
        self.status = status


def map_index(index: int, length: int, mode: str) -> int | None:
    out = ctypes.c_int64(0)
    status = _load().ndfo_map_index(index, length, MODES.index(mode), ctypes.byref(out))
    if status != 0:
        raise OracleError(status, "map_index failed")
    return None if out.value < 0 else out.value


def midpoint(lower: float, upper: float) -> float:
    return _load().ndfo_midpoint(lower, upper)


def interpolate(lower: float, upper: float, fraction: float) -> float:
    return _load().ndfo_interpolate(lower, upper, fraction)


def key32(bits: int) -> int:
    return _load().ndfo_key32(bits)


def key64(bits: int) -> int:
    return _load().ndfo_key64(bits)


def unkey32(key: int) -> int:
    return _load().ndfo_unkey32(key)


def unkey64(key: int) -> int:
    return _load().ndfo_unkey64(key)


def default_threads() -> int:
    value = os.environ.get("RAYON_NUM_THREADS")
    if value:
        return max(1, int(value))
    return os.cpu_count() or 1


def window_filter(
    array: np.ndarray,
    size: Sequence[int],
    *,
    footprint: Sequence[bool] | np.ndarray | None = None,
    origin: Sequence[int] | None = None,
    mode: Sequence[str] | str = "reflect",
    cval: float = 0.0,
    statistic: str = "median",
    rank: int = 0,
    q: float = 0.0,
    omit_nan: bool = False,
    ignore_inf: bool = False,
    min_count: int = 1,
    method: str = "direct",
    threads: int | None = None,
) -> np.ndarray:
    """Run the oracle on a C-contiguous float32/float64 array.

    ``statistic`` is median, rank (``rank``), percentile (``q`` in 0..100),
    min or max. ``method`` is direct, histogram (median and rank) or
    separable (min and max over a rectangle)."""
    array = np.ascontiguousarray(array)
    if array.dtype not in (np.float32, np.float64):
        raise TypeError("oracle supports float32 and float64")
    ndim = array.ndim
    if isinstance(mode, str):
        mode = (mode,) * ndim
    if origin is None:
        origin = (0,) * ndim
    out = np.empty_like(array)
    shape = (ctypes.c_size_t * max(ndim, 1))(*array.shape)
    sizes = (ctypes.c_size_t * max(ndim, 1))(*size)
    origins = (ctypes.c_ssize_t * max(ndim, 1))(*origin)
    modes = (ctypes.c_int32 * max(ndim, 1))(*(MODES.index(m) for m in mode))
    fp = None
    if footprint is not None:
        fp_arr = np.ascontiguousarray(np.asarray(footprint).ravel() != 0, dtype=np.uint8)
        fp = fp_arr.ctypes.data
    err = ctypes.create_string_buffer(256)
    if method == "separable":
        if statistic not in ("min", "max") or footprint is not None:
            raise ValueError("the separable method serves rectangular min and max only")
        status = _load().ndfo_extrema(
            array.dtype.itemsize, array.ctypes.data, out.ctypes.data, ndim, shape, sizes, origins, modes,
            float(cval), int(statistic == "max"), int(bool(omit_nan)), int(bool(ignore_inf)), int(min_count), err,
            len(err),
        )
        if status != 0:
            raise OracleError(status, err.value.decode())
        return out
    status = _load().ndfo_filter_q(
        array.dtype.itemsize,
        array.ctypes.data,
        out.ctypes.data,
        ndim,
        shape,
        sizes,
        fp,
        origins,
        modes,
        float(cval),
        STATISTICS[statistic],
        int(rank),
        float(q),
        int(bool(omit_nan)),
        int(bool(ignore_inf)),
        int(min_count),
        METHOD_HISTOGRAM if method == "histogram" else METHOD_DIRECT,
        default_threads() if threads is None else int(threads),
        err,
        len(err),
    )
    if status != 0:
        raise OracleError(status, err.value.decode())
    return out


def median(array: np.ndarray, size, **kwargs) -> np.ndarray:
    return window_filter(array, size, statistic="median", **kwargs)


def fixpix_median(
    array: np.ndarray,
    mask: np.ndarray,
    size: Sequence[int],
    *,
    footprint: Sequence[bool] | np.ndarray | None = None,
    origin: Sequence[int] | None = None,
    mode: Sequence[str] | str = "reflect",
    cval: float = float("nan"),
    ignore_inf: bool = False,
    min_count: int = 1,
    threads: int | None = None,
) -> np.ndarray:
    """Mask-directed median repair (``ndfo_fixpix_median``) of a C-contiguous array."""
    array = np.ascontiguousarray(array)
    if array.dtype not in (np.float32, np.float64):
        raise TypeError("oracle supports float32 and float64")
    bad = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
    ndim = array.ndim
    if isinstance(mode, str):
        mode = (mode,) * ndim
    if origin is None:
        origin = (0,) * ndim
    out = np.empty_like(array)
    lib = _load()
    if not hasattr(lib, "_fixpix_declared"):
        lib.ndfo_fixpix_median.argtypes = [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p,
            ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ctypes.c_int32), ctypes.c_double, ctypes.c_int32,
            ctypes.c_size_t, ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t,
        ]
        lib.ndfo_fixpix_median.restype = ctypes.c_int
        lib._fixpix_declared = True
    shape = (ctypes.c_size_t * max(ndim, 1))(*array.shape)
    sizes = (ctypes.c_size_t * max(ndim, 1))(*size)
    origins = (ctypes.c_ssize_t * max(ndim, 1))(*origin)
    modes = (ctypes.c_int32 * max(ndim, 1))(*(MODES.index(m) for m in mode))
    fp = None
    if footprint is not None:
        fp_arr = np.ascontiguousarray(np.asarray(footprint).ravel() != 0, dtype=np.uint8)
        fp = fp_arr.ctypes.data
    err = ctypes.create_string_buffer(256)
    status = lib.ndfo_fixpix_median(
        array.dtype.itemsize, array.ctypes.data, out.ctypes.data, bad.ctypes.data, bad.size, ndim, shape, sizes,
        fp, origins, modes, float(cval), int(bool(ignore_inf)), int(min_count),
        default_threads() if threads is None else int(threads), err, len(err),
    )
    if status != 0:
        raise OracleError(status, err.value.decode())
    return out


MESH_EDGES = ("shrink", "truncate")
MESH_THRESHOLDS = ("all", "deviation", "above")


def mesh_filter(background: np.ndarray, rms: np.ndarray, filter_size: Sequence[int], *, edge: str = "shrink",
                threshold_kind: str = "all", threshold: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """Median filter of float64 mesh grids (``ndfo_mesh_filter``)."""
    background = np.ascontiguousarray(background, dtype=np.float64)
    rms = np.ascontiguousarray(rms, dtype=np.float64)
    lib = _load()
    if not hasattr(lib, "_mesh_declared"):
        lib.ndfo_mesh_filter.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t), ctypes.c_int32, ctypes.c_int32,
            ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t,
        ]
        lib.ndfo_mesh_filter.restype = ctypes.c_int
        lib._mesh_declared = True
    ndim = background.ndim
    out_b = np.empty_like(background)
    out_r = np.empty_like(rms)
    err = ctypes.create_string_buffer(256)
    status = lib.ndfo_mesh_filter(
        background.ctypes.data, rms.ctypes.data, out_b.ctypes.data, out_r.ctypes.data, ndim,
        (ctypes.c_size_t * ndim)(*background.shape), (ctypes.c_size_t * ndim)(*filter_size),
        MESH_EDGES.index(edge), MESH_THRESHOLDS.index(threshold_kind), float(threshold), err, len(err),
    )
    if status != 0:
        raise OracleError(status, err.value.decode())
    return out_b, out_r

