"""Native layer: the MI355X counterpart of ndfiltrs' private ``_core`` module.

``window_filter`` keeps the reference signature
(``src/pyapi/filter.rs:265-314``, ``python/ndfiltrs/_core.pyi:34-46``) and its
error behaviour, and runs on the GPU through ``libndfgpu.so``. The input is
snapshotted while the GIL is held (``src/pyapi/filter.rs:219-229``); the GIL is
released while the device computes.
"""

from __future__ import annotations

import atexit
import ctypes
import json
import os
import threading
from collections import OrderedDict
from collections.abc import Sequence

import numpy as np

from . import _lib

_MSG_HIST_FOOTPRINT = "histogram reductions require rectangular size, not footprint"
_MSG_RANK0 = "filtering requires at least one dimension"
_MSG_EQUAL_RANK = "array, window, origin, and mode must have equal rank"
_MSG_SIZE0 = "window sizes must be positive"
_MSG_LENGTH = "input length does not match the configured shape"
_MSG_OVERFLOW = "shape or index arithmetic overflowed"


# Test hook: when set (a pair of uint32 / uint64 bit patterns), every plan
# created here fills its device output with that pattern before each
# host-buffer launch, so an output element no kernel writes cannot inherit a
# correct value from an earlier launch.
_OUTPUT_POISON: tuple[int, int] | None = None


def set_output_poison(bits: tuple[int, int] | None) -> None:
    """Enable (``(float32_bits, float64_bits)``) or disable (``None``) the
    output poison of plans created from now on."""
    global _OUTPUT_POISON
    _OUTPUT_POISON = None if bits is None else (int(bits[0]), int(bits[1]))


def _raise(status: int, message: str) -> None:
    """Map a status to the exception class the reference raises
    (``src/pyapi/filter.rs:37-45``)."""
    if status == _lib.ERR_MEMORY:
        raise MemoryError(message)
    if status == _lib.ERR_OVERFLOW:
        raise OverflowError(message)
    if status == _lib.ERR_DEVICE:
        raise RuntimeError(message)
    raise ValueError(message)


def _as_sizes(values: Sequence[int], name: str) -> list[int]:
    out = []
    for value in values:
        value = int(value)
        if value < 0:
            raise OverflowError(f"can't convert negative int to unsigned ({name})")
        out.append(value)
    return out


def _typed_array(array: object) -> np.ndarray:
    if (
        not isinstance(array, np.ndarray)
        or array.dtype not in (np.float32, np.float64)
        or not array.dtype.isnative
    ):
        raise TypeError("argument 'array': expected a float32 or float64 numpy.ndarray")
    return array


class Plan:
    """A validated filter for one shape (``WindowFilter``, ``src/plan.rs:236-428``).

    Owns a native plan; ``apply`` filters host arrays, ``apply_device``
    device buffers.
    """

    def __init__(
        self,
        dtype: np.dtype,
        shape: Sequence[int],
        kind: str,
        size: Sequence[int],
        footprint: Sequence[bool] | None,
        origin: Sequence[int],
        mode: Sequence[str],
        cval: float,
        omit_nan: bool,
        ignore_inf: bool,
        min_count: int,
        parameter: int | float,
    ) -> None:
        """``kind`` is a window-filter kind (``parameter`` the nonnegative
        rank of the rank kinds) or ``"percentile"``, the robust percentile
        (``RobustConfig``, ``src/robust.rs:698-722``) with ``parameter`` the
        percentile in [0, 100]."""
        self._handle = None
        dtype = np.dtype(dtype)
        robust = kind == "percentile"
        if not robust and kind not in _lib.KINDS:
            raise ValueError("unsupported window filter kind")
        histogram = kind in ("median_histogram", "rank_histogram")
        ndim = len(shape)
        size = _as_sizes(size, "size")
        origin = [int(value) for value in origin]
        min_count = _as_sizes([min_count], "min_count")[0]
        parameter = float(parameter) if robust else _as_sizes([parameter], "parameter")[0]
        modes = []
        for value in mode:
            if value not in _lib.MODES:
                raise ValueError("unsupported boundary mode")
            modes.append(_lib.MODES.index(value))
        if histogram and footprint is not None:
            raise ValueError(_MSG_HIST_FOOTPRINT)
        if ndim == 0:
            raise ValueError(_MSG_RANK0)
        if len(size) != ndim or len(origin) != ndim or len(modes) != ndim:
            raise ValueError(_MSG_EQUAL_RANK)
        fp = None
        if footprint is not None:
            flags = [bool(value) for value in footprint]
            if any(value == 0 for value in size):
                raise ValueError(_MSG_SIZE0)
            volume = 1
            for value in size:
                volume *= value
            if volume >= 2**64:
                raise OverflowError(_MSG_OVERFLOW)
            if len(flags) != volume:
                raise ValueError(_MSG_LENGTH)
            fp = np.asarray(flags, dtype=np.uint8)
        self._footprint = fp
        self.dtype = dtype
        self.shape = tuple(int(v) for v in shape)
        self.kind = kind
        self.len = int(np.prod(self.shape, dtype=np.int64)) if ndim else 0
        c_shape = (ctypes.c_size_t * ndim)(*self.shape)
        c_size = (ctypes.c_size_t * ndim)(*size)
        c_origin = (ctypes.c_ssize_t * ndim)(*origin)
        c_modes = (ctypes.c_int32 * ndim)(*modes)
        err = ctypes.create_string_buffer(256)
        handle = ctypes.c_void_p()
        if robust:
            status = _lib.lib.ndf_robust_plan_create(
                ctypes.byref(handle), dtype.itemsize, ndim, c_shape, c_size,
                None if fp is None else fp.ctypes.data, c_origin, c_modes, float(cval), _lib.ROBUST_KINDS[kind],
                parameter, int(bool(omit_nan)), int(bool(ignore_inf)), min_count, err, len(err),
            )
        else:
            status = _lib.lib.ndf_plan_create(
                ctypes.byref(handle),
                dtype.itemsize,
                ndim,
                c_shape,
                c_size,
                None if fp is None else fp.ctypes.data,
                c_origin,
                c_modes,
                float(cval),
                _lib.KINDS[kind],
                int(bool(omit_nan)),
                int(bool(ignore_inf)),
                min_count,
                parameter,
                err,
                len(err),
            )
        if status != _lib.OK:
            _raise(status, err.value.decode())
        self._handle = handle
        if _OUTPUT_POISON is not None:
            self.set_output_fill(_OUTPUT_POISON[0 if dtype.itemsize == 4 else 1])

    def set_output_fill(self, bits: int | None) -> None:
        """Fill the device output with ``bits`` before every host-buffer
        launch (``None`` turns it off)."""
        _lib.lib.ndf_plan_set_output_fill(self._handle, int(bits is not None), int(bits or 0))

    def last_error(self) -> None:
        """Wait for the last launch; raise if a kernel reported an internal error."""
        err = ctypes.create_string_buffer(256)
        status = _lib.lib.ndf_plan_last_error(self._handle, err, len(err))
        if status != _lib.OK:
            _raise(status, err.value.decode())

    def __del__(self) -> None:
        handle = getattr(self, "_handle", None)
        library = getattr(_lib, "lib", None) if _lib is not None else None
        if handle is not None and handle.value and library is not None:
            library.ndf_plan_destroy(handle)
            self._handle = None

    def set_engine(self, engine: str) -> None:
        err = ctypes.create_string_buffer(256)
        status = _lib.lib.ndf_plan_set_engine(self._handle, _lib.ENGINES[engine], err, len(err))
        if status != _lib.OK:
            _raise(status, err.value.decode())

    def describe(self) -> str:
        buffer = ctypes.create_string_buffer(512)
        _lib.lib.ndf_plan_describe(self._handle, buffer, len(buffer))
        return buffer.value.decode()

    def apply(self, array: np.ndarray) -> np.ndarray:
        """Filter a C-contiguous host array into a new owned array."""
        if array.shape != self.shape or array.dtype != self.dtype:
            raise ValueError(_MSG_LENGTH)
        err = ctypes.create_string_buffer(256)
        if self.len:
            # Snapshot: copied to the device while the GIL is held.
            status = _lib.lib_gil.ndf_plan_upload(self._handle, array.ctypes.data, self.len, err, len(err))
            if status != _lib.OK:
                _raise(status, err.value.decode())
        # PyArray_Zeros in the reference: a new owned C-contiguous array.
        output = np.zeros(self.shape, dtype=self.dtype)
        if self.len:
            status = _lib.lib.ndf_plan_compute_download(
                self._handle, output.ctypes.data, self.len, err, len(err)
            )
            if status != _lib.OK:
                _raise(status, err.value.decode())
        return output

    def apply_device(self, d_input: int, d_output: int, stream: int | None = None) -> None:
        """Filter device buffers (integer addresses) asynchronously on ``stream``."""
        err = ctypes.create_string_buffer(256)
        status = _lib.lib.ndf_plan_apply_device(
            self._handle, d_input, d_output, self.len, stream, err, len(err)
        )
        if status != _lib.OK:
            _raise(status, err.value.decode())

    def set_bracket(self, weights: np.ndarray | None, sort_len: int = 4096) -> None:
        """Enable the learned bracket (a 1346-float model) or disable it (None)."""
        err = ctypes.create_string_buffer(256)
        if weights is None:
            status = _lib.lib.ndf_plan_set_bracket(self._handle, None, 0, 0, err, len(err))
        else:
            w = np.ascontiguousarray(weights, dtype=np.float32)
            status = _lib.lib.ndf_plan_set_bracket(
                self._handle, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size, int(sort_len), err, len(err)
            )
        if status != _lib.OK:
            _raise(status, err.value.decode())

    def bracket_report(self) -> tuple[int, int]:
        """(tiles, tiles recomputed by the plain kernel) of the last launch."""
        tiles = ctypes.c_int64(0)
        fallback = ctypes.c_int64(0)
        status = _lib.lib.ndf_plan_bracket_report(self._handle, ctypes.byref(tiles), ctypes.byref(fallback))
        if status != _lib.OK:
            raise RuntimeError("bracket report failed")
        return int(tiles.value), int(fallback.value)

    def bracket_training_data(self, d_input: int, d_output: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Per-tile features, standardised target range and weight (see ndfgpu.h)."""
        tiles = int(_lib.lib.ndf_plan_bracket_tiles(self._handle))
        features = np.zeros((tiles, 6), dtype=np.float32)
        targets = np.zeros((tiles, 2), dtype=np.float32)
        weight = np.zeros(tiles, dtype=np.float32)
        err = ctypes.create_string_buffer(256)
        f32 = ctypes.POINTER(ctypes.c_float)
        status = _lib.lib.ndf_plan_bracket_training_data(
            self._handle,
            d_input,
            d_output,
            features.ctypes.data_as(f32),
            targets.ctypes.data_as(f32),
            weight.ctypes.data_as(f32),
            err,
            len(err),
        )
        if status != _lib.OK:
            _raise(status, err.value.decode())
        return features, targets, weight

    def repair_device(self, d_input: int, d_output: int, d_snapshot: int, d_positions: int, npos: int,
                      path: str = "auto", stream: int | None = None) -> None:
        """Mask-directed repair of device buffers (see ``ndf_plan_repair_device``);
        ``d_positions`` holds ``npos`` int64 flat indices."""
        err = ctypes.create_string_buffer(256)
        status = _lib.lib.ndf_plan_repair_device(
            self._handle, d_input, d_output, d_snapshot, d_positions, npos, self.len, _lib.REPAIR_PATHS[path],
            stream, err, len(err)
        )
        if status != _lib.OK:
            _raise(status, err.value.decode())

    def time_repair_device(self, d_input: int, d_output: int, d_snapshot: int, d_positions: int, npos: int,
                           path: str, warmup: int, iterations: int) -> float:
        """Mean milliseconds per repair (HIP events on one stream)."""
        err = ctypes.create_string_buffer(256)
        ms = ctypes.c_double(0.0)
        status = _lib.lib.ndf_plan_time_repair_device(
            self._handle, d_input, d_output, d_snapshot, d_positions, npos, self.len, _lib.REPAIR_PATHS[path],
            warmup, iterations, ctypes.byref(ms), err, len(err)
        )
        if status != _lib.OK:
            _raise(status, err.value.decode())
        return ms.value

    def time_device(self, d_input: int, d_output: int, warmup: int, iterations: int,
                    main_kernel: bool = False):
        """Mean milliseconds per launch (HIP events on one stream); with
        ``main_kernel`` also the mean duration of the launch's main kernel."""
        err = ctypes.create_string_buffer(256)
        ms = ctypes.c_double(0.0)
        main = ctypes.c_double(0.0)
        status = _lib.lib.ndf_plan_time_device(
            self._handle, d_input, d_output, self.len, warmup, iterations, ctypes.byref(ms), ctypes.byref(main),
            err, len(err)
        )
        if status != _lib.OK:
            _raise(status, err.value.decode())
        return (ms.value, main.value) if main_kernel else ms.value


# ---------------------------------------------------------------------------
# Plan reuse. A plan owns its device buffers and pinned staging, and planning
# a large window takes milliseconds, so public calls reuse plans keyed by
# everything that defines one (the reference plans per call but documents
# WindowFilter reuse, src/plan.rs:199-204). A plan is taken out of the cache
# for the duration of a call, so concurrent calls never share one.
# ---------------------------------------------------------------------------

_CACHE_PLANS = 4
_CACHE_BYTES = 4 << 30


class _PlanCache:
    def __init__(self) -> None:
        self._lock = threading.Lock()
        self._idle: OrderedDict[tuple, list[tuple[Plan, int]]] = OrderedDict()

    def take(self, key: tuple) -> Plan | None:
        with self._lock:
            plans = self._idle.get(key)
            if not plans:
                return None
            plan, _ = plans.pop()
            if not plans:
                del self._idle[key]
            return plan

    def give(self, key: tuple, plan: Plan, nbytes: int) -> None:
        with self._lock:
            self._idle.setdefault(key, []).append((plan, nbytes))
            self._idle.move_to_end(key)
            while self._idle:
                count = sum(len(v) for v in self._idle.values())
                total = sum(b for v in self._idle.values() for _, b in v)
                if count <= _CACHE_PLANS and total <= _CACHE_BYTES:
                    break
                oldest = next(iter(self._idle))
                victims = self._idle[oldest]
                victims.pop(0)
                if not victims:
                    del self._idle[oldest]

    def clear(self) -> None:
        with self._lock:
            self._idle.clear()


_PLANS = _PlanCache()
# Plans release their device memory before the interpreter tears modules down.
atexit.register(lambda: _PLANS.clear())


def clear_plan_cache() -> None:
    """Release the plans (and their device memory) kept for reuse."""
    _PLANS.clear()


# Learned-bracket models shipped with the package, by (dtype, window): a
# public call whose plan runs the slide engine on such a window uses the
# model (results are identical with or without it).
_MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")
_SHIPPED_MODELS = {
    ("float32", (63, 63)): "bracket_sky_63x63.json",
    ("float64", (31, 31)): "bracket_sky_31x31_f64.json",
}
_MODEL_WEIGHTS: dict[str, tuple[np.ndarray, dict]] = {}
_BRACKET_MIN_PIXELS = 1 << 20


def _shipped_model(dtype: np.dtype, size: Sequence[int]):
    """(weights, meta) of the shipped bracket model for this dtype and window,
    or None."""
    name = _SHIPPED_MODELS.get((np.dtype(dtype).name, tuple(int(v) for v in size)))
    if name is None:
        return None
    if name not in _MODEL_WEIGHTS:
        path = os.path.join(_MODEL_DIR, name)
        if not os.path.exists(path):
            return None
        with open(path) as handle:
            data = json.load(handle)
        _MODEL_WEIGHTS[name] = (np.asarray(data["weights"], dtype=np.float32), data.get("meta", {}))
    return _MODEL_WEIGHTS[name]


def _use_shipped_bracket(plan: "Plan", typed: np.ndarray, kind: str, size, footprint, omit_nan: bool) -> None:
    """Give a new plan the shipped bracket model when one matches: a 2-D
    rectangular NaN-omitting window on the slide engine, at least 2^20
    pixels. Results never depend on it (tiles it misses are recomputed)."""
    if (footprint is None and typed.ndim == 2 and omit_nan and typed.size >= _BRACKET_MIN_PIXELS
            and '"engine": "slide"' in plan.describe()):
        model = _shipped_model(typed.dtype, size)
        if model is not None:
            from ndfiltrs_amd import bracket

            try:
                plan.set_bracket(model[0], bracket.sort_len(model[1], kind))
            except ValueError:
                pass


def _cached_apply(typed: np.ndarray, kind: str, size, footprint, origin, mode, cval: float, omit_nan: bool,
                  ignore_inf: bool, min_count: int, parameter) -> np.ndarray:
    """Run one public call on a cached (or new) plan."""
    key = (kind, typed.dtype.str, typed.shape, tuple(int(v) for v in size),
           None if footprint is None else tuple(bool(v) for v in footprint), tuple(int(v) for v in origin),
           tuple(mode), np.float64(cval).tobytes(), bool(omit_nan), bool(ignore_inf), int(min_count),
           float(parameter) if kind == "percentile" else int(parameter))
    plan = _PLANS.take(key)
    if plan is None:
        plan = Plan(typed.dtype, typed.shape, kind, size, footprint, origin, mode, cval, omit_nan, ignore_inf,
                    min_count, parameter)
        _use_shipped_bracket(plan, typed, kind, size, footprint, omit_nan)
    output = plan.apply(typed)
    _PLANS.give(key, plan, 2 * typed.nbytes)
    return output


def window_filter(
    array: np.ndarray,
    kind: str,
    size: Sequence[int],
    footprint: Sequence[bool] | None,
    origin: Sequence[int],
    mode: Sequence[str],
    cval: float,
    omit_nan: bool,
    ignore_inf: bool,
    min_count: int,
    parameter: int,
) -> np.ndarray:
    """Apply one window reduction on the GPU.

    ``kind`` is one of ``median``, ``median_network``, ``median_histogram``,
    ``rank``, ``rank_histogram``, ``min`` or ``max``; ``parameter`` is the
    nonnegative rank for the rank kinds. The input must be C-contiguous. The
    reference's other kinds (means, variances) are not on this path and give
    ``ValueError("unsupported window filter kind")``.
    """
    if kind not in _lib.KINDS:
        raise ValueError("unsupported window filter kind")
    typed = _typed_array(array)
    if not typed.flags.c_contiguous:
        raise ValueError("native input must be C-contiguous")
    return _cached_apply(typed, kind, size, footprint, origin, mode, cval, omit_nan, ignore_inf, min_count,
                         parameter)


def robust_filter(
    array: np.ndarray,
    size: Sequence[int],
    footprint: Sequence[bool] | None,
    origin: Sequence[int],
    mode: Sequence[str],
    cval: float,
    kind: str,
    parameter: float,
    omit_nan: bool,
    ignore_inf: bool,
    min_count: int,
    sigma_lower: float,
    sigma_upper: float,
    maxiters: int | None,
    center: str,
    spread: str,
    estimator: str = "sex",
    modify_sample_size: bool = False,
    pearson: float = 2.5,
) -> np.ndarray:
    """Apply one robust window reduction on the GPU (``_core.robust_filter``,
    ``src/pyapi/robust.rs:86-158``).

    ``kind="percentile"`` (``parameter`` = q in [0, 100]) is implemented: the
    linear-interpolated percentile of the retained samples
    (``src/robust.rs:320-367``). The reference's other robust kinds (MAD,
    mode, biweight, sigma clipping) are not on this path and raise
    ``NotImplementedError``; the clipping arguments are accepted for
    signature compatibility.
    """
    if kind in _lib.ROBUST_OTHER_KINDS:
        raise NotImplementedError(f"robust filter kind {kind!r} is not implemented in ndfiltrs_amd")
    if kind not in _lib.ROBUST_KINDS:
        raise ValueError("unsupported robust filter kind")
    if not omit_nan and (ignore_inf or int(min_count) != 1):
        raise ValueError("ignore_inf and min_count require NaN omission")
    typed = _typed_array(array)
    if not typed.flags.c_contiguous or not typed.flags.aligned:
        raise ValueError("native input must be aligned and C-contiguous")
    return _cached_apply(typed, kind, size, footprint, origin, mode, cval, omit_nan, ignore_inf, min_count,
                         parameter)


def fixpix_median(
    array: np.ndarray,
    mask: np.ndarray,
    size: Sequence[int],
    footprint: Sequence[bool] | None,
    origin: Sequence[int],
    mode: Sequence[str],
    cval: float,
    ignore_inf: bool,
    min_count: int,
) -> np.ndarray:
    """Median repair of the masked pixels (``_core.fixpix_median``,
    ``src/pyapi/repair.rs:125-183``) on the GPU.

    ``mask`` is uint8 with one entry per element (nonzero = bad). The data
    and mask are snapshotted while the GIL is held; the GIL is released for
    the device work. Returns a new owned C-contiguous array.
    """
    typed = _typed_array(array)
    bad = np.asarray(mask)
    if bad.dtype != np.uint8:
        raise TypeError("argument 'mask': expected a uint8 numpy.ndarray")
    ndim = typed.ndim
    size = _as_sizes(size, "size")
    origin = [int(value) for value in origin]
    min_count = _as_sizes([min_count], "min_count")[0]
    modes = []
    for value in mode:
        if value not in _lib.MODES:
            raise ValueError("unsupported boundary mode")
        modes.append(_lib.MODES.index(value))
    if ndim == 0:
        raise ValueError(_MSG_RANK0)
    if len(size) != ndim or len(origin) != ndim or len(modes) != ndim:
        raise ValueError(_MSG_EQUAL_RANK)
    fp = None
    if footprint is not None:
        fp = np.asarray([bool(value) for value in footprint], dtype=np.uint8)
        volume = 1
        for value in size:
            volume *= value
        if 0 in size:
            raise ValueError(_MSG_SIZE0)
        if len(fp) != volume:
            raise ValueError(_MSG_LENGTH)
    if fp is not None:
        fp = fp.astype(bool).tolist()
    data = typed if typed.flags.c_contiguous else np.ascontiguousarray(typed)
    flags = bad if bad.flags.c_contiguous else np.ascontiguousarray(bad)
    if flags.size != data.size:
        raise ValueError(_MSG_LENGTH)
    key = ("fixpix", data.dtype.str, data.shape, tuple(size), None if fp is None else tuple(fp), tuple(origin),
           tuple(mode), np.float64(cval).tobytes(), bool(ignore_inf), int(min_count))
    plan = _PLANS.take(key)
    if plan is None:
        plan = Plan(data.dtype, data.shape, "median", size, fp, origin, mode, cval, True, ignore_inf, min_count, 0)
        # The dense repair runs the plan's whole-array launch, bracket included.
        _use_shipped_bracket(plan, data, "median", size, fp, True)
    err = ctypes.create_string_buffer(256)
    # Snapshot (GIL held): the data and the mask go to the device; the mask's
    # positions are compacted there.
    status = _lib.lib_gil.ndf_fixpix_upload(plan._handle, data.ctypes.data, flags.ctypes.data, plan.len, err,
                                            len(err))
    if status != _lib.OK:
        _raise(status, err.value.decode())
    # Every element is written by the native call (copied or repaired).
    output = np.empty(data.shape, dtype=data.dtype)
    status = _lib.lib.ndf_fixpix_compute_download(plan._handle, output.ctypes.data, plan.len, err, len(err))
    if status != _lib.OK:
        _raise(status, err.value.decode())
    _PLANS.give(key, plan, 2 * data.nbytes)
    return output


def get_num_threads() -> int:
    """Host thread count (``RAYON_NUM_THREADS`` or the CPU count)."""
    return int(_lib.lib.ndf_get_num_threads())


class DeviceArray:
    """A device allocation owned by Python (benchmark and test plumbing)."""

    def __init__(self, nbytes: int) -> None:
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        if _lib.lib.ndf_device_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) != _lib.OK:
            raise RuntimeError("device allocation failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        if _lib.lib.ndf_memcpy_to_device(self.ptr, array.ctypes.data, array.nbytes) != _lib.OK:
            raise RuntimeError("host-to-device copy failed")

    def download(self, array: np.ndarray) -> np.ndarray:
        if array.nbytes > self.nbytes or not array.flags.c_contiguous:
            raise ValueError("bad destination array")
        if _lib.lib.ndf_memcpy_to_host(array.ctypes.data, self.ptr, array.nbytes) != _lib.OK:
            raise RuntimeError("device-to-host copy failed")
        return array

    def fill(self, bits: int, element_bytes: int, count: int | None = None) -> None:
        """Fill ``count`` elements (default: the whole buffer) with ``bits``."""
        count = self.nbytes // element_bytes if count is None else int(count)
        if count * element_bytes > self.nbytes:
            raise ValueError("fill larger than the device buffer")
        if _lib.lib.ndf_device_fill(self.ptr, count, element_bytes, int(bits)) != _lib.OK:
            raise RuntimeError("device fill failed")

    def free(self) -> None:
        if self.ptr.value:
            _lib.lib.ndf_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self) -> None:
        self.free()


def device_count() -> int:
    count = ctypes.c_int32(0)
    _lib.lib.ndf_device_count(ctypes.byref(count))
    return int(count.value)


def synchronize() -> None:
    if _lib.lib.ndf_device_synchronize() != _lib.OK:
        raise RuntimeError("device synchronize failed")
