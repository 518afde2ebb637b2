"""Batches of 2-D images with per-image windows and boundary modes.

One :class:`Batch` validates every image as the single-image API would, then
filters all of them in one persistent launch (a work queue over every tile of
every image), captured once in a hipGraph and replayed. ``autotune`` times
the (sort length, tile-row cap) candidates on real buffers and keeps the
fastest. There is no reference counterpart: ndfiltrs filters one array per
call, so the reference-side equivalent is one ``nanmedian_filter`` call per
image.
"""

from __future__ import annotations

import ctypes
import json
from collections.abc import Sequence

import numpy as np

from . import _core, _lib


class Batch:
    def __init__(
        self,
        dtype,
        shapes: Sequence[Sequence[int]],
        sizes: Sequence[Sequence[int] | int],
        modes: Sequence[Sequence[str] | str],
        kind: str = "median",
        cval: float = 0.0,
        omit_nan: bool = True,
        ignore_inf: bool = False,
        min_count: int = 1,
        parameter: int = 0,
    ) -> None:
        self._handle = None
        dtype = np.dtype(dtype)
        if kind not in _lib.KINDS:
            raise ValueError("unsupported window filter kind")
        count = len(shapes)
        if len(sizes) != count or len(modes) != count:
            raise ValueError("shapes, sizes and modes must have one entry per image")
        flat_shapes, flat_sizes, flat_modes = [], [], []
        for shape, size, mode in zip(shapes, sizes, modes):
            if len(shape) != 2:
                raise ValueError("batch images must be 2-D")
            size = [int(size)] * 2 if np.ndim(size) == 0 else [int(v) for v in size]
            mode = [mode] * 2 if isinstance(mode, str) else list(mode)
            if len(size) != 2 or len(mode) != 2:
                raise ValueError("array, window, origin, and mode must have equal rank")
            for value in mode:
                if value not in _lib.MODES:
                    raise ValueError("unsupported boundary mode")
            flat_shapes += [int(v) for v in shape]
            flat_sizes += size
            flat_modes += [_lib.MODES.index(v) for v in mode]
        self.dtype = dtype
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.offsets = np.concatenate([[0], np.cumsum([a * b for a, b in self.shapes])]).astype(np.int64)
        err = ctypes.create_string_buffer(512)
        handle = ctypes.c_void_p()
        status = _lib.lib.ndf_batch_create(
            ctypes.byref(handle), dtype.itemsize, count,
            (ctypes.c_size_t * len(flat_shapes))(*flat_shapes), (ctypes.c_size_t * len(flat_sizes))(*flat_sizes),
            (ctypes.c_int32 * len(flat_modes))(*flat_modes), float(cval), _lib.KINDS[kind], int(bool(omit_nan)),
            int(bool(ignore_inf)), int(min_count), int(parameter), err, len(err),
        )
        if status != _lib.OK:
            _core._raise(status, err.value.decode())
        self._handle = handle
        self.len = int(_lib.lib.ndf_batch_length(handle))

    def __del__(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            _lib.lib.ndf_batch_destroy(handle)
            self._handle = None

    def _check(self, status: int, err) -> None:
        if status != _lib.OK:
            _core._raise(status, err.value.decode())

    def configure(self, sort_len: int = 2048, max_tile_rows: int = 16, use_graph: bool = True) -> None:
        err = ctypes.create_string_buffer(256)
        self._check(_lib.lib.ndf_batch_configure(self._handle, sort_len, max_tile_rows, int(use_graph), err,
                                                 len(err)), err)

    def autotune(self, d_input: int, d_output: int, repeats: int = 3) -> dict:
        err = ctypes.create_string_buffer(256)
        self._check(_lib.lib.ndf_batch_autotune(self._handle, d_input, d_output, self.len, repeats, err, len(err)),
                    err)
        return self.describe()

    def apply_device(self, d_input: int, d_output: int, stream: int | None = None) -> None:
        err = ctypes.create_string_buffer(256)
        self._check(_lib.lib.ndf_batch_apply_device(self._handle, d_input, d_output, self.len, stream, err,
                                                    len(err)), err)

    def time_device(self, d_input: int, d_output: int, warmup: int, iterations: int) -> float:
        err = ctypes.create_string_buffer(256)
        ms = ctypes.c_double(0.0)
        self._check(_lib.lib.ndf_batch_time_device(self._handle, d_input, d_output, self.len, warmup, iterations,
                                                   ctypes.byref(ms), err, len(err)), err)
        return ms.value

    def describe(self) -> dict:
        buffer = ctypes.create_string_buffer(4096)
        _lib.lib.ndf_batch_describe(self._handle, buffer, len(buffer))
        return json.loads(buffer.value.decode())

    def pack(self, images: Sequence[np.ndarray]) -> np.ndarray:
        """Concatenate images (C order) into one flat buffer."""
        if len(images) != len(self.shapes):
            raise ValueError("one image per batch entry")
        for image, shape in zip(images, self.shapes):
            if image.shape != shape or image.dtype != self.dtype:
                raise ValueError("input length does not match the configured shape")
        return np.concatenate([np.ascontiguousarray(i).ravel() for i in images]) if self.len else \
            np.zeros(0, self.dtype)

    def unpack(self, flat: np.ndarray) -> list[np.ndarray]:
        return [flat[self.offsets[i]:self.offsets[i + 1]].reshape(shape).copy()
                for i, shape in enumerate(self.shapes)]

    def apply(self, images: Sequence[np.ndarray]) -> list[np.ndarray]:
        """Filter host images through device buffers (blocking)."""
        flat = self.pack(images)
        d_in = _core.DeviceArray(flat.nbytes)
        d_out = _core.DeviceArray(flat.nbytes)
        d_in.upload(flat)
        poison = _core._OUTPUT_POISON
        if poison is not None and flat.size:
            d_out.fill(poison[0 if flat.itemsize == 4 else 1], flat.itemsize)
        self.apply_device(d_in.address, d_out.address)
        out = d_out.download(np.empty_like(flat))
        d_in.free()
        d_out.free()
        return self.unpack(out)
