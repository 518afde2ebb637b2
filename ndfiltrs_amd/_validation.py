"""Argument validation for the public filters.

Behaviour (accepted inputs, exception classes and message texts) follows
``python/ndfiltrs/_validation.py:16-211`` so that error handling is a drop-in.
"""

from __future__ import annotations

import numbers
import operator
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

MODES = ("reflect", "constant", "nearest", "mirror", "wrap")


def integer_index(value: object, message: str) -> int:
    """Return ``operator.index(value)``; booleans and non-integers raise TypeError."""
    if isinstance(value, (bool, np.bool_)):
        raise TypeError(message)
    try:
        return operator.index(value)
    except TypeError:
        raise TypeError(message) from None


@dataclass(frozen=True)
class FilterArguments:
    """Normalised arguments for the native layer."""

    array: NDArray[np.float32] | NDArray[np.float64]
    size: tuple[int, ...]
    mode: tuple[str, ...]
    cval: float
    min_count: int
    footprint: list[bool] | None = None
    origin: tuple[int, ...] = ()
    count: int = 1


def check_array(array: object) -> NDArray[np.float32] | NDArray[np.float64]:
    """Rank >= 1, float32/float64, native byte order, aligned, not masked."""
    if np.ma.isMaskedArray(array):
        raise TypeError("masked arrays are not supported")
    value = np.asarray(array)
    if value.ndim == 0:
        raise ValueError("array must have at least one dimension")
    if value.dtype.kind != "f" or value.dtype.itemsize not in (4, 8):
        raise TypeError("array dtype must be float32 or float64")
    if not value.dtype.isnative:
        raise TypeError("array dtype must be native-endian")
    if not value.flags.aligned:
        raise ValueError("array must be aligned")
    return value


def _sizes(size: object, ndim: int) -> tuple[int, ...]:
    scalar_message = "size must be an integer or a sequence of integers"
    try:
        scalar = integer_index(size, scalar_message)
    except TypeError:
        if isinstance(size, (str, bytes)) or not isinstance(size, Sequence):
            raise TypeError(scalar_message) from None
        entries = tuple(size)
        if len(entries) != ndim:
            raise ValueError("size sequence must have one entry per array dimension") from None
        values = tuple(
            integer_index(entry, "size entries must be integers, not booleans") for entry in entries
        )
    else:
        values = (scalar,) * ndim
    if any(value <= 0 for value in values):
        raise ValueError("size entries must be positive integers")
    return values


def _footprint(footprint: object, ndim: int) -> tuple[tuple[int, ...], list[bool], int]:
    if np.ma.isMaskedArray(footprint):
        raise TypeError("footprint must not be a masked array")
    entries = np.asarray(footprint)
    if entries.dtype.kind not in "biuf":
        raise TypeError("footprint must have a boolean or real numeric dtype")
    if entries.dtype.kind == "f" and not np.isfinite(entries).all():
        raise ValueError("footprint entries must be finite")
    selected = entries != 0
    if selected.ndim != ndim:
        raise ValueError("footprint must have the same rank as the array")
    if selected.size == 0:
        raise ValueError("footprint axes must be nonempty")
    count = int(np.count_nonzero(selected))
    if count == 0:
        raise ValueError("footprint must select at least one element")
    return tuple(int(n) for n in selected.shape), selected.ravel().tolist(), count


def _origins(origin: object, ndim: int) -> tuple[int, ...]:
    if np.ndim(origin) == 0:
        return (integer_index(origin, "origin must be an integer or a sequence"),) * ndim
    if isinstance(origin, np.ndarray):
        origin = origin.tolist()
    values = tuple(
        integer_index(entry, "origin entries must be integers, not booleans") for entry in origin
    )
    if len(values) != ndim:
        raise ValueError("origin sequence must have one entry per array dimension")
    return values


def window_geometry(
    size: object, footprint: object, origin: object, ndim: int
) -> tuple[tuple[int, ...], list[bool] | None, tuple[int, ...], int]:
    """Window shape, flat footprint (or None), origins and tap count.

    Pass ``size`` or ``footprint``, not both; with neither the window is 3 per
    axis. Nonzero footprint entries are taps.
    """
    if footprint is None:
        shape = _sizes(3 if size is None else size, ndim)
        flat = None
        count = int(np.prod(shape, dtype=object))
    else:
        if size is not None:
            raise ValueError("pass either size or footprint, not both")
        shape, flat, count = _footprint(footprint, ndim)
    origins = _origins(origin, ndim)
    for length, shift in zip(shape, origins):
        if not -(length // 2) <= shift <= (length - 1) // 2:
            raise ValueError("origin must satisfy -(size // 2) <= origin <= (size - 1) // 2")
    return shape, flat, origins, count


def _modes(mode: object, ndim: int) -> tuple[str, ...]:
    if isinstance(mode, str):
        values = (mode,) * ndim
    else:
        if not isinstance(mode, Sequence):
            raise TypeError("mode must be a string or a sequence of strings")
        values = tuple(mode)
        if len(values) != ndim:
            raise ValueError("mode sequence must have one entry per array dimension")
        if any(not isinstance(value, str) for value in values):
            raise TypeError("mode entries must be strings")
    if any(value not in MODES for value in values):
        raise ValueError(f"mode must use only {MODES}")
    return values


def filter_arguments(
    array: object,
    size: object,
    mode: object,
    cval: object,
    *,
    min_count: object,
    footprint: object = None,
    origin: object = 0,
) -> FilterArguments:
    """Validate public arguments; non-C-contiguous inputs are copied."""
    value = check_array(array)
    shape, flat, origins, count = window_geometry(size, footprint, origin, value.ndim)
    modes = _modes(mode, value.ndim)
    if isinstance(cval, (bool, np.bool_)) or not isinstance(cval, numbers.Real):
        raise TypeError("cval must be a non-boolean real scalar")
    cast_cval = float(value.dtype.type(cval))
    minimum = integer_index(min_count, "min_count must be an integer, not a boolean")
    native_max = np.iinfo(np.uintp).max
    volume = 1
    for length in shape:
        volume *= length
        if volume > native_max:
            raise OverflowError("window volume exceeds the native index range")
    if minimum < 1 or minimum > count:
        raise ValueError("min_count must be between 1 and the number of window elements")
    if not value.flags.c_contiguous or type(value) is not np.ndarray:
        value = np.require(value, dtype=value.dtype, requirements=("C", "A", "O", "W"))
    return FilterArguments(
        array=value,
        size=shape,
        mode=modes,
        cval=cast_cval,
        min_count=minimum,
        footprint=flat,
        origin=origins,
        count=count,
    )
