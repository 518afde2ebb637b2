"""Local percentiles on the GPU (drop-in signatures of ndfiltrs).

``percentile_filter`` and ``nanpercentile_filter`` keep the reference's
signatures, defaults, checks and messages (``python/ndfiltrs/_robust.py:21-
183``). They run ``_core.robust_filter(kind="percentile")``: each window's
retained samples are ranked in IEEE total order and the samples at
``floor(r)`` and ``floor(r) + 1`` are interpolated in float64 at
``r = (n - 1) * percentile / 100``, with the reference's overflow-safe
rules (``src/robust.rs:320-367``), then cast once to the dtype.
"""

from __future__ import annotations

import math
import numbers

import numpy as np
from numpy.typing import ArrayLike, NDArray

from . import _core
from ._api import BoundaryMode
from ._validation import filter_arguments

FloatArray = NDArray[np.float32] | NDArray[np.float64]


def _finite_real(value: object, name: str) -> float:
    """A real, non-boolean, finite scalar."""
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, numbers.Real):
        raise TypeError(f"{name} must be a non-boolean real number")
    number = float(value)
    if not math.isfinite(number):
        raise ValueError(f"{name} must be finite")
    return number


def _percentile(value: object) -> float:
    quantile = _finite_real(value, "percentile")
    if not 0 <= quantile <= 100:
        raise ValueError("percentile must lie between 0 and 100")
    return quantile


def _robust(arguments, quantile: float, omit_nan: bool, ignore_inf: bool) -> FloatArray:
    # The clipping arguments of the native signature are unused by the
    # percentile; they carry the reference's defaults.
    return _core.robust_filter(
        arguments.array,
        arguments.size,
        arguments.footprint,
        arguments.origin,
        arguments.mode,
        arguments.cval,
        kind="percentile",
        parameter=quantile,
        omit_nan=omit_nan,
        ignore_inf=ignore_inf,
        min_count=arguments.min_count,
        sigma_lower=3.0,
        sigma_upper=3.0,
        maxiters=5,
        center="median",
        spread="std",
    )


def percentile_filter(
    array: ArrayLike,
    percentile: float,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
) -> FloatArray:
    """Linear-interpolated percentile of every window, on the GPU.

    ``percentile`` lies in [0, 100]; 0 and 100 give the window's minimum and
    maximum in total order (-0.0 below +0.0). A window containing NaN gives
    canonical NaN; infinities take part, and interpolating between -inf and
    +inf gives NaN. Other parameters are those of ``median_filter`` without
    ``method``. Returns a new owned C-contiguous array of the input's shape
    and dtype.
    """
    quantile = _percentile(percentile)
    arguments = filter_arguments(array, size, mode, cval, footprint=footprint, origin=origin, min_count=1)
    return _robust(arguments, quantile, False, False)


def nanpercentile_filter(
    array: ArrayLike,
    percentile: float,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    ignore_inf: bool = False,
    min_count: int = 1,
) -> FloatArray:
    """:func:`percentile_filter` over the retained samples only: NaN (and,
    with ``ignore_inf``, infinities) are omitted, the rank uses the retained
    count, and fewer than ``min_count`` retained samples give canonical NaN.
    ``cval`` samples follow the same rules; boundary repeats count once per
    tap."""
    quantile = _percentile(percentile)
    if not isinstance(ignore_inf, (bool, np.bool_)):
        raise TypeError("ignore_inf must be a boolean")
    arguments = filter_arguments(
        array, size, mode, cval, footprint=footprint, origin=origin, min_count=min_count
    )
    return _robust(arguments, quantile, True, bool(ignore_inf))
