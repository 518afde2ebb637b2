"""Public median, rank and extrema filters (drop-in signatures of ndfiltrs).

Signatures, defaults and error behaviour follow ``python/ndfiltrs/_api.py``
(``median_filter`` :86-152, ``nanmedian_filter`` :155-225, ``min_filter``
:388-435, ``max_filter`` :438-485, ``nanmin_filter`` :488-551,
``nanmax_filter`` :554-617, ``rank_filter`` :945-1013, ``get_num_threads``
:1016-1028). Every method string is served by the same bit-exact GPU engines;
the strings keep their validation.
"""

from __future__ import annotations

from typing import Literal

import numpy as np
from numpy.typing import ArrayLike, NDArray

from . import _core
from ._validation import FilterArguments, filter_arguments, integer_index

BoundaryMode = Literal["reflect", "constant", "nearest", "mirror", "wrap"]

_MEDIAN_KINDS = {"direct": "median", "sortnet": "median_network", "histogram": "median_histogram"}


def _check_method(method: object, allowed: tuple[str, ...]) -> None:
    if not isinstance(method, str) or method not in allowed:
        raise ValueError(f"method must be one of {allowed}")


def _median_kind(method: object, arguments: FilterArguments) -> str:
    _check_method(method, ("direct", "sortnet", "histogram"))
    if method == "histogram" and arguments.footprint is not None:
        raise ValueError("method='histogram' requires size, not footprint")
    return _MEDIAN_KINDS[str(method)]


def _run(
    kind: str,
    arguments: FilterArguments,
    *,
    ignore_inf: bool | None = None,
    parameter: int = 0,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """One native call; ``ignore_inf=None`` selects the propagate policy."""
    return _core.window_filter(
        arguments.array,
        kind,
        arguments.size,
        arguments.footprint,
        arguments.origin,
        arguments.mode,
        arguments.cval,
        ignore_inf is not None,
        bool(ignore_inf),
        arguments.min_count,
        parameter,
    )


def median_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    method: Literal["direct", "sortnet", "histogram"] = "direct",
) -> NDArray[np.float32] | NDArray[np.float64]:
    """N-dimensional median filter on the GPU; NaN propagates.

    Parameters
    ----------
    array : array-like
        Aligned, native-endian float32 or float64 array of rank >= 1.
    size : int or tuple of int, optional
        Rectangular window; a scalar applies to every axis. Defaults to 3 per
        axis when ``footprint`` is None.
    footprint : array-like, optional
        Window mask of the array's rank; nonzero entries are taps. Mutually
        exclusive with ``size``.
    origin : int or tuple of int, optional
        SciPy window shift: tap ``k`` of an axis of length ``n`` reads
        ``center + k - (n // 2 + origin)``.
    mode : str or tuple of str, optional
        ``reflect``, ``constant``, ``nearest``, ``mirror`` or ``wrap`` per axis.
    cval : float, optional
        Sample value beyond ``constant`` boundaries, cast to the dtype.
    method : {'direct', 'sortnet', 'histogram'}, optional
        Accepted for compatibility; all give identical bits. ``histogram``
        rejects ``footprint``.

    Returns
    -------
    numpy.ndarray
        New owned C-contiguous array. A window containing NaN gives canonical
        NaN; infinities take part; an even count gives the f64 midpoint of
        the two middle values, cast once to the dtype.
    """
    arguments = filter_arguments(array, size, mode, cval, footprint=footprint, origin=origin, min_count=1)
    return _run(_median_kind(method, arguments), arguments)


def nanmedian_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    ignore_inf: bool = False,
    min_count: int = 1,
    method: Literal["direct", "sortnet", "histogram"] = "direct",
) -> NDArray[np.float32] | NDArray[np.float64]:
    """N-dimensional median filter on the GPU that omits NaN samples.

    Parameters are those of :func:`median_filter`, plus:

    ignore_inf : bool, optional
        Also omit +inf and -inf.
    min_count : int, optional
        Smallest number of retained samples for a result; fewer gives
        canonical NaN. Between 1 and the number of taps.

    ``cval`` samples follow the same omission rules as array samples.
    """
    if not isinstance(ignore_inf, (bool, np.bool_)):
        raise TypeError("ignore_inf must be a boolean")
    arguments = filter_arguments(
        array, size, mode, cval, footprint=footprint, origin=origin, min_count=min_count
    )
    return _run(_median_kind(method, arguments), arguments, ignore_inf=bool(ignore_inf))


def min_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """Smallest sample of each window, in IEEE total order, on the GPU.

    Total order puts -0.0 below +0.0, so a window holding both gives -0.0.
    A window containing NaN gives canonical NaN; infinities take part. A
    rectangular window runs the separable extrema engine. Parameters are
    those of :func:`median_filter` without ``method``.
    """
    arguments = filter_arguments(array, size, mode, cval, footprint=footprint, origin=origin, min_count=1)
    return _run("min", arguments)


def max_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """Largest sample of each window, in IEEE total order, on the GPU.

    A window holding both signed zeros gives +0.0. NaN propagates as
    canonical NaN; infinities take part. Parameters are those of
    :func:`min_filter`.
    """
    arguments = filter_arguments(array, size, mode, cval, footprint=footprint, origin=origin, min_count=1)
    return _run("max", arguments)


def nanmin_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    ignore_inf: bool = False,
    min_count: int = 1,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """:func:`min_filter` that omits NaN samples (and infinities when
    ``ignore_inf``); fewer than ``min_count`` retained samples give canonical
    NaN. ``cval`` samples follow the same rules, and boundary repeats count
    once per tap."""
    if not isinstance(ignore_inf, (bool, np.bool_)):
        raise TypeError("ignore_inf must be a boolean")
    arguments = filter_arguments(
        array, size, mode, cval, footprint=footprint, origin=origin, min_count=min_count
    )
    return _run("min", arguments, ignore_inf=bool(ignore_inf))


def nanmax_filter(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    ignore_inf: bool = False,
    min_count: int = 1,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """:func:`max_filter` that omits NaN samples; other rules as in
    :func:`nanmin_filter`."""
    if not isinstance(ignore_inf, (bool, np.bool_)):
        raise TypeError("ignore_inf must be a boolean")
    arguments = filter_arguments(
        array, size, mode, cval, footprint=footprint, origin=origin, min_count=min_count
    )
    return _run("max", arguments, ignore_inf=bool(ignore_inf))


def rank_filter(
    array: ArrayLike,
    rank: int,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    origin: int | tuple[int, ...] = 0,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    method: Literal["direct", "histogram"] = "direct",
) -> NDArray[np.float32] | NDArray[np.float64]:
    """Value of a given zero-based rank in each sorted window, on the GPU.

    ``rank`` satisfies ``-n <= rank < n`` for ``n`` taps; negative values
    count from the top. Ordering is IEEE total order (-0.0 before +0.0).
    NaN propagates. Other parameters are those of :func:`median_filter`.
    """
    position = integer_index(rank, "rank must be an integer, not a boolean")
    arguments = filter_arguments(array, size, mode, cval, footprint=footprint, origin=origin, min_count=1)
    if not -arguments.count <= position < arguments.count:
        raise ValueError("rank must satisfy -n <= rank < n for n window elements")
    _check_method(method, ("direct", "histogram"))
    if method == "histogram" and arguments.footprint is not None:
        raise ValueError("method='histogram' requires size, not footprint")
    return _run(
        "rank_histogram" if method == "histogram" else "rank",
        arguments,
        parameter=position % arguments.count,
    )


def get_num_threads() -> int:
    """Host thread count a CPU caller would use (``RAYON_NUM_THREADS`` or CPUs)."""
    return _core.get_num_threads()
