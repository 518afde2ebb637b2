"""Mask-directed pixel repair: ``fixpix`` with ``method="median"`` on the GPU.

Signature, defaults, option checks, messages and ownership rules follow
``python/ndfiltrs/_repair.py:22-250``. The median method evaluates only the
masked windows, over a snapshot of the data in which every masked pixel is
NaN (``src/repair.rs:107-126``, ``src/filter_nd.rs:174-215``), so repaired
values never feed another repair. The ``linear`` and ``mean`` methods are
validated like the reference but are not implemented in ndfiltrs_amd: they
raise ``NotImplementedError``.
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Literal, SupportsIndex

import numpy as np
from numpy.typing import ArrayLike, NDArray

from . import _core
from ._validation import check_array, filter_arguments

BoundaryMode = Literal["reflect", "constant", "nearest", "mirror", "wrap"]


def fixpix(
    array: ArrayLike,
    mask: ArrayLike | None = None,
    *,
    method: Literal["linear", "median", "mean"] | None = None,
    priority: Sequence[SupportsIndex] | None = None,
    size: int | tuple[int, ...] | None = None,
    footprint: ArrayLike | None = None,
    kernel: ArrayLike | None = None,
    origin: int | tuple[int, ...] | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] | None = None,
    cval: float | None = None,
    ignore_inf: bool | None = None,
    min_count: int | None = None,
    copy: bool = True,
) -> NDArray[np.float32] | NDArray[np.float64]:
    """Repair masked pixels by the median of their unmasked neighbours.

    Parameters
    ----------
    array : array-like
        Aligned, native-endian float32 or float64 data of rank >= 1; never
        modified.
    mask : array-like or None
        Same shape as ``array``; nonzero entries mark bad pixels. ``None``
        means no bad pixels. NaN and infinities are not masked automatically.
    method : {'linear', 'median', 'mean'} or None
        ``None`` selects ``'mean'`` when ``kernel`` is given, else
        ``'linear'``. Only ``'median'`` is implemented (on the GPU).
    size, footprint, origin, mode : optional
        Median window as in ``nanmedian_filter``; defaults 3 per axis, origin
        0 and ``'reflect'``.
    cval : float or None
        Constant-boundary sample; ``None`` selects NaN, so exterior samples
        are omitted under ``'constant'``.
    ignore_inf : bool or None
        Also omit infinities (``None``: ``False``). NaN and originally masked
        samples are always omitted.
    min_count : int or None
        Fewest retained samples for a result (``None``: 1); fewer gives NaN.
    copy : bool
        ``False`` may return the input itself when no pixel is masked.

    Returns
    -------
    numpy.ndarray
        Same shape and dtype; unmasked pixels unchanged; a new owned
        C-contiguous array unless ``copy=False`` returned the input.
    """
    value = check_array(array)
    if method is None:
        method = "linear" if kernel is None else "mean"
    if not isinstance(method, str) or method not in ("linear", "median", "mean"):
        raise ValueError("method must be 'linear', 'median', or 'mean'")
    if not isinstance(copy, (bool, np.bool_)):
        raise TypeError("copy must be a boolean")
    options = {
        "priority": (priority, ("linear",)),
        "size": (size, ("median", "mean")),
        "footprint": (footprint, ("median", "mean")),
        "kernel": (kernel, ("mean",)),
        "origin": (origin, ("median", "mean")),
        "mode": (mode, ("median", "mean")),
        "cval": (cval, ("median", "mean")),
        "ignore_inf": (ignore_inf, ("median", "mean")),
        "min_count": (min_count, ("median", "mean")),
    }
    for name, (option, methods) in options.items():
        if option is not None and method not in methods:
            allowed = " or ".join(repr(entry) for entry in methods)
            raise ValueError(f"{name} requires method={allowed}")
    if ignore_inf is not None and not isinstance(ignore_inf, (bool, np.bool_)):
        raise TypeError("ignore_inf must be a boolean")
    if method != "median":
        raise NotImplementedError(
            f"fixpix(method={method!r}) is not implemented in ndfiltrs_amd; method='median' is"
        )
    arguments = filter_arguments(
        value,
        size,
        "reflect" if mode is None else mode,
        np.nan if cval is None else cval,
        footprint=footprint,
        origin=0 if origin is None else origin,
        min_count=1 if min_count is None else min_count,
    )
    bad = None
    if mask is not None:
        if np.ma.isMaskedArray(mask):
            raise TypeError("masked arrays are not supported as masks")
        bad = np.asarray(mask)
        if bad.shape != value.shape:
            raise ValueError("mask must have the same shape as array")
        if bad.dtype.kind not in "buifc":
            raise TypeError("mask must have a boolean or numeric dtype")
    if bad is None or not np.any(bad):
        return value.copy(order="C") if copy else value
    if bad.dtype == np.bool_ and bad.flags.c_contiguous:
        bad = bad.view(np.uint8)  # one byte per element already, 0 or 1
    else:
        bad = np.ascontiguousarray(bad != 0, dtype=np.bool_).view(np.uint8)
    return _core.fixpix_median(
        arguments.array,
        bad,
        arguments.size,
        arguments.footprint,
        arguments.origin,
        arguments.mode,
        arguments.cval,
        bool(ignore_inf),
        arguments.min_count,
    )
