"""Median filter of mesh-background grids on the GPU.

The reference's ``mesh_background`` (``python/ndfiltrs/_background.py:69-
268``) estimates one background and RMS value per mesh, fills invalid
meshes, median-filters both grids and interpolates them back to full
resolution. The grid median filter (``MeshBackground::filter``,
``src/background.rs:797-855``) is the stage on the median path; it is built
here, with the reference's ``filter_size``, ``filter_edge``,
``filter_threshold`` and ``threshold_kind`` arguments and checks. The mesh
statistics, fill and spline stages are not part of this package.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from typing import Literal

import numpy as np
from numpy.typing import ArrayLike, NDArray

from . import _core, _lib
from ._robust import _finite_real
from ._validation import integer_index

EDGES = ("shrink", "truncate")
THRESHOLD_KINDS = ("deviation", "above")


def _filter_sizes(value: object, ndim: int) -> tuple[int, ...]:
    """Positive odd sizes, a scalar or one per grid axis."""
    message = "filter_size must be an integer or a sequence of integers"
    if isinstance(value, np.ndarray) and value.ndim == 1:
        value = value.tolist()
    if isinstance(value, Sequence) and not isinstance(value, (str, bytes)):
        sizes = tuple(integer_index(entry, message) for entry in value)
        if len(sizes) != ndim:
            raise ValueError("filter_size must have one entry per array dimension")
    else:
        sizes = (integer_index(value, message),) * ndim
    if any(size <= 0 for size in sizes):
        raise ValueError("filter_size entries must be positive")
    if any(size % 2 == 0 for size in sizes):
        raise ValueError("filter_size entries must be odd")
    return sizes


def _choice(value: object, name: str, options: tuple[str, ...]) -> str:
    if not isinstance(value, str) or value not in options:
        raise ValueError(f"{name} must be one of {options}")
    return value


def mesh_median_filter(
    background: ArrayLike,
    rms: ArrayLike,
    filter_size: int | Sequence[int] = 3,
    *,
    filter_edge: Literal["shrink", "truncate"] = "shrink",
    filter_threshold: float | None = None,
    threshold_kind: Literal["deviation", "above"] = "deviation",
) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
    """Median-filter a mesh background grid and its RMS grid on the GPU.

    Parameters
    ----------
    background, rms : array-like
        Grids of one value per mesh, the same shape, cast to float64.
    filter_size : int or sequence of int, optional
        Odd window size in meshes, scalar or one per axis. Default 3; 1 on
        every axis returns copies.
    filter_edge : {'shrink', 'truncate'}, optional
        'shrink' (SExtractor) narrows the window at the grid edge so it stays
        centred; 'truncate' (photutils) drops window positions outside the
        grid.
    filter_threshold : float or None, optional
        None replaces every mesh. Otherwise ``threshold_kind`` decides.
    threshold_kind : {'deviation', 'above'}, optional
        'deviation' replaces meshes whose median differs from their value by
        at least ``filter_threshold``; 'above' replaces meshes whose value is
        above it. A replaced mesh takes the RMS median too.

    Returns
    -------
    (numpy.ndarray, numpy.ndarray)
        New float64 background and RMS grids. A window holding NaN gives NaN;
        an even count averages the two middle values.
    """
    grid = np.ascontiguousarray(background, dtype=np.float64)
    spread = np.ascontiguousarray(rms, dtype=np.float64)
    if grid.shape != spread.shape:
        raise ValueError("background and rms grids must have the same shape")
    if grid.ndim == 0:
        raise ValueError("mesh grids must have at least one dimension")
    sizes = _filter_sizes(filter_size, grid.ndim)
    edge = _choice(filter_edge, "filter_edge", EDGES)
    kind = _choice(threshold_kind, "threshold_kind", THRESHOLD_KINDS)
    threshold = None if filter_threshold is None else _finite_real(filter_threshold, "filter_threshold")
    out_grid = np.empty_like(grid)
    out_spread = np.empty_like(spread)
    ndim = grid.ndim
    err = ctypes.create_string_buffer(256)
    status = _lib.lib.ndf_mesh_median_filter(
        grid.ctypes.data,
        spread.ctypes.data,
        out_grid.ctypes.data,
        out_spread.ctypes.data,
        ndim,
        (ctypes.c_size_t * ndim)(*grid.shape),
        (ctypes.c_size_t * ndim)(*sizes),
        EDGES.index(edge),
        0 if threshold is None else 1 + THRESHOLD_KINDS.index(kind),
        0.0 if threshold is None else threshold,
        err,
        len(err),
    )
    if status != _lib.OK:
        _core._raise(status, err.value.decode())
    return out_grid, out_spread
