"""Grey-scale morphology with flat structuring elements, on the GPU.

Signatures and defaults follow ``python/ndfiltrs/_morphology.py:17-343``.
Erosion is :func:`min_filter`; dilation is :func:`max_filter` over the
structuring element mirrored through its centre (SciPy's
``grey_dilation`` convention); the compound operators combine those two in
the input dtype with NumPy arithmetic, as the reference does.
"""

from __future__ import annotations

import numpy as np
from numpy.typing import ArrayLike, NDArray

from ._api import BoundaryMode, max_filter, min_filter
from ._validation import check_array, window_geometry

FloatArray = NDArray[np.float32] | NDArray[np.float64]


def grey_dilation(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Maximum over the mirrored structuring element.

    The footprint is reversed along every axis and each origin becomes
    ``-origin`` (minus one more on an even axis), so the result matches
    ``scipy.ndimage.grey_dilation`` for flat elements. NaN propagates.
    """
    value = check_array(array)
    shape, _, origins, _ = window_geometry(size, footprint, origin, value.ndim)
    if footprint is not None:
        footprint = np.asarray(footprint)[(slice(None, None, -1),) * value.ndim]
    mirrored = tuple(-shift - (1 - length % 2) for length, shift in zip(shape, origins))
    return max_filter(value, size, footprint=footprint, mode=mode, cval=cval, origin=mirrored)


def grey_opening(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Erosion then dilation: removes bright features smaller than the
    element. The boundary mode applies to both passes."""
    eroded = min_filter(array, size, footprint=footprint, mode=mode, cval=cval, origin=origin)
    return grey_dilation(eroded, size, footprint=footprint, mode=mode, cval=cval, origin=origin)


def grey_closing(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Dilation then erosion: removes dark features smaller than the element."""
    dilated = grey_dilation(array, size, footprint=footprint, mode=mode, cval=cval, origin=origin)
    return min_filter(dilated, size, footprint=footprint, mode=mode, cval=cval, origin=origin)


def white_tophat(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Input minus its grey opening (bright features smaller than the
    element); opposite infinities give NaN."""
    value = check_array(array)
    return value - grey_opening(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin)


def black_tophat(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Grey closing minus the input (dark features smaller than the element)."""
    value = check_array(array)
    return grey_closing(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin) - value


def morphological_gradient(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Dilation minus erosion: the local max - min range."""
    value = check_array(array)
    dilated = grey_dilation(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin)
    return dilated - min_filter(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin)


def morphological_laplace(
    array: ArrayLike,
    size: int | tuple[int, ...] | None = None,
    *,
    footprint: ArrayLike | None = None,
    mode: BoundaryMode | tuple[BoundaryMode, ...] = "reflect",
    cval: float = 0.0,
    origin: int | tuple[int, ...] = 0,
) -> FloatArray:
    """Dilation plus erosion minus twice the input, in the input dtype."""
    value = check_array(array)
    dilated = grey_dilation(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin)
    eroded = min_filter(value, size, footprint=footprint, mode=mode, cval=cval, origin=origin)
    return dilated + eroded - 2 * value
