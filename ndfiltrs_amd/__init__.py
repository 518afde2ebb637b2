"""ndfiltrs_amd: NaN-aware window selection filters on MI355X (gfx950).

Drop-in for ndfiltrs' ``median_filter``, ``nanmedian_filter``,
``rank_filter``, ``percentile_filter``, ``nanpercentile_filter``,
``min_filter``, ``max_filter``, ``nanmin_filter``, ``nanmax_filter``, the
grey-scale morphology built on them, the mesh-background grid median
filter, ``get_num_threads`` and ``fixpix(method="median")``; results are bit-exact with the reference CPU
semantics. Computation runs in hand-written HIP kernels in ``libndfgpu.so``;
there is no CPU fallback.
"""

from . import _core
from ._background import mesh_median_filter
from ._api import (
    get_num_threads,
    max_filter,
    median_filter,
    min_filter,
    nanmax_filter,
    nanmedian_filter,
    nanmin_filter,
    rank_filter,
)
from ._morphology import (
    black_tophat,
    grey_closing,
    grey_dilation,
    grey_opening,
    morphological_gradient,
    morphological_laplace,
    white_tophat,
)
from ._repair import fixpix
from ._robust import nanpercentile_filter, percentile_filter

__all__ = [
    "_core",
    "black_tophat",
    "fixpix",
    "get_num_threads",
    "grey_closing",
    "grey_dilation",
    "grey_opening",
    "max_filter",
    "median_filter",
    "mesh_median_filter",
    "min_filter",
    "morphological_gradient",
    "morphological_laplace",
    "nanmax_filter",
    "nanmedian_filter",
    "nanmin_filter",
    "nanpercentile_filter",
    "percentile_filter",
    "rank_filter",
    "white_tophat",
]
__version__ = "0.2.0"
