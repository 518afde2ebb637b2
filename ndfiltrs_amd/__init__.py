"""ndfiltrs_amd: NaN-aware median and rank window filters on MI355X (gfx950).

Drop-in for ndfiltrs' ``median_filter``, ``nanmedian_filter``,
``rank_filter``, ``get_num_threads`` and ``fixpix(method="median")``; results are bit-exact with the
reference CPU semantics. Computation runs in hand-written HIP kernels in
``libndfgpu.so``; there is no CPU fallback.
"""

from . import _core
from ._api import get_num_threads, median_filter, nanmedian_filter, rank_filter
from ._repair import fixpix

__all__ = [
    "_core",
    "fixpix",
    "get_num_threads",
    "median_filter",
    "nanmedian_filter",
    "rank_filter",
]
__version__ = "0.1.0"
