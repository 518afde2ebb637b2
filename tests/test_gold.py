"""The reference's gold scene (tests/data/gold_v1 of the reference) pins the
oracle: the committed subset is the reference's data bit for bit, the exact
controls reconstruct (tests/test_gold_image.py:127-140 of the reference), and
median repair and 3x3 filters over the real defect kinds equal an
independent window-by-window construction (:87-124, :143-170)."""

from __future__ import annotations

import hashlib
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle
from tests import _cases

HERE = os.path.dirname(os.path.abspath(__file__))


def load_gold() -> dict[str, np.ndarray]:
    with np.load(os.path.join(HERE, "golden", "gold_v1.npz"), allow_pickle=False) as archive:
        return {name: archive[name] for name in archive.files}


GOLD = load_gold()
with open(os.path.join(HERE, "golden", "gold_v1.json")) as _handle:
    RECORD = json.load(_handle)


def median_of(values: list[float]) -> float:
    """Median by the midpoint contract (docs/semantics.md:74-85), NaN when empty."""
    if not values:
        return math.nan
    ordered = sorted(values, key=lambda v: (v, math.copysign(1.0, v)))
    n = len(ordered)
    return ordered[n // 2] if n % 2 else _cases._mid(ordered[n // 2 - 1], ordered[n // 2])


def repair_reference(array: np.ndarray, mask: np.ndarray, min_count: int) -> np.ndarray:
    """fixpix(method='median', size=3, mode='constant', ignore_inf=True): each
    masked pixel is the median of the unmasked finite samples of its 3x3
    window clipped to the array (constant mode with cval NaN omits the
    exterior), NaN below min_count; other pixels unchanged."""
    out = array.copy()
    for y, x in np.argwhere(mask):
        rows = slice(max(0, y - 1), min(array.shape[0], y + 2))
        cols = slice(max(0, x - 1), min(array.shape[1], x + 2))
        patch = array[rows, cols]
        usable = (~mask[rows, cols]) & np.isfinite(patch)
        values = [float(v) for v in patch[usable]]
        out[y, x] = array.dtype.type(median_of(values)) if len(values) >= min_count else array.dtype.type(np.nan)
    return out


def test_subset_is_the_reference_data() -> None:
    for name, digest in RECORD["sha256"].items():
        assert hashlib.sha256(GOLD[name].tobytes(order="C")).hexdigest() == digest, name
    assert int(GOLD["mask_sparse"].sum()) == 81 and int(GOLD["mask_stress"].sum()) == 618


@pytest.mark.parametrize("name", ["control_flat", "control_ramp"])
def test_controls_reconstruct_exactly(name: str) -> None:
    truth = GOLD[name]
    mask = GOLD["control_mask"]
    corrupted = truth.copy()
    corrupted[mask] = np.nan
    np.testing.assert_array_equal(oracle.fixpix_median(corrupted, mask, (3, 3)), truth)
    np.testing.assert_array_equal(oracle.window_filter(truth, (3, 3))[1:-1, 1:-1], truth[1:-1, 1:-1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ["sparse", "stress"])
def test_repair_matches_window_construction(dtype, case: str) -> None:
    array = GOLD[f"observed_{case}"].astype(dtype)
    mask = GOLD[f"mask_{case}"]
    got = oracle.fixpix_median(array, mask, (3, 3), mode="constant", ignore_inf=True, min_count=3)
    want = repair_reference(array, mask, 3)
    np.testing.assert_array_equal(_cases.bits_of(got), _cases.bits_of(want))


@pytest.mark.parametrize("index", range(3))
def test_filter_windows_match_brute_force(index: int) -> None:
    r0, r1, c0, c1 = RECORD["clean_windows"][index]
    for array, omit in ((GOLD[f"clean_window_{index}"], False), (GOLD["observed_stress"][r0:r1, c0:c1], True)):
        array = np.ascontiguousarray(array)
        got = oracle.window_filter(array, (3, 3), omit_nan=omit)
        want = _cases.brute_force(array, (3, 3), origin=(0, 0), mode=("reflect", "reflect"), cval=0.0,
                                  omit_nan=omit, ignore_inf=False, min_count=1)
        np.testing.assert_array_equal(_cases.bits_of(got), _cases.bits_of(want))
