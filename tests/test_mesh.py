"""Mesh-background grid median filter (MeshBackground::filter,
src/background.rs:797-855): the oracle is pinned by the reference's literal
test (src/background.rs:1366-1388) and by an independent NumPy brute force,
which each of four planted mistakes makes fail; the GPU kernel gives the
oracle's bits. The brute force and the literals run on the CPU; the
GPU comparisons are marked gpu."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import oracle

LITERALS = [
    # (edge, threshold kind, threshold, expected) on the grid [1, 100, 3, 4, 5], filter 3.
    ("shrink", "all", 0.0, [1.0, 3.0, 4.0, 4.0, 5.0]),
    ("truncate", "all", 0.0, [50.5, 3.0, 4.0, 4.0, 4.5]),
    ("shrink", "deviation", 10.0, [1.0, 3.0, 3.0, 4.0, 5.0]),
    ("truncate", "above", 50.0, [1.0, 3.0, 3.0, 4.0, 5.0]),
]
GRID = np.array([1.0, 100.0, 3.0, 4.0, 5.0])

MUTATION: str | None = None


def brute(background, rms, size, edge, kind, threshold):
    """Window-by-window restatement with NumPy's median."""
    out_b, out_r = background.copy(), rms.copy()
    if all(s == 1 for s in size):
        return out_b, out_r
    for index in np.ndindex(background.shape):
        slices = []
        for k, n, s in zip(index, background.shape, size):
            half = s // 2
            if MUTATION == "shrink_as_truncate" or edge == "truncate":
                lo, hi = max(k - half, 0), min(k + half, n - 1)
            else:
                half = min(half, k, n - 1 - k)
                lo, hi = k - half, k + half
            slices.append(slice(lo, hi + 1))
        window = background[tuple(slices)].ravel()
        median = np.nan if np.isnan(window).any() and MUTATION != "nan_skipped" else np.median(window[~np.isnan(window)])
        if MUTATION == "upper_middle":
            ordered = np.sort(window)
            median = ordered[len(ordered) // 2]
        replace = True
        if kind == "deviation":
            replace = abs(median - background[index]) >= threshold
            if MUTATION == "strict_deviation":
                replace = abs(median - background[index]) > threshold
        elif kind == "above":
            replace = background[index] > threshold
        if replace:
            out_b[index] = median
            out_r[index] = np.median(rms[tuple(slices)].ravel()) if not np.isnan(rms[tuple(slices)]).any() else np.nan
    return out_b, out_r


@pytest.mark.parametrize("case", LITERALS, ids=[f"{c[0]}-{c[1]}" for c in LITERALS])
def test_oracle_literals(case) -> None:
    edge, kind, threshold, expected = case
    got, _ = oracle.mesh_filter(GRID, GRID, (3,), edge=edge, threshold_kind=kind, threshold=threshold)
    np.testing.assert_array_equal(got, expected)
    got, _ = brute(GRID, GRID, (3,), edge, kind, threshold)
    np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("mutation", ["shrink_as_truncate", "upper_middle", "strict_deviation", "nan_skipped"])
def test_planted_mistakes_fail_a_literal_or_a_nan_case(mutation: str) -> None:
    global MUTATION
    MUTATION = mutation
    try:
        failures = 0
        for edge, kind, threshold, expected in LITERALS:
            got, _ = brute(GRID, GRID, (3,), edge, kind, threshold)
            failures += not np.array_equal(got, expected)
        # A NaN mesh poisons every window holding it (median_of, src/background.rs:941-943).
        nan_grid = np.array([1.0, np.nan, 3.0, 4.0])
        got, _ = brute(nan_grid, nan_grid, (3,), "shrink", "all", 0.0)
        failures += not np.array_equal(got, [1.0, np.nan, np.nan, 4.0], equal_nan=True)
        # A deviation exactly at the threshold is replaced.
        tie = np.array([1.0, 2.0, 3.0, 10.0, 5.0])
        got, _ = brute(tie, tie, (3,), "shrink", "deviation", 5.0)
        failures += not np.array_equal(got, [1.0, 2.0, 3.0, 5.0, 5.0])
        assert failures > 0, mutation
    finally:
        MUTATION = None


def grids(seed: int, shape):
    rng = np.random.default_rng(seed)
    background = np.round(rng.normal(100.0, 5.0, size=shape), 1)
    rms = np.round(rng.uniform(1.0, 3.0, size=shape), 2)
    background[rng.random(shape) < 0.05] = np.nan
    return background, rms


CASES = list(itertools.product([(7,), (9, 6), (4, 5, 3)], ["shrink", "truncate"],
                               [("all", 0.0), ("deviation", 2.0), ("above", 101.0)]))


@pytest.mark.parametrize(("shape", "edge", "threshold"), CASES)
def test_oracle_equals_brute_force(shape, edge, threshold) -> None:
    background, rms = grids(len(shape) + 3, shape)
    for size in ((3,) * len(shape), (5,) + (1,) * (len(shape) - 1), tuple(1 + 2 * (i % 2) for i in range(len(shape)))):
        want = brute(background, rms, size, edge, *threshold)
        got = oracle.mesh_filter(background, rms, size, edge=edge, threshold_kind=threshold[0],
                                 threshold=threshold[1])
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)


def test_python_validation() -> None:
    import ndfiltrs_amd as nf

    with pytest.raises(ValueError, match="odd"):
        nf.mesh_median_filter(np.ones((4, 4)), np.ones((4, 4)), 2)
    with pytest.raises(ValueError, match="filter_edge"):
        nf.mesh_median_filter(np.ones(4), np.ones(4), filter_edge="clip")
    with pytest.raises(ValueError, match="filter_threshold must be finite"):
        nf.mesh_median_filter(np.ones(4), np.ones(4), filter_threshold=np.nan)
    with pytest.raises(TypeError):
        nf.mesh_median_filter(np.ones(4), np.ones(4), filter_threshold="1")


@pytest.mark.gpu
@pytest.mark.parametrize(("shape", "edge", "threshold"), CASES)
def test_gpu_equals_oracle(shape, edge, threshold) -> None:
    import ndfiltrs_amd as nf

    background, rms = grids(len(shape) + 7, shape)
    for size in ((3,) * len(shape), (5,) + (1,) * (len(shape) - 1), (1,) * len(shape)):
        want = oracle.mesh_filter(background, rms, size, edge=edge, threshold_kind=threshold[0],
                                  threshold=threshold[1])
        got = nf.mesh_median_filter(background, rms, size, filter_edge=edge,
                                    filter_threshold=None if threshold[0] == "all" else threshold[1],
                                    threshold_kind="deviation" if threshold[0] == "all" else threshold[0])
        for a, b in zip(got, want):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
def test_gpu_literals_and_overflowing_midpoint() -> None:
    import ndfiltrs_amd as nf

    for edge, kind, threshold, expected in LITERALS:
        got, _ = nf.mesh_median_filter(GRID, GRID, 3, filter_edge=edge,
                                       filter_threshold=None if kind == "all" else threshold,
                                       threshold_kind="deviation" if kind == "all" else kind)
        np.testing.assert_array_equal(got, expected)
    big = np.finfo(np.float64).max
    grid = np.array([big, big, -big, big])
    got, _ = nf.mesh_median_filter(grid, grid, 3, filter_edge="truncate")
    want, _ = oracle.mesh_filter(grid, grid, (3,), edge="truncate")
    np.testing.assert_array_equal(got, want)
    assert np.isfinite(got).all()
