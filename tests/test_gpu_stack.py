"""GPU parity of the stack engine (box windows, sorted slices merged per
output) against the CPU oracle, bit for bit.

The engine is forced on every box shape it has networks for, with inputs
full of ties, signed zeros, infinities and NaN payloads, every boundary mode,
origin and retention policy, and shapes that leave partial column tiles, runs
shorter than the window and single rows. Outputs are poisoned before every
launch (conftest), so a skipped write fails the comparison.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases

pytestmark = pytest.mark.gpu

# (window size, dtypes the engine compiles it for)
BOXES = [
    ((3, 3), (np.float32, np.float64)),
    ((5, 5), (np.float32, np.float64)),
    ((7, 7), (np.float32, np.float64)),
    ((9, 9), (np.float32,)),
    ((3, 3, 3), (np.float32, np.float64)),
    ((3, 5, 5), (np.float32,)),
    ((3, 7, 7), (np.float32,)),
]


def plan_for(values: np.ndarray, size, *, origin=None, mode="reflect", cval=0.0, kind="median", omit=True,
             ignore_inf=False, min_count=1, parameter=0, engine="stack") -> _core.Plan:
    ndim = values.ndim
    origin = (0,) * ndim if origin is None else tuple(origin)
    modes = (mode,) * ndim if isinstance(mode, str) else tuple(mode)
    plan = _core.Plan(values.dtype, values.shape, kind, list(size), None, origin, modes, cval, omit, ignore_inf,
                      min_count, parameter)
    plan.set_engine(engine)
    return plan


def check(actual: np.ndarray, expected: np.ndarray, context: str) -> None:
    _cases.assert_no_poison(actual, context)
    a, e = _cases.bits_of(actual), _cases.bits_of(expected)
    if not np.array_equal(a, e):
        bad = np.argwhere(a != e)
        index = tuple(bad[0])
        raise AssertionError(f"{context}: {len(bad)} mismatches, first at {index}: got {actual[index]!r} "
                             f"expected {expected[index]!r}")


def shapes_for(size) -> list[tuple[int, ...]]:
    if len(size) == 2:
        # partial column tiles, a single row, fewer rows than the window
        return [(45, 150), (1, 70), (3, 64), (70, 129)]
    return [(6, 20, 71), (1, 9, 40), (2, 3, 65), (11, 13, 17)]


@pytest.mark.parametrize(("size", "dtypes"), BOXES, ids=["x".join(map(str, s)) for s, _ in BOXES])
def test_stack_matches_oracle_on_special_values(size, dtypes) -> None:
    for dtype, shape in itertools.product(dtypes, shapes_for(size)):
        values = _cases.special_sample(shape, dtype, seed=sum(shape) + len(size))
        count = int(np.prod(size))
        origin = tuple((s - 1) // 2 if i % 2 else -(s // 2) for i, s in enumerate(size))
        for mode, policy in itertools.product(_cases.MODES, _cases.POLICIES):
            function, extra = _cases.policy_kwargs(policy, count)
            omit = function == "nanmedian_filter"
            ignore_inf = bool(extra.get("ignore_inf", False))
            min_count = int(extra.get("min_count", 1))
            kwargs = {"size": size, "origin": origin, "mode": mode, "cval": 0.5, **extra}
            expected = _cases.oracle_call(function, values, kwargs)
            plan = plan_for(values, size, origin=origin, mode=mode, cval=0.5, omit=omit, ignore_inf=ignore_inf,
                            min_count=min_count if omit else 1)
            check(plan.apply(values), expected, f"{dtype.__name__} {shape} {mode} {policy}")


@pytest.mark.parametrize(("size", "dtypes"), BOXES, ids=["x".join(map(str, s)) for s, _ in BOXES])
def test_stack_ranks_and_percentiles(size, dtypes) -> None:
    count = int(np.prod(size))
    for dtype in dtypes:
        shape = shapes_for(size)[0]
        values = _cases.special_sample(shape, dtype, seed=11 + count)
        for rank in sorted({0, 1, count // 3, count // 2, count - 2, count - 1}):
            expected = oracle.window_filter(values, size, mode="wrap", statistic="rank", rank=rank, omit_nan=True,
                                            min_count=2)
            plan = plan_for(values, size, mode="wrap", kind="rank", omit=True, min_count=2, parameter=rank)
            check(plan.apply(values), expected, f"{dtype.__name__} omit rank {rank}")
            expected = oracle.window_filter(values, size, mode="mirror", statistic="rank", rank=rank)
            plan = plan_for(values, size, mode="mirror", kind="rank", omit=False, parameter=rank)
            check(plan.apply(values), expected, f"{dtype.__name__} propagate rank {rank}")
        for q in (0.0, 10.0, 37.5, 50.0, 84.0, 100.0):
            expected = oracle.window_filter(values, size, mode="nearest", statistic="percentile", q=q,
                                            omit_nan=True, ignore_inf=True)
            plan = plan_for(values, size, mode="nearest", kind="percentile", omit=True, ignore_inf=True, parameter=q)
            check(plan.apply(values), expected, f"{dtype.__name__} percentile {q}")


@pytest.mark.parametrize("nan_fraction", [0.0, 0.3, 0.6, 0.95])
def test_stack_median_with_many_omitted_samples(nan_fraction: float) -> None:
    """Windows with many omitted samples read ranks far below W/2, past the
    short select of the top ranks (the wave takes the full-half select)."""
    rng = np.random.default_rng(int(nan_fraction * 100))
    for size, dtype in (((5, 5), np.float64), ((3, 7, 7), np.float32), ((9, 9), np.float32)):
        shape = (70, 133) if len(size) == 2 else (5, 30, 70)
        values = rng.normal(0.0, 1.0, size=shape).astype(dtype)
        values[rng.random(shape) < nan_fraction] = np.nan
        # NaN blocks: whole windows without a retained sample
        values[..., 10:30, 40:60] = np.nan
        for min_count in (1, 3):
            expected = oracle.window_filter(values, size, mode="reflect", statistic="median", omit_nan=True,
                                            min_count=min_count)
            plan = plan_for(values, size, mode="reflect", min_count=min_count)
            check(plan.apply(values), expected, f"{size} nan {nan_fraction} min_count {min_count}")


def test_stack_equals_network_and_slide_on_sky() -> None:
    """The auto engine for a 5x5 box is the stack engine; its bits equal the
    network and slide engines' on a sky image with 1 % NaN."""
    rng = np.random.default_rng(4)
    values = rng.normal(1000.0, 10.0, size=(300, 333)).astype(np.float32)
    values[rng.random(values.shape) < 0.01] = np.nan
    auto = plan_for(values, (5, 5), engine="auto")
    assert '"engine": "stack"' in auto.describe()
    first = auto.apply(values)
    for engine in ("network", "slide", "generic"):
        check(plan_for(values, (5, 5), engine=engine).apply(values), first, engine)
    expected = oracle.window_filter(values, (5, 5), mode="reflect", statistic="median", omit_nan=True)
    check(first, expected, "oracle")


def test_stack_runs_split_columns() -> None:
    """Small images split each column into runs; a long 3-D stack keeps
    whole columns. Both equal the oracle."""
    rng = np.random.default_rng(9)
    for shape in ((40, 12, 70), (130, 4, 64)):
        values = rng.poisson(200.0, size=shape).astype(np.float32)
        values[rng.random(shape) < 0.05] = np.nan
        values[rng.random(shape) < 0.001] = np.inf
        plan = plan_for(values, (3, 7, 7), mode="wrap", ignore_inf=True)
        description = plan.describe()
        expected = oracle.window_filter(values, (3, 7, 7), mode="wrap", statistic="median", omit_nan=True,
                                        ignore_inf=True)
        check(plan.apply(values), expected, f"{shape} {description}")


def test_stack_rejects_other_windows() -> None:
    values = np.zeros((20, 20), dtype=np.float32)
    for size in ((4, 4), (3, 5), (11, 11)):
        plan = _core.Plan(values.dtype, values.shape, "median", list(size), None, (0, 0), ("reflect",) * 2, 0.0,
                          True, False, 1, 0)
        with pytest.raises(ValueError, match="stack engine needs"):
            plan.set_engine("stack")
    f64 = np.zeros((6, 20, 20), dtype=np.float64)
    plan = _core.Plan(f64.dtype, f64.shape, "median", [3, 7, 7], None, (0, 0, 0), ("reflect",) * 3, 0.0, True,
                      False, 1, 0)
    with pytest.raises(ValueError, match="stack engine needs"):
        plan.set_engine("stack")
