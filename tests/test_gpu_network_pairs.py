"""GPU parity of the network engine's two-row step (windows at rows r and
r + 1 share their core taps; each merges the sorted core with its own sorted
extra taps) against the CPU oracle, bit for bit.

Ring and ring-like footprints take the two-row step (`row_pair` in the plan
description); the tests cover odd row counts (the last pair's second row is
past the array), 3-D footprints, every boundary mode and policy, heavy NaN
(the wide selects), ranks and percentiles (the full networks), and equality
with the one-output network and the generic engine.
"""

from __future__ import annotations

import itertools
import json

import numpy as np
import pytest

from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases

pytestmark = pytest.mark.gpu


def ring(inner: int, outer: int) -> np.ndarray:
    yy, xx = np.mgrid[-outer:outer + 1, -outer:outer + 1]
    r2 = yy * yy + xx * xx
    return (r2 >= inner * inner) & (r2 <= outer * outer)


def plan_for(values, footprint, *, origin=None, mode="reflect", cval=0.0, kind="median", omit=True,
             ignore_inf=False, min_count=1, parameter=0, engine="network") -> _core.Plan:
    ndim = values.ndim
    origin = (0,) * ndim if origin is None else tuple(origin)
    modes = (mode,) * ndim if isinstance(mode, str) else tuple(mode)
    plan = _core.Plan(values.dtype, values.shape, kind, list(footprint.shape), footprint.ravel().tolist(), origin,
                      modes, cval, omit, ignore_inf, min_count, parameter)
    plan.set_engine(engine)
    return plan


def check(actual, expected, context: str) -> None:
    _cases.assert_no_poison(actual, context)
    a, e = _cases.bits_of(actual), _cases.bits_of(expected)
    if not np.array_equal(a, e):
        bad = np.argwhere(a != e)
        index = tuple(bad[0])
        raise AssertionError(f"{context}: {len(bad)} mismatches, first at {index}: got {actual[index]!r} "
                             f"expected {expected[index]!r}")


# Footprints that take the two-row step (classes 48+20, 64+24, 96+32, 32+12).
FOOTPRINTS = {
    "ring4_6": ring(4, 6),
    "ring3_5": ring(3, 5),
    "ring5_7": ring(5, 7),
    "ring4_7": ring(4, 7),
    "ring2_4": ring(2, 4),
}


def test_ring_footprints_take_the_two_row_step() -> None:
    values = np.zeros((40, 40), dtype=np.float32)
    for name, fp in FOOTPRINTS.items():
        info = json.loads(plan_for(values, fp).describe())
        assert info["row_pair"][0] > 0, (name, info)
    # float64, and footprints the step would not make cheaper, keep one output
    # per step
    info = json.loads(plan_for(values.astype(np.float64), FOOTPRINTS["ring2_4"]).describe())
    assert info["row_pair"] == [0, 0]
    for fp in (ring(2, 3), ring(6, 8), np.ones((4, 4), dtype=bool)):
        assert json.loads(plan_for(values, fp).describe())["row_pair"] == [0, 0]


@pytest.mark.parametrize("name", sorted(FOOTPRINTS))
def test_two_row_step_matches_oracle_on_special_values(name: str) -> None:
    fp = FOOTPRINTS[name]
    count = int(fp.sum())
    for shape in ((37, 70), (1, 90), (66, 129)):
        values = _cases.special_sample(shape, np.float32, seed=sum(shape) + count)
        origin = ((fp.shape[0] - 1) // 2, -(fp.shape[1] // 2))
        for mode, policy in itertools.product(_cases.MODES, _cases.POLICIES):
            function, extra = _cases.policy_kwargs(policy, count)
            omit = function == "nanmedian_filter"
            kwargs = {"footprint": fp, "origin": origin, "mode": mode, "cval": -1.5, **extra}
            expected = _cases.oracle_call(function, values, kwargs)
            plan = plan_for(values, fp, origin=origin, mode=mode, cval=-1.5, omit=omit,
                            ignore_inf=bool(extra.get("ignore_inf", False)),
                            min_count=int(extra.get("min_count", 1)) if omit else 1)
            assert json.loads(plan.describe())["row_pair"][0] > 0
            check(plan.apply(values), expected, f"{name} {shape} {mode} {policy}")


@pytest.mark.parametrize("name", ["ring4_6", "ring2_4", "ring4_7"])
def test_two_row_step_ranks_and_percentiles(name: str) -> None:
    fp = FOOTPRINTS[name]
    count = int(fp.sum())
    values = _cases.special_sample((45, 77), np.float32, seed=count)
    for rank in sorted({0, 1, count // 2, count - 1}):
        expected = oracle.window_filter(values, fp.shape, footprint=fp, mode="wrap", statistic="rank", rank=rank,
                                        omit_nan=True)
        check(plan_for(values, fp, mode="wrap", kind="rank", parameter=rank).apply(values), expected,
              f"{name} rank {rank}")
    for q in (0.0, 16.0, 50.0, 84.0, 100.0):
        expected = oracle.window_filter(values, fp.shape, footprint=fp, mode="mirror", statistic="percentile", q=q,
                                        omit_nan=True, ignore_inf=True)
        plan = plan_for(values, fp, mode="mirror", kind="percentile", ignore_inf=True, parameter=q)
        check(plan.apply(values), expected, f"{name} percentile {q}")


def test_two_row_step_with_nan_blocks_and_3d() -> None:
    rng = np.random.default_rng(12)
    values = rng.poisson(200.0, size=(97, 140)).astype(np.float32)
    values[rng.random(values.shape) < 0.3] = np.nan
    values[20:50, 30:60] = np.nan
    values[rng.random(values.shape) < 0.01] = np.inf
    fp = FOOTPRINTS["ring4_6"]
    expected = oracle.window_filter(values, fp.shape, footprint=fp, origin=(1, 0), mode="wrap",
                                    statistic="median", omit_nan=True, ignore_inf=True)
    pair = plan_for(values, fp, origin=(1, 0), mode="wrap", ignore_inf=True).apply(values)
    check(pair, expected, "nan blocks")
    generic = plan_for(values, fp, origin=(1, 0), mode="wrap", ignore_inf=True, engine="generic").apply(values)
    check(pair, generic, "pair vs generic")
    # 3-D: a ring in each of two planes
    cube = rng.poisson(200.0, size=(4, 31, 70)).astype(np.float32)
    cube[rng.random(cube.shape) < 0.05] = np.nan
    fp3 = np.stack([ring(3, 5)] * 2)
    info = json.loads(plan_for(cube, fp3).describe())
    assert info["row_pair"][0] > 0
    expected = oracle.window_filter(cube, fp3.shape, footprint=fp3, mode="reflect", statistic="median",
                                    omit_nan=True)
    check(plan_for(cube, fp3).apply(cube), expected, "3-D rings")
