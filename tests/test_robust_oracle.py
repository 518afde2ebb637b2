"""The oracle's percentile and extrema, pinned without the reference binary.

Percentile (src/robust.rs:320-367): the reference's literal answers
(src/robust.rs:876-892, :1042-1053; tests/test_robust.py:23-191, :322-334),
SciPy's generic filter with NumPy's linear percentile on the reference's own
test array, NumPy's percentile on same-sign windows, an exact-rational bound
on opposite-sign interpolation, and an independent brute force that each of
five planted mistakes makes fail a literal answer.

Extrema (src/reduction.rs:110-119, src/extrema.rs:125-436): the oracle's two
algorithms (direct total-order extreme, separable van Herk passes) agree bit
for bit on special values in every mode and policy, agree with SciPy's
minimum/maximum filters on finite data, and give the reference's literal
signed-zero, infinity and repetition answers.
"""

from __future__ import annotations

import itertools
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import ndimage

from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases

MODES = _cases.MODES
DTYPES = (np.float32, np.float64)
TINY = 5e-324


def pct(values, q, size=None, omit=False, mode="wrap", **kwargs):
    values = np.asarray(values)
    size = size or values.shape
    return oracle.window_filter(values, size, statistic="percentile", q=q, omit_nan=omit, mode=mode, **kwargs)


# ---- percentile: literal answers of the reference --------------------------


def test_interpolation_literals() -> None:
    big = np.finfo(np.float64).max
    assert oracle.interpolate(-big, big, 0.5) == 0.0
    assert oracle.interpolate(big, big, 0.75) == big
    assert oracle.interpolate(2.0, math.inf, 0.0) == 2.0
    assert oracle.interpolate(2.0, math.inf, 0.5) == math.inf
    assert oracle.interpolate(-math.inf, 2.0, 1.0) == 2.0
    assert math.isnan(oracle.interpolate(-math.inf, math.inf, 0.5))
    assert np.float64(oracle.interpolate(-0.0, 0.0, 0.5)).view(np.uint64) == 0


def literal_cases() -> list[tuple[str, object, object]]:
    """(name, call, expected): the reference's literal percentile answers.
    `call` maps a percentile function f(values, q, size, omit, mode, **kw)
    to an output array."""
    cases = []
    for q, want in ((0.0, 1.0), (25.0, 2.5), (50.0, 4.0), (100.0, 9.0)):
        cases.append((f"nine_one_four_q{q}", lambda f, q=q: f([9.0, 1.0, 4.0], q, None, False, "wrap"), want))
    cases.append(("subnormal_cancellation", lambda f: f([-TINY, 2 * TINY, np.nan], 50.0, None, True, "wrap"), 0.0))
    for dtype in DTYPES:
        d = np.dtype(dtype).name
        cases.append((f"even_retained_{d}", lambda f, t=dtype: f(np.array([np.nan, 1, 3, 9, 11], t), 25.0, (5,),
                                                                 True, "wrap"), 2.5))
        cases.append((f"min_count_{d}", lambda f, t=dtype: f(np.array([np.nan, 1, 3, 9, 11], t), 25.0, (5,), True,
                                                            "wrap", min_count=5), math.nan))
        big = float(np.finfo(dtype).max)
        for q, want in ((0, -big), (25, -big / 2), (50, 0.0), (100, big)):
            cases.append((f"extremes_q{q}_{d}", lambda f, t=dtype, q=q, b=big: f(np.array([-b, np.nan, b], t), q,
                                                                                  None, True, "wrap"), want))
        cases.append((f"inf_q0_{d}", lambda f, t=dtype: f(np.array([-np.inf, np.nan, np.inf], t), 0.0, None, True,
                                                          "wrap"), -math.inf))
        cases.append((f"inf_q100_{d}", lambda f, t=dtype: f(np.array([-np.inf, np.nan, np.inf], t), 100.0, None,
                                                            True, "wrap"), math.inf))
        cases.append((f"inf_q50_{d}", lambda f, t=dtype: f(np.array([-np.inf, np.nan, np.inf], t), 50.0, None,
                                                           True, "wrap"), math.nan))
        cases.append((f"all_inf_{d}", lambda f, t=dtype: f(np.full(3, np.inf, t), 25.0, None, False, "wrap"),
                      math.inf))
        cases.append((f"ignore_inf_{d}", lambda f, t=dtype: f(np.array([1, 3, np.inf], t), 50.0, None, True, "wrap",
                                                              ignore_inf=True), 2.0))
        cases.append((f"nan_propagates_{d}", lambda f, t=dtype: f(np.array([1, np.nan, 3], t), 50.0, (3,), False,
                                                                  "reflect"), math.nan))
        cases.append((f"zero_q0_{d}", lambda f, t=dtype: f(np.array([0.0, -0.0, 0.0], t), 0.0, None, False, "wrap"),
                      -0.0))
        cases.append((f"zero_q100_{d}", lambda f, t=dtype: f(np.array([0.0, -0.0, 0.0], t), 100.0, None, False,
                                                             "wrap"), 0.0))
        cases.append((f"repeats_count_{d}", lambda f, t=dtype: f(np.array([5.0], t), 50.0, (3,), True, "reflect",
                                                                 min_count=3), 5.0))
        cases.append((f"cval_counts_{d}", lambda f, t=dtype: f(np.array([np.nan], t), 50.0, (3,), True, "constant",
                                                               cval=2.0, min_count=2), 2.0))
    return cases


def oracle_pct(values, q, size, omit, mode, **kwargs):
    values = np.asarray(values, dtype=np.float64) if not isinstance(values, np.ndarray) else values
    return pct(values, q, size or values.shape, omit, mode, **kwargs)


def check_literal(out: np.ndarray, want: float) -> None:
    out = np.asarray(out)
    if math.isnan(want):
        assert np.all(_cases.bits_of(out) == (0x7FC00000 if out.dtype == np.float32 else 0x7FF8000000000000))
    else:
        expected = np.full(out.shape, want, dtype=out.dtype)
        assert np.array_equal(_cases.bits_of(out), _cases.bits_of(expected)), (out, want)


@pytest.mark.parametrize("case", literal_cases(), ids=[c[0] for c in literal_cases()])
def test_oracle_percentile_literals(case) -> None:
    _, call, want = case
    check_literal(call(oracle_pct), want)


def test_mixed_modes_and_size_one_axes() -> None:
    values = np.array([[[1.0, 2.0, 3.0]], [[10.0, 20.0, 30.0]]])
    out = pct(values, 25.0, (3, 1, 1), mode=("constant", "nearest", "wrap"), cval=5.0)
    np.testing.assert_array_equal(out, [[[3, 3.5, 4]], [[3, 3.5, 4]]])


def test_adjacent_extremes_stay_finite() -> None:
    for dtype in DTYPES:
        big = np.finfo(dtype).max
        adjacent = np.nextafter(dtype(big), dtype(0))
        out = pct(np.array([adjacent, np.nan, big], dtype), 50.0, omit=True)
        assert np.all(np.isfinite(out)) and np.all((out >= adjacent) & (out <= big))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("q", [0.0, 12.5, 50.0, 87.5, 100.0])
def test_scipy_generic_filter_with_numpy_percentile(dtype, mode: str, q: float) -> None:
    array = np.array([[3, -4, 8], [1, 12, -2]], dtype=dtype)
    expected = ndimage.generic_filter(array, lambda v: np.percentile(v, q, method="linear"), size=(3, 5), mode=mode,
                                      cval=1.25)
    got = oracle.window_filter(array, (3, 5), statistic="percentile", q=q, mode=mode, cval=1.25)
    np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("seed", range(6))
def test_numpy_percentile_on_same_sign_windows(seed: int) -> None:
    """NumPy's linear percentile interpolates from the nearer end with the
    same operations as the reference for same-sign samples; its only
    difference, the branch taken at a fraction of exactly 0.5, is avoided."""
    rng = np.random.default_rng(seed)
    values = rng.uniform(1.0, 1000.0, size=(23, 19))
    size = (3, 5)
    count = 15
    for q in (7.0, 33.0, 61.0, 91.0):
        fraction = q / 100.0 * (count - 1) % 1.0
        assert fraction != 0.5
        got = oracle.window_filter(values, size, statistic="percentile", q=q, mode="reflect")
        want = ndimage.generic_filter(values, lambda v: np.percentile(v, q, method="linear"), size=size,
                                      mode="reflect")
        np.testing.assert_array_equal(got, want)


def test_opposite_sign_interpolation_is_within_two_units_of_the_endpoints() -> None:
    rng = np.random.default_rng(11)
    for _ in range(3000):
        lower = -float(rng.uniform(0, 1)) * 10.0 ** float(rng.integers(-300, 300))
        upper = float(rng.uniform(0, 1)) * 10.0 ** float(rng.integers(-300, 300))
        fraction = float(rng.uniform(0, 1))
        got = oracle.interpolate(lower, upper, fraction)
        exact = Fraction(lower) * (1 - Fraction(fraction)) + Fraction(upper) * Fraction(fraction)
        assert lower <= got <= upper
        # Rounding of 1 - fraction, of each product and of their sum: within
        # two units of the larger endpoint's last place.
        scale = np.spacing(max(abs(lower), abs(upper)))
        assert abs(Fraction(got) - exact) <= 2 * Fraction(scale), (lower, upper, fraction)


def test_percentile_is_monotone_in_q() -> None:
    values = _cases.special_sample((31, 29), np.float64, seed=4, fraction=0.1)
    previous = None
    for q in np.linspace(0.0, 100.0, 41):
        out = oracle.window_filter(values, (3, 4), statistic="percentile", q=float(q), omit_nan=True,
                                   ignore_inf=True, mode="mirror")
        if previous is not None:
            both = ~np.isnan(out) & ~np.isnan(previous)
            assert np.all(out[both] >= previous[both])
        previous = out


# ---- an independent brute force, and planted mistakes ----------------------

MUTATION: str | None = None


def _lerp(lower: float, upper: float, fraction: float) -> float:
    """Interpolation from docs of src/robust.rs:320-356, written from its
    rules: endpoints exact, opposite infinities NaN, one infinity wins,
    opposite signs by a two-term compensated sum of rounded products,
    same signs from the nearer end."""
    if MUTATION == "naive_lerp":
        return lower + (upper - lower) * fraction
    if fraction == 0.0 or (lower == upper and math.copysign(1, lower) == math.copysign(1, upper)):
        return lower
    if fraction == 1.0:
        return upper
    if math.isinf(lower) or math.isinf(upper):
        if lower == -math.inf and upper == math.inf:
            return math.nan
        return lower if math.isinf(lower) else upper
    if math.copysign(1, lower) != math.copysign(1, upper):
        if math.isinf(upper - lower):
            return min(max((lower / 2 + (upper / 2 - lower / 2) * fraction) * 2, lower), upper)
        a, b = lower * (1 - fraction), upper * fraction
        if (a != 0 and abs(a) < 2.2250738585072014e-308) or (b != 0 and abs(b) < 2.2250738585072014e-308):
            # Products below the normal range: the exact sum, rounded once.
            return float(Fraction(lower) * Fraction(1 - fraction) + Fraction(upper) * Fraction(fraction))
        return float(Fraction(a) + Fraction(b))
    diff = upper - lower
    result = lower + diff * fraction if fraction <= 0.5 else upper - diff * (1 - fraction)
    return min(max(result, lower), upper)


def brute_percentile(values, q, size, omit, mode, cval=0.0, ignore_inf=False, min_count=1):
    array = np.asarray(values, dtype=np.float64) if not isinstance(values, np.ndarray) else values
    size = size or array.shape
    mode = (mode,) * array.ndim if isinstance(mode, str) else tuple(mode)
    anchors = [n // 2 for n in size]
    taps = list(itertools.product(*[range(n) for n in size]))
    out = np.empty_like(array)
    for center in np.ndindex(array.shape):
        samples, nan = [], False
        for tap in taps:
            mapped = [_cases._map(c + t - a, n, m) for c, t, a, n, m in zip(center, tap, anchors, array.shape, mode)]
            value = float(array.dtype.type(cval)) if None in mapped else float(array[tuple(mapped)])
            if math.isnan(value):
                nan = True
                continue
            if omit and ignore_inf and math.isinf(value):
                continue
            samples.append(value)
        key = (lambda v: v) if MUTATION == "value_order" else (lambda v: (v, math.copysign(1.0, v)))
        samples.sort(key=key)
        n = len(samples) if MUTATION != "rank_over_taps" else len(taps)
        if (not omit and nan) or (omit and len(samples) < min_count) or not samples:
            out[center] = np.nan
            continue
        rank = q / 100.0 * (n - 1)
        lower = math.ceil(rank) if MUTATION == "ceil_rank" else math.floor(rank)
        lower = min(lower, len(samples) - 1)
        upper = min(lower + 1, len(samples) - 1)
        value = _lerp(samples[lower], samples[upper], rank - lower)
        out[center] = array.dtype.type(value) if not math.isnan(value) else np.nan
    return out


@pytest.mark.parametrize("case", literal_cases(), ids=[c[0] for c in literal_cases()])
def test_brute_force_passes_the_literals(case) -> None:
    _, call, want = case
    check_literal(call(brute_percentile), want)


@pytest.mark.parametrize("mutation", ["naive_lerp", "ceil_rank", "rank_over_taps", "value_order"])
def test_planted_percentile_mistakes_are_caught(mutation: str) -> None:
    global MUTATION
    MUTATION = mutation
    try:
        failures = 0
        for _, call, want in literal_cases():
            try:
                check_literal(call(brute_percentile), want)
            except AssertionError:
                failures += 1
        assert failures > 0, mutation
    finally:
        MUTATION = None


@pytest.mark.parametrize("seed", range(8))
def test_oracle_equals_brute_force_on_special_values(seed: int) -> None:
    rng = np.random.default_rng(700 + seed)
    shape = tuple(int(rng.integers(1, 9)) for _ in range(2))
    dtype = DTYPES[seed % 2]
    values = _cases.special_sample(shape, dtype, seed)
    size = tuple(int(rng.integers(1, 5)) for _ in range(2))
    mode = tuple(MODES[int(rng.integers(0, 5))] for _ in range(2))
    for q in (0.0, 30.0, 50.0, 100.0 / 3, 100.0):
        for omit in (False, True):
            got = oracle.window_filter(values, size, statistic="percentile", q=q, omit_nan=omit, mode=mode,
                                       cval=-0.0, origin=(0, 0))
            size_centered = size
            want = brute_percentile(values, q, size_centered, omit, mode, cval=-0.0)
            # brute_percentile anchors at n // 2 (origin 0), as the oracle does.
            assert np.array_equal(_cases.bits_of(got), _cases.bits_of(want)), (q, omit, mode, size)


def test_percentile_validation_messages() -> None:
    with pytest.raises(oracle.OracleError, match="percentile must be finite and between zero and 100"):
        oracle.window_filter(np.ones(4), (3,), statistic="percentile", q=101.0)
    with pytest.raises(oracle.OracleError, match="min_count must be between one and the number"):
        oracle.window_filter(np.ones(4), (3,), statistic="percentile", q=5.0, omit_nan=True, min_count=4)
    # The native plan of the product validates in the reference's order too
    # (plan creation runs on the host; no kernel is launched).
    with pytest.raises(ValueError, match="percentile must be finite and between zero and 100"):
        _core.Plan(np.float64, (4,), "percentile", (3,), None, (0,), ("reflect",), 0.0, False, False, 1, -1.0)
    with pytest.raises(ValueError, match="min_count must be between one and the number of window elements"):
        _core.Plan(np.float64, (4,), "percentile", (3,), None, (0,), ("reflect",), 0.0, True, False, 4, 50.0)
    with pytest.raises(ValueError, match="ignore_inf and min_count require NaN omission"):
        _core.Plan(np.float64, (4,), "percentile", (3,), None, (0,), ("reflect",), 0.0, False, True, 1, 50.0)
    with pytest.raises(ValueError, match="origin must satisfy"):
        _core.Plan(np.float64, (4,), "percentile", (3,), None, (2,), ("reflect",), 0.0, False, False, 1, 500.0)


# ---- extrema ----------------------------------------------------------------

EXTREMA_GEOMETRIES = [
    ((23,), (4,), (-2,)),
    ((17, 13), (3, 5), (0, 2)),
    ((9, 11), (6, 1), (2, 0)),
    ((5, 7, 6), (3, 2, 4), (-1, 0, 1)),
    ((1, 9), (4, 4), (1, -2)),
]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(("shape", "size", "origin"), EXTREMA_GEOMETRIES)
def test_separable_extrema_equal_direct(dtype, shape, size, origin) -> None:
    values = _cases.special_sample(shape, dtype, seed=len(shape) * 31 + size[0])
    volume = int(np.prod(size))
    policies = [(False, False, 1), (True, False, 1), (True, False, volume), (True, True, 1), (True, True, volume),
                (True, False, max(1, volume // 2))]
    for statistic in ("min", "max"):
        for mode, (omit, ignore_inf, min_count), cval in itertools.product(MODES, policies, (np.nan, -0.0, 2.5)):
            kwargs = dict(statistic=statistic, mode=mode, origin=origin, cval=cval, omit_nan=omit,
                          ignore_inf=ignore_inf, min_count=min_count)
            direct = oracle.window_filter(values, size, **kwargs)
            separable = oracle.window_filter(values, size, method="separable", **kwargs)
            assert np.array_equal(_cases.bits_of(direct), _cases.bits_of(separable)), (statistic, mode, omit,
                                                                                      ignore_inf, min_count, cval)


@pytest.mark.parametrize("mode", MODES)
def test_extrema_match_scipy_on_finite_data(mode: str) -> None:
    rng = np.random.default_rng(5)
    values = rng.normal(size=(19, 23)) + 0.5  # no signed zeros: SciPy does not order them
    footprint = np.array([[0, 1, 0, 1], [1, 1, 0, 0], [0, 1, 1, 1]], dtype=bool)
    for statistic, scipy_filter in (("min", ndimage.minimum_filter), ("max", ndimage.maximum_filter)):
        for kwargs in ({"size": (3, 4), "origin": (1, -2)}, {"footprint": footprint, "origin": (0, 1)}):
            want = scipy_filter(values, mode=mode, cval=-0.75, **kwargs)
            fp = kwargs.get("footprint")
            got = oracle.window_filter(values, fp.shape if fp is not None else kwargs["size"], footprint=fp,
                                       origin=kwargs["origin"], statistic=statistic, mode=mode, cval=-0.75)
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("method", ["direct", "separable"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_extrema_literals(method: str, dtype) -> None:
    def run(values, statistic, size=(3,), mode="wrap", **kwargs):
        return oracle.window_filter(np.array(values, dtype), size, statistic=statistic, mode=mode, method=method,
                                    **kwargs)

    # src/reduction.rs:137-160 and tests/test_statistics.py:215-242.
    assert np.signbit(run([0.0, -0.0, 1.0], "min")).all()
    assert not np.signbit(run([-0.0, 0.0, -1.0], "max")).any()
    for values in ([-0.0, 0.0, -0.0], [0.0, -0.0, 0.0], [-0.0] * 3):
        assert np.signbit(run(values, "min")).all()
    for values in ([-0.0, 0.0, -0.0], [0.0, -0.0, 0.0], [0.0] * 3):
        assert not np.signbit(run(values, "max")).any()
    np.testing.assert_array_equal(run([-np.inf, 2, np.inf], "min"), -np.inf)
    np.testing.assert_array_equal(run([-np.inf, 2, np.inf], "max"), np.inf)
    assert np.signbit(run([0.0], "min", (3,), "constant", cval=-0.0)).all()
    assert not np.signbit(run([-0.0], "max", (3,), "constant", cval=0.0)).any()
    # tests/test_statistics.py:292-316: empty support and repetitions.
    for statistic in ("min", "max"):
        for invalid in (np.nan, np.inf, -np.inf):
            out = run(np.full(3, invalid), statistic, (3,), "reflect", omit_nan=True, ignore_inf=True)
            check_literal(out, math.nan)
        for min_count, want in ((2, 7.0), (3, math.nan)):
            out = run([np.nan], statistic, (3,), "constant", cval=7.0, omit_nan=True, min_count=min_count)
            check_literal(out, want)
        check_literal(run([7.0], statistic, (3,), "reflect", omit_nan=True, min_count=3), 7.0)
        check_literal(run([1.0, np.nan, 3.0], statistic, (3,), "reflect"), math.nan)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("statistic,q", [("percentile", 0.0), ("percentile", 10.0), ("percentile", 50.0),
                                         ("percentile", 100.0 / 7), ("percentile", 100.0), ("min", 0.0),
                                         ("max", 0.0)])
def test_histogram_method_equals_direct(dtype, statistic: str, q: float) -> None:
    """The block rank histogram serves percentiles and extrema too (the
    whole-image parity reference of the benchmark); it equals the direct
    method bit for bit."""
    values = _cases.special_sample((47, 61), dtype, seed=9, fraction=0.1)
    for mode, omit, ignore_inf, min_count in (("reflect", True, False, 1), ("constant", False, False, 1),
                                              ("wrap", True, True, 9), ("mirror", True, False, 40)):
        kwargs = dict(statistic=statistic, q=q, mode=mode, omit_nan=omit, ignore_inf=ignore_inf,
                      min_count=min_count, cval=0.5)
        direct = oracle.window_filter(values, (7, 6), **kwargs)
        histogram = oracle.window_filter(values, (7, 6), method="histogram", **kwargs)
        assert np.array_equal(_cases.bits_of(direct), _cases.bits_of(histogram)), (statistic, q, mode)
