"""GPU parity: the HIP engines against the CPU oracle, bit for bit.

Every test calls the C-ABI library (libndfgpu.so) on the device. Inputs carry
ties, signed zeros, infinities, NaN payloads and the largest finite value
(tests/test_histogram_median.py:34-46 law); every engine that can run a
geometry (slide, generic, network for small windows, stack for box windows)
is forced in turn and must give the oracle's bits.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import ndfiltrs_amd as nf
from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    count = _core.device_count()
    assert count > 0, "no HIP device visible: GPU tests need an MI355X"


def run_plan(values: np.ndarray, function: str, kwargs: dict, rank: int | None = None, engine: str = "auto",
             method: str = "direct") -> np.ndarray:
    """Call the native plan directly so an engine can be forced."""
    ndim = values.ndim
    footprint = kwargs.get("footprint")
    if footprint is not None:
        fp = np.asarray(footprint) != 0
        size, flat, count = list(fp.shape), fp.ravel().tolist(), int(fp.sum())
    else:
        size = kwargs.get("size", 3)
        size = [int(size)] * ndim if np.ndim(size) == 0 else list(size)
        flat, count = None, int(np.prod(size))
    origin = kwargs.get("origin", 0)
    origin = [int(origin)] * ndim if np.ndim(origin) == 0 else list(origin)
    mode = kwargs.get("mode", "reflect")
    mode = [mode] * ndim if isinstance(mode, str) else list(mode)
    cval = float(values.dtype.type(kwargs.get("cval", 0.0)))
    histogram = method == "histogram"
    if function == "rank_filter":
        kind, omit, ignore_inf, min_count, parameter = (
            "rank_histogram" if histogram else "rank", False, False, 1, int(rank) % count)
    elif function == "median_filter":
        kind, omit, ignore_inf, min_count, parameter = (
            "median_histogram" if histogram else "median", False, False, 1, 0)
    else:
        kind, omit = ("median_histogram" if histogram else "median"), True
        ignore_inf, min_count, parameter = bool(kwargs.get("ignore_inf", False)), int(kwargs.get("min_count", 1)), 0
    plan = _core.Plan(values.dtype, values.shape, kind, size, flat, origin, mode, cval, omit, ignore_inf,
                      min_count, parameter)
    if engine != "auto":
        try:
            plan.set_engine(engine)
        except ValueError as error:
            # The network engine takes small windows only; other engines must
            # always be available where the tests force them.
            if engine in ("network", "stack") and f"{engine} engine needs" in str(error):
                return None
            raise
    return plan.apply(np.ascontiguousarray(values))


# Every engine that can run a geometry must give the oracle's bits.
ENGINES = ("slide", "generic", "network", "stack")


def assert_bits(actual: np.ndarray | None, expected: np.ndarray, context: str = "") -> None:
    if actual is None:  # engine not applicable to this geometry
        return
    _cases.assert_no_poison(actual, context)
    a, e = _cases.bits_of(actual), _cases.bits_of(expected)
    if not np.array_equal(a, e):
        bad = np.argwhere(a != e)
        index = tuple(bad[0])
        raise AssertionError(
            f"{context}: {len(bad)} mismatches, first at {index}: got {actual[index]!r} "
            f"({int(a[index]):x}) expected {expected[index]!r} ({int(e[index]):x})"
        )


KAT = _cases.load_kat()


@pytest.mark.parametrize("entry", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_gpu_matches_golden_cases(entry: dict) -> None:
    array = _cases.kat_input(entry)
    kwargs = _cases.kat_kwargs(entry)
    _cases.check_kat(entry, _cases.public_call(nf, entry["function"], array, kwargs, entry.get("rank")))
    for engine in ("generic", "slide", "network", "stack"):
        try:
            out = run_plan(array, entry["function"], kwargs, entry.get("rank"), engine)
        except ValueError as error:
            assert "cannot stage" in str(error)
            continue
        if out is not None:
            _cases.check_kat(entry, out)


GEOMETRIES = [
    ((150,), (1,), (0,)),
    ((150,), (8,), (-4,)),
    ((150,), (33,), (16,)),
    ((70,), (101,), (0,)),
    ((70, 66), (3, 3), (0, 0)),
    ((70, 66), (6, 5), (-3, 2)),
    ((70, 66), (2, 41), (0, -20)),
    ((1, 90), (4, 7), (1, 0)),
    ((9, 40, 37), (3, 2, 5), (-1, 0, 2)),
    ((9, 40, 37), (12, 1, 4), (5, 0, -2)),
    ((70, 66), (11, 11), (2, -3)),
    ((9, 40, 37), (3, 7, 7), (1, 0, -2)),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize(("shape", "size", "origin"), GEOMETRIES)
def test_engines_match_oracle_on_special_values(dtype, shape, size, origin) -> None:
    values = _cases.special_sample(shape, dtype, seed=len(shape) * 100 + size[0])
    count = int(np.prod(size))
    for mode, policy in itertools.product(_cases.MODES, _cases.POLICIES):
        function, extra = _cases.policy_kwargs(policy, count)
        kwargs = {"size": size, "origin": origin, "mode": mode, "cval": 0.5, **extra}
        expected = oracle_direct = _cases.oracle_call(function, values, kwargs)
        for engine in ENGINES:
            assert_bits(run_plan(values, function, kwargs, engine=engine), expected,
                        f"{engine} {function} {mode} {policy}")
        del oracle_direct
    for rank in sorted({0, count // 3, count - 1}):
        kwargs = {"size": size, "origin": origin, "mode": "wrap", "cval": -0.0}
        expected = _cases.oracle_call("rank_filter", values, kwargs, rank)
        for engine in ENGINES:
            assert_bits(run_plan(values, "rank_filter", kwargs, rank, engine=engine), expected, f"{engine} rank {rank}")
        assert_bits(run_plan(values, "rank_filter", kwargs, rank, method="histogram"), expected, "rank histogram kind")


@pytest.mark.parametrize("cval", [np.nan, np.inf, -np.inf, -0.0, 0.0, 3.0])
@pytest.mark.parametrize("policy", list(_cases.POLICIES))
def test_constant_values_follow_the_sample_policy(cval: float, policy: str) -> None:
    values = _cases.special_sample((40, 45), np.float64, seed=7)
    function, extra = _cases.policy_kwargs(policy, 20)
    kwargs = {"size": (5, 4), "origin": (2, -2), "mode": ("constant", "wrap"), "cval": cval, **extra}
    expected = _cases.oracle_call(function, values, kwargs)
    assert_bits(_cases.public_call(nf, function, values, kwargs), expected, "public")
    assert_bits(run_plan(values, function, kwargs, engine="generic"), expected, "generic")


def annulus(inner: int, outer: int) -> np.ndarray:
    yy, xx = np.mgrid[-outer:outer + 1, -outer:outer + 1]
    r2 = yy**2 + xx**2
    return (r2 >= inner**2) & (r2 <= outer**2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode", _cases.MODES)
def test_footprints_match_oracle(dtype, mode: str) -> None:
    rng = np.random.default_rng(17)
    values = _cases.special_sample((57, 83), dtype, seed=23, fraction=0.08)
    footprints = [
        annulus(4, 6),
        np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool),
        np.indices((3, 4)).sum(axis=0) % 3 != 0,
        rng.random((7, 5)) < 0.5,
        np.array([[1, 1]], dtype=bool),
    ]
    for index, footprint in enumerate(footprints):
        footprint = footprint.copy()
        footprint.flat[0] = True if index == 3 else footprint.flat[0]
        count = int(footprint.sum())
        origin = (footprint.shape[0] - 1) // 2, -(footprint.shape[1] // 2)
        for function, extra in (("median_filter", {}), ("nanmedian_filter", {"ignore_inf": True, "min_count": 2})):
            kwargs = {"footprint": footprint, "origin": origin, "mode": mode, "cval": 1.25, **extra}
            expected = _cases.oracle_call(function, values, kwargs)
            for engine in ENGINES:
                assert_bits(run_plan(values, function, kwargs, engine=engine), expected, f"{engine} fp{index}")
        kwargs = {"footprint": footprint, "mode": mode, "cval": 1.25}
        rank = count // 2
        expected = _cases.oracle_call("rank_filter", values, kwargs, rank)
        assert_bits(_cases.public_call(nf, "rank_filter", values, kwargs, rank), expected, f"rank fp{index}")


def test_rank_five_mixed_modes_uses_the_generic_engine() -> None:
    values = np.arange(48.0, dtype=np.float64).reshape(2, 2, 2, 2, 3)
    values[0, 1, 0, 1, 2] = np.nan
    kwargs = {"size": (3, 3, 3, 3, 5), "mode": ("reflect", "constant", "nearest", "mirror", "wrap"), "cval": -4.0,
              "min_count": 7}
    expected = _cases.oracle_call("nanmedian_filter", values, kwargs)
    assert_bits(nf.nanmedian_filter(values, **kwargs), expected, "rank-5")


def test_three_dimensional_policies_match_oracle() -> None:
    values = np.array([[[1.0, np.nan], [np.inf, 4.0]], [[5.0, 6.0], [7.0, -np.inf]]], dtype=np.float64)
    kwargs = {"size": (3, 1, 3), "mode": ("reflect", "constant", "wrap"), "cval": np.nan, "ignore_inf": True,
              "min_count": 2}
    expected = _cases.oracle_call("nanmedian_filter", values, kwargs)
    assert_bits(nf.nanmedian_filter(values, **kwargs), expected, "3-d")


@pytest.mark.parametrize("seed", range(24))
def test_generated_geometries_match_oracle(seed: int) -> None:
    rng = np.random.default_rng(5000 + seed)
    ndim = int(rng.integers(1, 4))
    longest = {1: 120, 2: 50, 3: 12}[ndim]
    shape = tuple(int(rng.integers(1, longest + 1)) for _ in range(ndim))
    size = tuple(int(rng.integers(1, 10)) for _ in range(ndim))
    origin = tuple(int(rng.integers(-(s // 2), (s - 1) // 2 + 1)) for s in size)
    mode = tuple(_cases.MODES[int(rng.integers(0, 5))] for _ in range(ndim))
    dtype = (np.float32, np.float64)[seed % 2]
    values = _cases.special_sample(shape, dtype, seed)
    cval = float([-np.inf, -0.0, 0.0, 1.0, np.inf, np.nan][seed % 6])
    kwargs = {"size": size, "origin": origin, "mode": mode, "cval": cval}
    if seed % 2:
        kwargs.update(ignore_inf=bool(seed % 3), min_count=int(rng.integers(1, int(np.prod(size)) + 1)))
        function = "nanmedian_filter"
    else:
        function = "median_filter"
    expected = _cases.oracle_call(function, values, kwargs)
    for method in ("direct", "sortnet", "histogram"):
        assert_bits(getattr(nf, function)(values, method=method, **kwargs), expected, method)
    assert_bits(run_plan(values, function, kwargs, engine="generic"), expected, "generic")


def sky(shape, dtype, seed: int, nan_fraction: float = 0.02) -> np.ndarray:
    rng = np.random.default_rng(seed)
    values = rng.normal(1000.0, 10.0, size=shape)
    values[rng.random(shape) < nan_fraction] = np.nan
    return values.astype(dtype)


@pytest.mark.parametrize(
    ("shape", "size", "dtype", "mode"),
    [
        ((300, 260), (63, 63), np.float32, "reflect"),
        ((280, 300), (31, 31), np.float64, "nearest"),
        ((200, 333), (45, 61), np.float32, "mirror"),
        ((130, 140), (64, 2), np.float64, "wrap"),
    ],
)
def test_large_windows_match_oracle(shape, size, dtype, mode: str) -> None:
    values = sky(shape, dtype, seed=size[0] + shape[1])
    kwargs = {"size": size, "mode": mode}
    expected = _cases.oracle_call("nanmedian_filter", values, kwargs, method="histogram")
    assert_bits(nf.nanmedian_filter(values, **kwargs), expected, "nanmedian")
    count = size[0] * size[1]
    for q in (0.10, 0.16, 0.84):
        rank = int(np.floor(q * (count - 1) + 0.5))
        expected = oracle.window_filter(values, size, mode=mode, statistic="rank", rank=rank, omit_nan=True,
                                        method="histogram")
        plan = _core.Plan(values.dtype, values.shape, "rank", size, None, (0, 0), (mode, mode), 0.0, True, False, 1,
                          rank)
        assert_bits(plan.apply(values), expected, f"omit rank {rank}")
    expected = _cases.oracle_call("rank_filter", values, kwargs, count // 7, method="histogram")
    assert_bits(nf.rank_filter(values, count // 7, **kwargs), expected, "propagate rank")


@pytest.mark.parametrize(
    ("shape", "size", "dtype", "mode"),
    [
        ((400, 300), (31, 31), np.float32, "reflect"),
        ((330, 270), (63, 63), np.float32, "nearest"),
        ((300, 200), (31, 31), np.float64, "mirror"),
        ((9, 120, 140), (3, 9, 9), np.float32, "wrap"),
    ],
)
def test_count_gate_keeps_tiles_with_results(shape, size, dtype, mode: str) -> None:
    """Propagate and high-min_count plans skip the sort of tiles whose
    windows all lack a result: NaNs cover the left part of the image (every
    window there has one), the right part is clean, and one band holds
    single NaNs so windows there sit just below or at min_count."""
    rng = np.random.default_rng(sum(shape) + size[-1])
    values = rng.normal(1000.0, 10.0, size=shape)
    width = shape[-1]
    left = values[..., : width // 3]
    left[rng.random(left.shape) < 0.05] = np.nan
    band = values[..., width // 2: width // 2 + 8]
    band[rng.random(band.shape) < 0.002] = np.nan
    values = values.astype(dtype)
    count = int(np.prod(size))
    modes = (mode,) * len(shape)
    kwargs = {"size": size, "mode": mode}
    expected = _cases.oracle_call("median_filter", values, kwargs, method="histogram")
    assert_bits(nf.median_filter(values, **kwargs), expected, "propagate median")
    rank = count // 5
    expected = _cases.oracle_call("rank_filter", values, kwargs, rank, method="histogram")
    assert_bits(nf.rank_filter(values, rank, **kwargs), expected, "propagate rank")
    for min_count in (count - 1, count // 2 + 1):
        expected = oracle.window_filter(values, size, mode=mode, statistic="median", omit_nan=True,
                                        min_count=min_count, method="histogram")
        plan = _core.Plan(values.dtype, values.shape, "median", size, None, (0,) * len(shape), modes, 0.0, True,
                          False, min_count, 0)
        assert_bits(plan.apply(values), expected, f"omit min_count {min_count}")
    # Everything covered: every tile takes the gate.
    values[..., :] = np.where(rng.random(shape) < 0.3, np.nan, values)
    expected = _cases.oracle_call("rank_filter", values, kwargs, rank, method="histogram")
    assert_bits(nf.rank_filter(values, rank, **kwargs), expected, "all NaN windows")


def test_window_larger_than_array() -> None:
    values = _cases.special_sample((13, 17), np.float32, seed=99)
    for mode in _cases.MODES:
        kwargs = {"size": (40, 51), "mode": mode, "cval": 2.0}
        expected = _cases.oracle_call("nanmedian_filter", values, kwargs)
        assert_bits(nf.nanmedian_filter(values, **kwargs), expected, mode)


def test_results_do_not_depend_on_engine_or_repetition() -> None:
    values = sky((190, 250), np.float32, seed=3)
    plan = _core.Plan(np.float32, values.shape, "median", (21, 21), None, (0, 0), ("reflect",) * 2, 0.0, True, False,
                      1, 0)
    first = plan.apply(values)
    again = plan.apply(values)
    plan.set_engine("generic")
    generic = plan.apply(values)
    assert_bits(again, first, "repeat")
    assert_bits(generic, first, "engines")


def test_public_call_snapshots_input_and_returns_owned_output() -> None:
    source = np.arange(80, dtype=np.float32).reshape(8, 10) - 40
    source[2, :5] = [np.nan, np.inf, -np.inf, -0.0, 0.0]
    for array in (source.T, source[::-1, ::2], np.asfortranarray(source)):
        array.flags.writeable = False
        original = array.copy()
        options = dict(size=(2, 3), origin=(-1, 1), mode=("mirror", "constant"))
        expected = _cases.oracle_call("median_filter", np.ascontiguousarray(array), options)
        for method in ("direct", "sortnet", "histogram"):
            actual = nf.median_filter(array, method=method, **options)
            assert_bits(actual, expected, method)
            assert actual.flags.owndata and actual.flags.c_contiguous and actual.flags.writeable
            assert not np.shares_memory(actual, array)
        assert_bits(array, original, "input untouched")


def test_output_never_aliases_a_patched_allocator(monkeypatch) -> None:
    external = np.full(3, 99.0)
    monkeypatch.setattr(np, "empty", lambda *a, **k: external)
    result = nf.median_filter(np.array([1.0, 2.0, 3.0]), 1)
    np.testing.assert_array_equal(result, [1.0, 2.0, 3.0])
    np.testing.assert_array_equal(external, [99.0, 99.0, 99.0])
    assert result is not external and result.flags.owndata
