"""Pins the CPU oracle: golden vectors, closed forms, SciPy and brute force.

These run without a GPU. The oracle is then trusted as the checker of the
GPU kernels (tests/test_gpu_parity.py).
"""

from __future__ import annotations

import itertools
import math
import struct

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle
from tests import _cases


KAT = _cases.load_kat()


def test_boundary_tables_match_reference_literals() -> None:
    tables = KAT["boundary"]
    for key in ("length4", "length2"):
        table = tables[key]
        length = 4 if key == "length4" else 2
        for mode in ("reflect", "mirror", "nearest", "wrap"):
            got = [oracle.map_index(c, length, mode) for c in table["coordinates"]]
            assert got == table[mode], (key, mode)
    constant = tables["constant_length2"]
    assert [oracle.map_index(c, 2, "constant") for c in constant["coordinates"]] == constant["mapped"]
    for entry in tables["many_periods"]:
        assert oracle.map_index(entry["index"], entry["length"], entry["mode"]) == entry["mapped"]


@pytest.mark.parametrize("mode", ["reflect", "nearest", "mirror", "wrap"])
def test_singleton_axes_map_to_zero(mode: str) -> None:
    for index in (-1_000_000, -1, 0, 1, 1_000_000):
        assert oracle.map_index(index, 1, mode) == 0


def test_map_index_rejects_empty_axis() -> None:
    with pytest.raises(oracle.OracleError):
        oracle.map_index(0, 0, "reflect")


def test_map_index_matches_a_folding_model() -> None:
    # Reflect and mirror are the index of a periodic extension: reflect has
    # period 2n with the edge repeated, mirror period 2n-2 without.
    for n in range(2, 7):
        for index in range(-25, 25):
            r = index % (2 * n)
            assert oracle.map_index(index, n, "reflect") == (r if r < n else 2 * n - 1 - r)
            r = index % (2 * n - 2)
            assert oracle.map_index(index, n, "mirror") == (r if r < n else 2 * n - 2 - r)
            assert oracle.map_index(index, n, "wrap") == index % n
            assert oracle.map_index(index, n, "nearest") == min(max(index, 0), n - 1)


@pytest.mark.parametrize("entry", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_oracle_matches_golden_cases(entry: dict) -> None:
    array = _cases.kat_input(entry)
    for method in ("direct", "histogram"):
        if method == "histogram" and entry["kwargs"].get("footprint") is not None:
            continue
        out = _cases.oracle_call(entry["function"], array, _cases.kat_kwargs(entry), entry.get("rank"), method)
        _cases.check_kat(entry, out)


def _f32_bits(value: float) -> int:
    return struct.unpack("<I", struct.pack("<f", value))[0]


def test_keys_round_trip_and_follow_total_order() -> None:
    rng = np.random.default_rng(11)
    specials32 = [-np.inf, -3.4e38, -1.0, -1e-45, -0.0, 0.0, 1e-45, 1.0, 3.4e38, np.inf]
    values = np.array(specials32 + list(rng.normal(size=200)), dtype=np.float32)
    raw = np.append(values.view(np.uint32), rng.integers(0, 2**32, size=500, dtype=np.uint64).astype(np.uint32))
    for b in raw:
        assert oracle.unkey32(oracle.key32(int(b))) == int(b)
    ordered = np.array(sorted(values.tolist(), key=lambda v: (v, math.copysign(1.0, v))), dtype=np.float32)
    keys = [oracle.key32(int(b)) for b in ordered.view(np.uint32)]
    assert keys == sorted(keys)
    assert oracle.key32(_f32_bits(-0.0)) < oracle.key32(_f32_bits(0.0))
    raw64 = rng.integers(0, 2**63, size=500, dtype=np.uint64)
    for b in raw64:
        assert oracle.unkey64(oracle.key64(int(b))) == int(b)
    doubles = np.sort(rng.normal(size=300) * 1e10)
    keys64 = [oracle.key64(int(b)) for b in doubles.view(np.uint64)]
    assert keys64 == sorted(keys64)


def test_midpoint_contract_cases() -> None:
    nan = float("nan")
    assert math.copysign(1.0, oracle.midpoint(-0.0, 0.0)) == 1.0
    assert math.copysign(1.0, oracle.midpoint(-0.0, -0.0)) == -1.0
    assert math.isnan(oracle.midpoint(-math.inf, math.inf))
    assert oracle.midpoint(math.inf, math.inf) == math.inf
    assert oracle.midpoint(1.0, math.inf) == math.inf
    assert oracle.midpoint(-math.inf, 5.0) == -math.inf
    assert oracle.midpoint(2.0, 4.0) == 3.0
    assert oracle.midpoint(-3.0, 5.0) == 1.0
    big = np.finfo(np.float64).max
    assert math.isfinite(oracle.midpoint(np.nextafter(big, 0.0), big))
    # Same-sign branch rounds lo + (hi - lo)/2, not (lo + hi)/2: these differ.
    lo, hi = 1.0, 1.0 + 3 * 2.0**-52
    assert oracle.midpoint(lo, hi) == lo + (hi - lo) / 2.0
    del nan


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode", _cases.MODES)
def test_finite_odd_windows_match_scipy(dtype, mode: str) -> None:
    rng = np.random.default_rng(3)
    values = rng.normal(size=(23, 29)).astype(dtype)
    for size, origin in (((3, 5), (0, 0)), ((5, 3), (1, -1)), ((1, 7), (0, 2)), ((9, 9), (0, 0))):
        window = {"size": size, "origin": origin, "mode": mode, "cval": -7.0}
        expected = ndimage.median_filter(values, **window)
        for method in ("direct", "histogram"):
            got = _cases.oracle_call("median_filter", values, window, method=method)
            np.testing.assert_array_equal(got, expected)
        count = size[0] * size[1]
        for rank in (0, 1, count // 3, count - 1):
            expected = ndimage.rank_filter(values, rank, **window)
            got = _cases.oracle_call("rank_filter", values, window, rank)
            np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_odd_footprints_match_scipy(dtype) -> None:
    rng = np.random.default_rng(4)
    values = rng.normal(size=(17, 19)).astype(dtype)
    yy, xx = np.mgrid[-6:7, -6:7]
    annulus = ((yy**2 + xx**2) >= 16) & ((yy**2 + xx**2) <= 36)
    assert annulus.sum() == 68
    plus = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool)
    for footprint in (plus, annulus):
        count = int(footprint.sum())
        for mode in _cases.MODES:
            window = {"footprint": footprint, "mode": mode, "cval": 2.5}
            for rank in (0, count // 2, count - 1):
                expected = ndimage.rank_filter(values, rank, **window)
                got = _cases.oracle_call("rank_filter", values, window, rank)
                np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("seed", range(40))
def test_tiny_inputs_match_brute_force(seed: int) -> None:
    rng = np.random.default_rng(1000 + seed)
    ndim = int(rng.integers(1, 4))
    shape = tuple(int(rng.integers(1, 6)) for _ in range(ndim))
    size = tuple(int(rng.integers(1, 5)) for _ in range(ndim))
    origin = tuple(int(rng.integers(-(s // 2), (s - 1) // 2 + 1)) for s in size)
    mode = tuple(_cases.MODES[int(rng.integers(0, 5))] for _ in range(ndim))
    dtype = (np.float32, np.float64)[seed % 2]
    values = _cases.special_sample(shape, dtype, seed, fraction=0.3)
    footprint = None
    if seed % 3 == 0:
        footprint = rng.random(size) < 0.6
        footprint.flat[0] = True
    count = int(np.prod(size)) if footprint is None else int(footprint.sum())
    omit = bool(seed % 4 >= 2)
    ignore_inf = bool(seed % 5 == 0) and omit
    min_count = int(rng.integers(1, count + 1)) if omit else 1
    cval = float(dtype([-np.inf, -0.0, 0.0, 1.5, np.inf, np.nan][seed % 6]))
    common = dict(footprint=footprint, origin=origin, mode=mode, cval=cval)
    expected = _cases.brute_force(
        values, size, omit_nan=omit, ignore_inf=ignore_inf, min_count=min_count, **common
    )
    got = oracle.window_filter(values, size, omit_nan=omit, ignore_inf=ignore_inf, min_count=min_count, **common)
    np.testing.assert_array_equal(_cases.bits_of(got), _cases.bits_of(expected))
    rank = int(rng.integers(0, count))
    expected = _cases.brute_force(
        values, size, omit_nan=omit, ignore_inf=ignore_inf, min_count=min_count, rank=rank, **common
    )
    got = oracle.window_filter(
        values,
        size,
        statistic="rank",
        rank=rank,
        omit_nan=omit,
        ignore_inf=ignore_inf,
        min_count=min_count,
        **common,
    )
    np.testing.assert_array_equal(_cases.bits_of(got), _cases.bits_of(expected))


GEOMETRIES = [
    ((150,), (1,), (0,)),
    ((150,), (8,), (-4,)),
    ((70,), (101,), (0,)),
    ((40, 36), (3, 3), (0, 0)),
    ((40, 36), (6, 5), (-3, 2)),
    ((40, 36), (2, 41), (0, -20)),
    ((1, 90), (4, 7), (1, 0)),
    ((9, 20, 17), (3, 2, 5), (-1, 0, 2)),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize(("shape", "size", "origin"), GEOMETRIES)
def test_direct_equals_histogram(dtype, shape, size, origin) -> None:
    """Two independent algorithms (per-output select vs block rank histogram)."""
    values = _cases.special_sample(shape, dtype, seed=len(shape) * 100 + size[0])
    count = int(np.prod(size))
    for mode, policy in itertools.product(_cases.MODES, _cases.POLICIES):
        function, extra = _cases.policy_kwargs(policy, count)
        kwargs = {"size": size, "origin": origin, "mode": mode, "cval": 0.5, **extra}
        direct = _cases.oracle_call(function, values, kwargs)
        hist = _cases.oracle_call(function, values, kwargs, method="histogram")
        np.testing.assert_array_equal(_cases.bits_of(hist), _cases.bits_of(direct))
    for rank in (0, count // 3, count - 1):
        kwargs = {"size": size, "origin": origin, "mode": "wrap", "cval": -0.0}
        direct = _cases.oracle_call("rank_filter", values, kwargs, rank)
        hist = _cases.oracle_call("rank_filter", values, kwargs, rank, method="histogram")
        np.testing.assert_array_equal(_cases.bits_of(hist), _cases.bits_of(direct))


def test_result_is_thread_count_invariant() -> None:
    values = _cases.special_sample((60, 70), np.float32, seed=7)
    outs = [
        oracle.window_filter(values, (9, 12), mode="wrap", threads=t, method=m)
        for t in (1, 3, 8)
        for m in ("direct", "histogram")
    ]
    for out in outs[1:]:
        np.testing.assert_array_equal(_cases.bits_of(out), _cases.bits_of(outs[0]))


def _kat_failures() -> list[str]:
    """Names of the golden cases the brute force gets wrong."""
    failures = []
    for entry in KAT["cases"]:
        values = _cases.kat_input(entry)
        kwargs = _cases.kat_kwargs(entry)
        ndim = values.ndim
        if "footprint" in kwargs:
            fp = np.asarray(kwargs["footprint"]) != 0
            fp = fp.reshape(fp.shape if fp.ndim == ndim else (-1,) * 0 + fp.shape)
            size = list(fp.shape)
        else:
            fp = None
            size = kwargs.get("size", 3)
            size = [int(size)] * ndim if np.ndim(size) == 0 else list(size)
        origin = kwargs.get("origin", 0)
        origin = [int(origin)] * ndim if np.ndim(origin) == 0 else list(origin)
        mode = kwargs.get("mode", "reflect")
        mode = [mode] * ndim if isinstance(mode, str) else list(mode)
        omit = entry["function"] == "nanmedian_filter"
        count = int(fp.sum()) if fp is not None else int(np.prod(size))
        rank = entry.get("rank") if entry["function"] == "rank_filter" else None
        out = _cases.brute_force(
            values, size, footprint=fp, origin=origin, mode=mode,
            cval=float(values.dtype.type(kwargs.get("cval", 0.0))), omit_nan=omit,
            ignore_inf=bool(kwargs.get("ignore_inf", False)) if omit else False,
            min_count=int(kwargs.get("min_count", 1)) if omit else count,
            rank=None if rank is None else int(rank) % count,
        )
        try:
            _cases.check_kat(entry, out)
        except AssertionError:
            failures.append(entry["name"])
    return failures


PLANTED = ["upper_middle", "naive_midpoint", "lower_middle", "signed_zero_lower", "reflect_as_mirror",
           "keep_negative_inf", "cval_not_a_sample"]


def test_brute_force_passes_every_golden_case() -> None:
    assert _kat_failures() == []


@pytest.mark.parametrize("mutation", PLANTED)
def test_planted_mistakes_are_caught(mutation: str, monkeypatch) -> None:
    """Each plausible mistake, planted in the CPU brute force, fails at least
    one golden case: the golden set pins the midpoint contract, the signed
    zero rule, the boundary tables, the ignore-inf policy and cval-as-sample."""
    monkeypatch.setattr(_cases, "MUTATION", mutation)
    assert _kat_failures(), f"no golden case detects {mutation}"


def test_oracle_error_messages() -> None:
    x = np.zeros((4, 4))
    cases = [
        (dict(size=(0, 3)), "window sizes must be positive"),
        (dict(size=(3, 3), origin=(2, 0)), "origin must satisfy -(size // 2) <= origin <= (size - 1) // 2"),
        (dict(size=(3, 3), omit_nan=True, min_count=10), "min_count must be between 1 and the number of window elements"),
        (dict(size=(3, 3), footprint=np.zeros((3, 3))), "footprint must select at least one element"),
        (dict(size=(3, 3), footprint=np.ones((3, 3)), method="histogram"),
         "histogram reductions require rectangular size, not footprint"),
    ]
    for kwargs, message in cases:
        with pytest.raises(oracle.OracleError, match=message.replace("(", r"\(").replace(")", r"\)")):
            oracle.window_filter(x, **kwargs)
    big = np.zeros((70000, 1))
    with pytest.raises(oracle.OracleError, match="overflow"):
        oracle.window_filter(big, (70000, 70000), method="histogram")
