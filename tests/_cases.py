"""Shared inputs and call adapters for the parity tests."""

from __future__ import annotations

import json
import math
import os
from collections.abc import Sequence

import numpy as np

from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
MODES = ("reflect", "constant", "nearest", "mirror", "wrap")
POLICIES = {
    "plain": None,
    "omit": {"ignore_inf": False, "min_count": 1},
    "finite": {"ignore_inf": True, "min_count": 1},
    "all": {"ignore_inf": False, "min_count": -1},
}


# Output poison: NaN payloads that no filter writes (outputs are input
# values, f64 midpoints cast once, or the canonical NaN).
POISON32 = 0x7FA5A5A5
POISON64 = 0x7FF4A5A5A5A5A5A5


def bits_of(array: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(array).view(np.uint32 if array.dtype == np.float32 else np.uint64)


def poison_bits(dtype) -> int:
    return POISON32 if np.dtype(dtype) == np.float32 else POISON64


def assert_no_poison(array: np.ndarray, context: str = "") -> None:
    """Fail when any output still holds the poison pattern (a skipped write)."""
    left = int(np.count_nonzero(bits_of(array) == poison_bits(array.dtype)))
    assert left == 0, f"{context}: {left} outputs were never written"


def special_sample(shape: Sequence[int], dtype, seed: int, fraction: float = 0.15) -> np.ndarray:
    """Quarter-step values with many ties, plus NaN (incl. a negative payload),
    +-inf, +-0 and the largest finite value at ``fraction`` of positions
    (the input law of tests/test_histogram_median.py:34-46)."""
    rng = np.random.default_rng(seed)
    values = (rng.integers(-4, 5, size=shape) / 4).astype(dtype)
    unsigned = np.uint32 if dtype == np.float32 else np.uint64
    payload = 0xFFC01234 if dtype == np.float32 else 0xFFF8000000001234
    special = np.array([np.nan, np.inf, -np.inf, -0.0, 0.0, np.finfo(dtype).max], dtype=dtype)
    special = np.append(special, np.array([payload], dtype=unsigned).view(dtype))
    chosen = rng.random(shape) < fraction
    values[chosen] = rng.choice(special, size=int(chosen.sum()))
    return values


def policy_kwargs(policy: str, count: int) -> tuple[str, dict]:
    options = POLICIES[policy]
    if options is None:
        return "median_filter", {}
    min_count = count if options["min_count"] == -1 else options["min_count"]
    return "nanmedian_filter", {"ignore_inf": options["ignore_inf"], "min_count": min_count}


def _window(kwargs: dict, ndim: int) -> tuple[list[int], np.ndarray | None, int]:
    footprint = kwargs.get("footprint")
    if footprint is not None:
        fp = np.asarray(footprint) != 0
        return list(fp.shape), fp, int(fp.sum())
    size = kwargs.get("size", 3)
    size = [int(size)] * ndim if np.ndim(size) == 0 else [int(s) for s in size]
    return size, None, int(np.prod(size))


def oracle_call(function: str, array: np.ndarray, kwargs: dict, rank: int | None = None, method: str = "direct"):
    """Evaluate a public-API call with the CPU oracle."""
    ndim = array.ndim
    size, fp, count = _window(kwargs, ndim)
    origin = kwargs.get("origin", 0)
    origin = [int(origin)] * ndim if np.ndim(origin) == 0 else [int(o) for o in origin]
    mode = kwargs.get("mode", "reflect")
    mode = [mode] * ndim if isinstance(mode, str) else list(mode)
    cval = float(array.dtype.type(kwargs.get("cval", 0.0)))
    common = dict(footprint=fp, origin=origin, mode=mode, cval=cval, method=method)
    if function == "median_filter":
        return oracle.window_filter(array, size, omit_nan=False, **common)
    if function == "nanmedian_filter":
        return oracle.window_filter(
            array,
            size,
            omit_nan=True,
            ignore_inf=bool(kwargs.get("ignore_inf", False)),
            min_count=int(kwargs.get("min_count", 1)),
            **common,
        )
    if function == "rank_filter":
        return oracle.window_filter(array, size, statistic="rank", rank=int(rank) % count, **common)
    raise ValueError(function)


def public_call(nf, function: str, array: np.ndarray, kwargs: dict, rank: int | None = None):
    if function == "rank_filter":
        return nf.rank_filter(array, rank, **kwargs)
    return getattr(nf, function)(array, **kwargs)


def load_kat() -> dict:
    with open(os.path.join(HERE, "golden", "kat.json")) as handle:
        return json.load(handle)


def kat_input(entry: dict) -> np.ndarray:
    dtype = np.dtype(entry["dtype"])
    unsigned = np.uint32 if dtype == np.float32 else np.uint64
    raw = np.array([int(b, 16) for b in entry["input"]], dtype=unsigned)
    return raw.view(dtype).reshape(entry["shape"])


def kat_kwargs(entry: dict) -> dict:
    kwargs = dict(entry["kwargs"])
    for key in ("size", "origin", "mode"):
        if isinstance(kwargs.get(key), list):
            kwargs[key] = tuple(kwargs[key])
    if isinstance(kwargs.get("footprint"), list):
        kwargs["footprint"] = np.asarray(kwargs["footprint"])
    return kwargs


def check_kat(entry: dict, output: np.ndarray) -> None:
    got = bits_of(output).ravel()
    if "expected_all" in entry:
        expected = np.array([int(b, 16) for b in entry["expected_all"]], dtype=got.dtype)
        np.testing.assert_array_equal(got, expected, err_msg=entry["name"])
    else:
        for index, value in entry["expected_at"].items():
            assert int(got[int(index)]) == int(value, 16), (entry["name"], index, hex(int(got[int(index)])), value)


# ---------------------------------------------------------------------------
# Independent brute force (tiny inputs): coordinate loops + Python sorting,
# written from the contract (docs/semantics.md:74-89), sharing no code with
# the C oracle.
# ---------------------------------------------------------------------------


# Planted mistakes for the sensitivity test (tests/test_oracle.py): each name
# switches one plausible error into the brute force below. None = correct.
MUTATION: str | None = None


def _map(index: int, length: int, mode: str) -> int | None:
    if MUTATION == "reflect_as_mirror" and mode == "reflect":
        mode = "mirror"
    if 0 <= index < length:
        return index
    if mode == "constant":
        return None
    if mode == "nearest":
        return min(max(index, 0), length - 1)
    if mode == "wrap":
        return index % length
    if length == 1:
        return 0
    if mode == "reflect":
        folded = index % (2 * length)
        return folded if folded < length else 2 * length - 1 - folded
    folded = index % (2 * length - 2)
    return folded if folded < length else 2 * length - 2 - folded


def _total_order(value: float) -> tuple:
    # -0.0 sorts before +0.0; only non-NaN values reach here.
    return (value, math.copysign(1.0, value))


def _mid(lo: float, hi: float) -> float:
    if MUTATION == "naive_midpoint":
        return (lo + hi) / 2.0
    if MUTATION == "signed_zero_lower" and lo == 0.0 and hi == 0.0:
        return lo
    if lo == hi and math.copysign(1.0, lo) == math.copysign(1.0, hi):
        return lo
    if lo == 0.0 and hi == 0.0:
        return 0.0
    if math.isinf(lo) and math.isinf(hi):
        return math.nan
    if math.isinf(lo) or math.isinf(hi):
        return lo if math.isinf(lo) else hi
    if (lo < 0) == (hi < 0) and math.copysign(1.0, lo) == math.copysign(1.0, hi):
        return lo + (hi - lo) / 2.0
    return (lo + hi) / 2.0


def brute_force(
    array: np.ndarray,
    size: Sequence[int],
    *,
    footprint: np.ndarray | None = None,
    origin: Sequence[int],
    mode: Sequence[str],
    cval: float,
    omit_nan: bool,
    ignore_inf: bool,
    min_count: int,
    rank: int | None = None,
) -> np.ndarray:
    """Median (rank=None) or rank filter by explicit loops over every output."""
    shape = array.shape
    fp = np.ones(size, dtype=bool) if footprint is None else np.asarray(footprint, dtype=bool)
    anchors = [n // 2 + o for n, o in zip(size, origin)]
    taps = [tuple(int(i) - a for i, a in zip(index, anchors)) for index in np.argwhere(fp)]
    out = np.empty_like(array)
    nan = array.dtype.type(np.nan)
    for center in np.ndindex(shape):
        samples = []
        saw_nan = False
        for tap in taps:
            mapped = [_map(c + t, n, m) for c, t, n, m in zip(center, tap, shape, mode)]
            if None in mapped and MUTATION == "cval_not_a_sample":
                continue
            value = float(cval) if None in mapped else float(array[tuple(mapped)])
            if math.isnan(value):
                saw_nan = True
                continue
            if omit_nan and ignore_inf and math.isinf(value):
                if not (MUTATION == "keep_negative_inf" and value < 0):
                    continue
            samples.append(value)
        samples.sort(key=_total_order)
        if (not omit_nan and saw_nan) or (omit_nan and len(samples) < min_count):
            out[center] = nan
        elif rank is not None:
            out[center] = samples[rank] if rank < len(samples) else nan
        elif len(samples) % 2:
            out[center] = samples[len(samples) // 2]
        elif MUTATION == "upper_middle":
            out[center] = samples[len(samples) // 2]
        else:
            m = _mid(samples[len(samples) // 2 - 1], samples[len(samples) // 2])
            if MUTATION == "lower_middle":
                m = samples[len(samples) // 2 - 1]
            out[center] = nan if math.isnan(m) else m
    return out
