"""fixpix(method="median"): oracle pinning and host-side validation (CPU).

The oracle's mask-directed repair (``ndfo_fixpix_median``) is pinned three
ways: the reference's literal answers (``tests/golden/kat.json``
``fixpix_median``), the reference tests' own mask-exclusion identity
(``tests/test_fixpix_methods.py:13-38``: a full NaN-omitting median of the
data with masked pixels set to NaN, kept only at masked pixels), computed
here by the oracle's other two algorithms (direct over every output and the
block rank histogram), and a numpy brute force on tiny inputs.
"""

from __future__ import annotations

import numpy as np
import pytest

import ndfiltrs_amd as nf
from oracle import oracle
from tests import _cases

KAT = _cases.load_kat()
MODES = ("reflect", "constant", "nearest", "mirror", "wrap")


def _kat_arrays(entry: dict) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    dtype = np.dtype(entry["dtype"])
    u = np.uint32 if dtype == np.float32 else np.uint64
    shape = tuple(entry["shape"])
    values = np.array([int(b, 16) for b in entry["input"]], dtype=u).view(dtype).reshape(shape)
    expected = np.array([int(b, 16) for b in entry["expected_all"]], dtype=u).reshape(shape)
    mask = np.array(entry["mask"], dtype=bool).reshape(shape)
    return values, mask, expected


def _kat_kwargs(entry: dict, ndim: int) -> dict:
    kw = dict(entry["kwargs"])
    size = kw.pop("size")
    kw["size"] = (size,) * ndim if isinstance(size, int) else tuple(size)
    return kw


def mask_exclusion(values, mask, size, method="direct", **kw):
    snapshot = np.array(values, copy=True)
    snapshot[mask] = np.nan
    filtered = oracle.window_filter(snapshot, size, omit_nan=True, method=method, **kw)
    out = np.array(values, copy=True)
    out[mask] = filtered[mask]
    return out


def bits(a):
    return _cases.bits_of(np.ascontiguousarray(a))


@pytest.mark.parametrize("entry", KAT["fixpix_median"], ids=[c["name"] for c in KAT["fixpix_median"]])
def test_oracle_matches_reference_literals(entry: dict) -> None:
    values, mask, expected = _kat_arrays(entry)
    kw = _kat_kwargs(entry, values.ndim)
    size = kw.pop("size")
    out = oracle.fixpix_median(values, mask, size, **kw)
    assert np.array_equal(bits(out), expected)


@pytest.mark.parametrize("dtype", (np.float32, np.float64))
@pytest.mark.parametrize("mode", MODES)
def test_oracle_equals_mask_exclusion_identity(dtype, mode: str) -> None:
    rng = np.random.default_rng(7)
    values = _cases.special_sample((23, 17), dtype, seed=11)
    mask = rng.random(values.shape) < 0.2
    cval = 4.0 if mode == "constant" else np.nan
    for size, origin in (((3, 3), (0, 0)), ((5, 4), (1, -1)), ((1, 7), (0, 2))):
        got = oracle.fixpix_median(values, mask, size, mode=mode, cval=cval, origin=origin, threads=3)
        for method in ("direct", "histogram"):
            want = mask_exclusion(values, mask, size, method=method, mode=mode, cval=cval, origin=origin)
            assert np.array_equal(bits(got), bits(want)), (size, origin, method)


def test_oracle_footprint_3d_and_policies() -> None:
    rng = np.random.default_rng(3)
    values = _cases.special_sample((6, 9, 8), np.float32, seed=5)
    mask = rng.random(values.shape) < 0.3
    fp = rng.random((3, 2, 3)) < 0.6
    fp[1, 1, 1] = True
    count = int(fp.sum())
    for ignore_inf, min_count in ((False, 1), (True, 2), (True, count)):
        got = oracle.fixpix_median(values, mask, fp.shape, footprint=fp, mode=("wrap", "mirror", "nearest"),
                                   ignore_inf=ignore_inf, min_count=min_count)
        want = mask_exclusion(values, mask, fp.shape, footprint=fp, mode=("wrap", "mirror", "nearest"),
                              cval=np.nan, ignore_inf=ignore_inf, min_count=min_count)
        assert np.array_equal(bits(got), bits(want))


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_numpy_brute_force(seed: int) -> None:
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 9))
    values = rng.integers(-3, 4, size=n).astype(np.float64)
    values[rng.random(n) < 0.2] = np.nan
    mask = rng.random(n) < 0.5
    k = int(rng.integers(1, 6))
    snapshot = values.copy()
    snapshot[mask] = np.nan
    half = k // 2
    want = values.copy()
    for i in np.flatnonzero(mask):
        window = [snapshot[min(max(i + d - half, 0), n - 1)] for d in range(k)]  # nearest mode
        kept = np.sort([v for v in window if not np.isnan(v)])
        if kept.size == 0:
            want[i] = np.nan
        elif kept.size % 2:
            want[i] = kept[kept.size // 2]
        else:
            lo, hi = kept[kept.size // 2 - 1], kept[kept.size // 2]
            want[i] = lo + (hi - lo) / 2
    got = oracle.fixpix_median(values, mask, (k,), mode="nearest")
    assert np.array_equal(got, want, equal_nan=True)


def test_oracle_rejects_mask_length_mismatch() -> None:
    with pytest.raises(oracle.OracleError) as info:
        oracle.fixpix_median(np.zeros(4), np.zeros(3, dtype=bool), (3,))
    assert info.value.status == 3


@pytest.mark.parametrize(
    ("kwargs", "error"),
    [
        ({"method": "unknown"}, ValueError),
        ({"method": 1}, (TypeError, ValueError)),
        ({"method": "linear", "size": 3}, ValueError),
        ({"method": "linear", "mode": "reflect"}, ValueError),
        ({"method": "linear", "cval": 0.0}, ValueError),
        ({"method": "linear", "ignore_inf": False}, ValueError),
        ({"method": "linear", "min_count": 1}, ValueError),
        ({"method": "median", "size": 0}, ValueError),
        ({"method": "median", "size": True}, TypeError),
        ({"method": "median", "size": (3, 3, 3)}, ValueError),
        ({"method": "median", "mode": "bad"}, ValueError),
        ({"method": "median", "mode": ("reflect",)}, ValueError),
        ({"method": "median", "mode": 1}, TypeError),
        ({"method": "median", "cval": True}, TypeError),
        ({"method": "median", "cval": 1 + 2j}, TypeError),
        ({"method": "median", "ignore_inf": 1}, TypeError),
        ({"method": "median", "min_count": 0}, ValueError),
        ({"method": "median", "min_count": 10}, ValueError),
        ({"method": "median", "min_count": True}, TypeError),
        ({"method": "median", "min_count": 1.5}, TypeError),
        ({"method": "median", "copy": 1}, TypeError),
    ],
)
def test_configuration_is_validated_before_empty_noop(kwargs, error) -> None:
    """tests/test_fixpix_methods.py:353-386."""
    values = np.empty((0, 2), dtype=np.float64)
    for mask in (None, np.empty(values.shape, dtype=bool)):
        with pytest.raises(error):
            nf.fixpix(values, mask, **kwargs)


def test_priority_and_window_volume_checks_without_mask() -> None:
    """tests/test_fixpix_methods.py:442-447."""
    values = np.ones((2, 2))
    with pytest.raises(ValueError, match="priority"):
        nf.fixpix(values, method="median", priority=(1, 0))
    with pytest.raises(OverflowError, match="window volume"):
        nf.fixpix(values, method="median", size=(2**63 + 1, 3), copy=False)


def test_mask_checks_and_noop_ownership() -> None:
    values = np.arange(24.0).reshape(4, 6)[::-1, ::2]
    with pytest.raises(ValueError, match="same shape"):
        nf.fixpix(values, np.zeros((2, 2)), method="median")
    with pytest.raises(TypeError, match="masked arrays"):
        nf.fixpix(values, np.ma.masked_array(np.zeros(values.shape)), method="median")
    with pytest.raises(TypeError, match="boolean or numeric"):
        nf.fixpix(values, np.zeros(values.shape, dtype="U1"), method="median")
    values.setflags(write=False)
    assert nf.fixpix(values, np.zeros(values.shape), method="median", copy=False) is values
    copied = nf.fixpix(values, None, method="median")
    assert copied.flags.owndata and copied.flags.writeable and copied.flags.c_contiguous
    assert np.array_equal(copied, values)


def test_linear_and_mean_are_not_in_this_build() -> None:
    values = np.ones(3)
    mask = np.array([False, True, False])
    with pytest.raises(NotImplementedError):
        nf.fixpix(values, mask)
    with pytest.raises(NotImplementedError):
        nf.fixpix(values, mask, kernel=np.ones(3))


def brute_fixpix(values, mask, size, *, mode="reflect", cval=np.nan, ignore_inf=False, min_count=1, origin=None,
                 mutation=None):
    """Loops over every masked pixel (src/filter_nd.rs:174-215 read literally):
    samples come from the data with every masked pixel set to NaN; NaN is
    dropped, infinities too under ignore_inf; fewer than min_count retained
    samples give NaN; the median is the middle sample or the contract
    midpoint, in float64, cast once. ``mutation`` plants one mistake."""
    shape = values.shape
    size = (size,) * values.ndim if isinstance(size, int) else tuple(size)
    modes = (mode,) * values.ndim if isinstance(mode, str) else tuple(mode)
    origin = (0,) * values.ndim if origin is None else tuple(origin)
    anchors = [n // 2 + o for n, o in zip(size, origin)]
    taps = [tuple(int(i) - a for i, a in zip(index, anchors)) for index in np.ndindex(size)]
    live = values.astype(np.float64)
    if mutation != "no_snapshot":
        live[mask] = np.nan
    out = values.copy()
    for center in np.ndindex(shape):
        if not mask[center]:
            continue
        samples = []
        for tap in taps:
            mapped = [_cases._map(c + t, n, m) for c, t, n, m in zip(center, tap, shape, modes)]
            if None in mapped:
                if mutation == "cval_dropped":
                    continue
                value = float(values.dtype.type(cval))
            else:
                value = float(live[tuple(mapped)])
            if np.isnan(value) or (ignore_inf and np.isinf(value) and mutation != "inf_kept"):
                continue
            samples.append(value)
        samples.sort(key=_cases._total_order)
        enough = len(samples) >= min_count or mutation == "min_count_ignored"
        if not samples or not enough:
            result = np.nan
        elif len(samples) % 2:
            result = samples[len(samples) // 2]
        else:
            result = _cases._mid(samples[len(samples) // 2 - 1], samples[len(samples) // 2])
        out[center] = result
        if mutation == "sequential":
            live[center] = result
    return out


def test_brute_force_passes_every_fixpix_golden_case() -> None:
    for entry in KAT["fixpix_median"]:
        values, mask, expected = _kat_arrays(entry)
        out = brute_fixpix(values, mask, **entry["kwargs"])
        assert np.array_equal(bits(out), expected), entry["name"]


@pytest.mark.parametrize("mutation", ["no_snapshot", "sequential", "cval_dropped", "inf_kept",
                                      "min_count_ignored"])
def test_planted_repair_mistakes_fail_a_golden_case(mutation: str) -> None:
    """Each plausible mistake in a repair restatement fails at least one of
    the reference's literal answers, so those answers pin the oracle."""
    failed = []
    for entry in KAT["fixpix_median"]:
        values, mask, expected = _kat_arrays(entry)
        out = brute_fixpix(values, mask, mutation=mutation, **entry["kwargs"])
        if not np.array_equal(bits(out), expected):
            failed.append(entry["name"])
    assert failed, f"no golden case catches {mutation}"


@pytest.mark.parametrize("mode", MODES)
def test_oracle_matches_brute_force_on_small_arrays(mode: str) -> None:
    rng = np.random.default_rng(MODES.index(mode))
    for dtype in (np.float32, np.float64):
        values = _cases.special_sample((9, 11), dtype, seed=MODES.index(mode) + 7)
        mask = rng.random(values.shape) < 0.3
        for size, origin, policy in (((3, 3), (0, 0), (False, 1)), ((4, 5), (1, -1), (True, 3)),
                                     ((1, 6), (0, 2), (True, 1))):
            ignore_inf, min_count = policy
            got = oracle.fixpix_median(values, mask, size, mode=mode, cval=2.0, origin=origin,
                                       ignore_inf=ignore_inf, min_count=min_count)
            want = brute_fixpix(values, mask, size, mode=mode, cval=2.0, origin=origin, ignore_inf=ignore_inf,
                                min_count=min_count)
            assert np.array_equal(bits(got), bits(want)), (dtype, size, origin, policy)
