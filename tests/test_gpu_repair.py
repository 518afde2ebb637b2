"""GPU parity of fixpix(method="median") against the oracle, bit for bit.

Covers the reference's literal answers, every boundary mode and dtype,
footprints with origins, 3-D data, the NaN/inf policies, both device paths
(sparse positions on the generic engine, dense whole-image engine + scatter)
forced through the C ABI, and unsorted/duplicate positions.
"""

from __future__ import annotations

import numpy as np
import pytest

import ndfiltrs_amd as nf
from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases
from tests.test_repair import KAT, _kat_arrays, _kat_kwargs, bits

pytestmark = pytest.mark.gpu
MODES = ("reflect", "constant", "nearest", "mirror", "wrap")


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


@pytest.mark.parametrize("entry", KAT["fixpix_median"], ids=[c["name"] for c in KAT["fixpix_median"]])
def test_reference_literals(entry: dict) -> None:
    values, mask, expected = _kat_arrays(entry)
    kw = _kat_kwargs(entry, values.ndim)
    out = nf.fixpix(values, mask, method="median", **kw)
    assert np.array_equal(bits(out), expected)


@pytest.mark.parametrize("dtype", (np.float32, np.float64))
@pytest.mark.parametrize("mode", MODES)
def test_modes_and_fractions_match_oracle(dtype, mode: str) -> None:
    rng = np.random.default_rng(101)
    values = _cases.special_sample((97, 131), dtype, seed=17)
    cval = -2.5 if mode == "constant" else None
    for fraction, size in ((0.004, 5), (0.05, 3), (0.3, 9), (0.01, 31), (0.02, 45)):
        mask = rng.random(values.shape) < fraction
        print(f"fixpix {np.dtype(dtype).name} {mode} fraction={fraction} size={size}", flush=True)
        got = nf.fixpix(values, mask, method="median", size=size, mode=mode, cval=cval)
        want = oracle.fixpix_median(values, mask, (size, size), mode=mode,
                                    cval=np.nan if cval is None else cval)
        assert np.array_equal(bits(got), bits(want)), (fraction, size)
        assert np.array_equal(bits(got[~mask]), bits(values[~mask]))


def test_footprint_origin_3d_and_policies() -> None:
    rng = np.random.default_rng(5)
    values = _cases.special_sample((9, 40, 33), np.float32, seed=23)
    values[rng.random(values.shape) < 0.01] = np.inf
    mask = rng.random(values.shape) < 0.1
    fp = rng.random((3, 5, 4)) < 0.5
    fp[1, 2, 2] = True
    count = int(fp.sum())
    for ignore_inf, min_count in ((False, 1), (True, 3), (True, count)):
        print(f"fixpix 3-D footprint ignore_inf={ignore_inf} min_count={min_count}", flush=True)
        got = nf.fixpix(values, mask, method="median", footprint=fp, origin=(0, 1, -1),
                        mode=("wrap", "reflect", "constant"), cval=7.0, ignore_inf=ignore_inf,
                        min_count=min_count)
        want = oracle.fixpix_median(values, mask, fp.shape, footprint=fp, origin=(0, 1, -1),
                                    mode=("wrap", "reflect", "constant"), cval=7.0, ignore_inf=ignore_inf,
                                    min_count=min_count)
        assert np.array_equal(bits(got), bits(want))


def test_input_snapshot_and_ownership() -> None:
    values = np.asarray([1.0, 99.0, 3.0], dtype=np.float64)
    mask = np.array([False, True, False])
    original = values.copy()
    out = nf.fixpix(values, mask, method="median", copy=False)
    assert out is not values and not np.shares_memory(out, values)
    assert np.array_equal(values, original)
    assert out[1] == 2.0
    assert out.flags.owndata and out.flags.c_contiguous
    strided = np.arange(4 * 6 * 8, dtype=np.float32).reshape(4, 6, 8)[::-1, ::2, 1::2]
    smask = np.zeros(strided.shape, dtype=np.int8)
    smask[1, 1, 1] = 3
    got = nf.fixpix(strided, smask, method="median")
    want = oracle.fixpix_median(np.ascontiguousarray(strided), smask != 0, (3, 3, 3))
    assert np.array_equal(bits(got), bits(want))


@pytest.mark.parametrize("dtype", (np.float32, np.float64))
@pytest.mark.parametrize("size", (5, 31, 41))
def test_sparse_and_dense_paths_agree(dtype, size: int) -> None:
    """Both device paths through the C ABI, on unsorted positions with duplicates."""
    rng = np.random.default_rng(size)
    shape = (150, 170)
    values = _cases.special_sample(shape, dtype, seed=size)
    flat = rng.choice(values.size, size=values.size // 40, replace=False)
    positions = np.concatenate([flat, flat[:50]]).astype(np.int64)
    rng.shuffle(positions)
    mask = np.zeros(values.size, dtype=bool)
    mask[flat] = True
    mask = mask.reshape(shape)
    want = oracle.fixpix_median(values, mask, (size, size), mode="mirror")
    plan = _core.Plan(values.dtype, shape, "median", [size, size], None, [0, 0], ["mirror"] * 2, np.nan, True,
                      False, 1, 0)
    nbytes = values.nbytes
    d_in, d_out, d_snap = _core.DeviceArray(nbytes), _core.DeviceArray(nbytes), _core.DeviceArray(nbytes)
    d_pos = _core.DeviceArray(positions.nbytes)
    d_in.upload(values)
    d_pos.upload(positions)
    for path in ("sparse", "dense", "auto"):
        print(f"repair path={path} {np.dtype(dtype).name} size={size}", flush=True)
        d_out.fill(_cases.poison_bits(values.dtype), values.itemsize)
        plan.repair_device(d_in.address, d_out.address, d_snap.address, d_pos.address, positions.size, path)
        got = d_out.download(np.empty_like(values))
        _cases.assert_no_poison(got, path)
        assert np.array_equal(bits(got), bits(want)), path


def test_position_outside_the_array_is_an_error() -> None:
    values = np.ones((16, 16), dtype=np.float32)
    plan = _core.Plan(values.dtype, values.shape, "median", [3, 3], None, [0, 0], ["reflect"] * 2, np.nan, True,
                      False, 1, 0)
    positions = np.array([3, values.size, 5], dtype=np.int64)
    d_in, d_out, d_snap = (_core.DeviceArray(values.nbytes) for _ in range(3))
    d_pos = _core.DeviceArray(positions.nbytes)
    d_in.upload(values)
    d_pos.upload(positions)
    for path in ("sparse", "dense"):
        with pytest.raises(ValueError):
            plan.repair_device(d_in.address, d_out.address, d_snap.address, d_pos.address, positions.size, path)


def _device_repair(plan, values: np.ndarray, positions: np.ndarray, path: str) -> np.ndarray:
    d_in, d_out, d_snap = (_core.DeviceArray(values.nbytes) for _ in range(3))
    d_pos = _core.DeviceArray(positions.nbytes)
    d_in.upload(values)
    d_pos.upload(positions)
    d_out.fill(_cases.poison_bits(values.dtype), values.itemsize)
    plan.repair_device(d_in.address, d_out.address, d_snap.address, d_pos.address, positions.size, path)
    out = d_out.download(np.empty_like(values))
    _cases.assert_no_poison(out, path)
    return out


def test_sparse_thread_fallback_above_the_wave_tap_limit() -> None:
    """65x65 = 4225 taps: the sparse path uses one generic thread per position."""
    rng = np.random.default_rng(65)
    values = _cases.special_sample((70, 90), np.float32, seed=65)
    mask = rng.random(values.shape) < 0.01
    positions = np.flatnonzero(mask).astype(np.int64)
    want = oracle.fixpix_median(values, mask, (65, 65), mode="wrap")
    plan = _core.Plan(values.dtype, values.shape, "median", [65, 65], None, [0, 0], ["wrap"] * 2, np.nan, True,
                      False, 1, 0)
    for path in ("sparse", "dense"):
        print(f"repair 65x65 path={path}", flush=True)
        assert np.array_equal(bits(_device_repair(plan, values, positions, path)), bits(want)), path


@pytest.mark.parametrize("dtype", (np.float32, np.float64))
def test_rank_plan_repair_matches_mask_exclusion(dtype) -> None:
    """The device repair takes any plan: an Omit rank-3 plan on a 7x5 window."""
    rng = np.random.default_rng(3)
    values = _cases.special_sample((60, 75), dtype, seed=3)
    mask = rng.random(values.shape) < 0.05
    positions = np.flatnonzero(mask).astype(np.int64)
    snapshot = values.copy()
    snapshot[mask] = np.nan
    filtered = oracle.window_filter(snapshot, (7, 5), statistic="rank", rank=3, omit_nan=True, mode="nearest",
                                    min_count=2)
    want = values.copy()
    want[mask] = filtered[mask]
    plan = _core.Plan(values.dtype, values.shape, "rank", [7, 5], None, [0, 0], ["nearest"] * 2, 0.0, True, False,
                      2, 3)
    for path in ("sparse", "dense"):
        print(f"repair rank-3 {np.dtype(dtype).name} path={path}", flush=True)
        assert np.array_equal(bits(_device_repair(plan, values, positions, path)), bits(want)), path


@pytest.mark.parametrize("dtype", (np.float32, np.float64))
def test_rank_four_and_five_filters_and_repair(dtype) -> None:
    """Arrays of rank 4 and 5 run the any-rank forms of the generic engine and
    of the sparse repair kernel (run-time coordinate arrays)."""
    rng = np.random.default_rng(45)
    for shape, size, origin, mode in (((4, 5, 6, 7), (3, 1, 3, 2), (0, 0, 1, 0), ("reflect", "wrap", "nearest",
                                                                                   "mirror")),
                                      ((3, 4, 3, 5, 4), (2, 3, 1, 3, 3), (0, -1, 0, 1, 0), "constant")):
        values = _cases.special_sample(shape, dtype, seed=len(shape))
        print(f"rank-{len(shape)} {np.dtype(dtype).name} nanmedian / rank / fixpix", flush=True)
        got = nf.nanmedian_filter(values, size, origin=origin, mode=mode, cval=1.5)
        want = oracle.window_filter(values, size, origin=origin, mode=mode, cval=1.5, omit_nan=True)
        assert np.array_equal(bits(got), bits(want))
        count = int(np.prod(size))
        got = nf.rank_filter(values, count // 3, size, origin=origin, mode=mode, cval=1.5)
        want = oracle.window_filter(values, size, origin=origin, mode=mode, cval=1.5, statistic="rank",
                                    rank=count // 3)
        assert np.array_equal(bits(got), bits(want))
        mask = rng.random(shape) < 0.2
        got = nf.fixpix(values, mask, method="median", size=size, origin=origin, mode=mode, cval=1.5)
        want = oracle.fixpix_median(values, mask, size, origin=origin, mode=mode, cval=1.5)
        assert np.array_equal(bits(got), bits(want))


def test_in_place_repair_and_snapshot_aliasing() -> None:
    rng = np.random.default_rng(9)
    values = _cases.special_sample((80, 96), np.float32, seed=9)
    mask = rng.random(values.shape) < 0.05
    positions = np.flatnonzero(mask).astype(np.int64)
    want = oracle.fixpix_median(values, mask, (5, 5), mode="reflect")
    plan = _core.Plan(values.dtype, values.shape, "median", [5, 5], None, [0, 0], ["reflect"] * 2, np.nan, True,
                      False, 1, 0)
    d_buf, d_snap = _core.DeviceArray(values.nbytes), _core.DeviceArray(values.nbytes)
    d_pos = _core.DeviceArray(positions.nbytes)
    d_pos.upload(positions)
    for path in ("sparse", "dense"):
        print(f"in-place repair path={path}", flush=True)
        d_buf.upload(values)
        plan.repair_device(d_buf.address, d_buf.address, d_snap.address, d_pos.address, positions.size, path)
        assert np.array_equal(bits(d_buf.download(np.empty_like(values))), bits(want)), path
    with pytest.raises(ValueError, match="alias"):
        plan.repair_device(d_buf.address, d_snap.address, d_buf.address, d_pos.address, positions.size)
