"""GPU parity at each benchmark configuration's exact window, boundary mode,
dtype and NaN policy, on images of the configuration's own law at sizes the
CPU oracle finishes in seconds, plus the reference's gold scene.

Every comparison is bit for bit against the oracle, and every device output
buffer starts filled with a poison pattern that no filter writes.
"""

from __future__ import annotations

import numpy as np
import pytest

import ndfiltrs_amd as nf
from ndfiltrs_amd import _core, bracket, synthetic
from oracle import oracle
from tests import _cases
from tests.test_gold import GOLD, RECORD, repair_reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


def assert_bits(actual: np.ndarray, expected: np.ndarray, context: str) -> None:
    _cases.assert_no_poison(actual, context)
    a, e = _cases.bits_of(actual), _cases.bits_of(expected)
    bad = np.flatnonzero(a.ravel() != e.ravel())
    assert bad.size == 0, f"{context}: {bad.size} outputs differ, first at flat index {bad[0]}"


def plan_for(image, kind, size, mode, parameter=0, footprint=None, origin=None, ignore_inf=False, omit=True):
    ndim = image.ndim
    flat = None
    if footprint is not None:
        size = list(footprint.shape)
        flat = footprint.ravel().tolist()
    origin = origin or (0,) * ndim
    return _core.Plan(image.dtype, image.shape, kind, size, flat, origin, (mode,) * ndim, 0.0, omit, ignore_inf, 1,
                      parameter)


def engines_match(image, expected, context, **plan_kwargs) -> None:
    """Every engine that can run the geometry, forced in turn, gives `expected`."""
    for engine in ("auto", "network", "slide", "generic"):
        plan = plan_for(image, **plan_kwargs)
        if engine != "auto":
            try:
                plan.set_engine(engine)
            except ValueError:
                continue
        print(f"{context}: engine {engine}", flush=True)
        assert_bits(plan.apply(image), expected, f"{context} {engine}")


def device_run(plan, image, weights=None, sort_len=4096) -> tuple[np.ndarray, tuple[int, int]]:
    d_in = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    try:
        d_in.upload(image)
        d_out.fill(_cases.poison_bits(image.dtype), image.itemsize)
        plan.set_bracket(weights, sort_len) if weights is not None else plan.set_bracket(None)
        plan.apply_device(d_in.address, d_out.address)
        plan.last_error()
        return d_out.download(np.empty_like(image)), plan.bracket_report()
    finally:
        d_in.free()
        d_out.free()


def test_c1_5x5_reflect_float32() -> None:
    image = synthetic.sky_small((512, 512), seed=101)
    expected = oracle.window_filter(image, (5, 5), mode="reflect", omit_nan=True)
    assert_bits(nf.nanmedian_filter(image, 5, mode="reflect", method="direct"), expected, "C1 public")
    engines_match(image, expected, "C1", kind="median", size=(5, 5), mode="reflect")


@pytest.mark.parametrize("method", ["direct", "histogram"])
def test_c2_31x31_nearest_float64(method: str) -> None:
    image = synthetic.sky_gradient_crop((512, 640), seed=102)
    expected = oracle.window_filter(image, (31, 31), mode="nearest", omit_nan=True, method="histogram")
    assert_bits(nf.nanmedian_filter(image, 31, mode="nearest", method=method), expected, f"C2 median {method}")


@pytest.mark.parametrize("rank", [154, 806])
def test_c2_ranks_float64(rank: int) -> None:
    image = synthetic.sky_gradient_crop((512, 640), seed=103)
    # The NaN-omitting rank of the native boundary (kind 'rank', Omit).
    expected = oracle.window_filter(image, (31, 31), mode="nearest", statistic="rank", rank=rank, omit_nan=True,
                                    method="histogram")
    engines_match(image, expected, f"C2 omit rank {rank}", kind="rank", size=(31, 31), mode="nearest",
                  parameter=rank)
    # The public rank_filter propagates NaN: with 2 % NaN and 3 NaN columns
    # nearly every 961-tap window is NaN.
    expected = oracle.window_filter(image, (31, 31), mode="nearest", statistic="rank", rank=rank,
                                    method="histogram")
    assert_bits(nf.rank_filter(image, rank, 31, mode="nearest"), expected, f"C2 public rank {rank}")


def annulus() -> np.ndarray:
    yy, xx = np.mgrid[-6:7, -6:7]
    r2 = yy * yy + xx * xx
    return (r2 >= 16) & (r2 <= 36)


def test_c3a_annulus_wrap_ignore_inf() -> None:
    image = synthetic.sky_annulus((512, 512), seed=104)
    assert np.isinf(image).any() and np.isnan(image).any()
    ring = annulus()
    expected = oracle.window_filter(image, ring.shape, footprint=ring, origin=(1, 0), mode="wrap", omit_nan=True,
                                    ignore_inf=True)
    got = nf.nanmedian_filter(image, footprint=ring, origin=(1, 0), mode="wrap", ignore_inf=True)
    assert_bits(got, expected, "C3a public")
    engines_match(image, expected, "C3a", kind="median", size=None, mode="wrap", footprint=ring, origin=(1, 0),
                  ignore_inf=True)


def test_c3b_cube_3x7x7_wrap_ignore_inf() -> None:
    image = synthetic.cube((8, 96, 160), seed=105)
    expected = oracle.window_filter(image, (3, 7, 7), mode="wrap", omit_nan=True, ignore_inf=True)
    assert_bits(nf.nanmedian_filter(image, (3, 7, 7), mode="wrap", ignore_inf=True), expected, "C3b public")
    engines_match(image, expected, "C3b", kind="median", size=(3, 7, 7), mode="wrap", ignore_inf=True)


@pytest.mark.parametrize("kind,rank", [("median", 0), ("rank", 397)])
def test_c5_trained_bracket_matches_oracle(kind: str, rank: int) -> None:
    image = synthetic.sky_quadratic_crop((1024, 1024), seed=106)
    weights, meta = bracket.load(bracket.model_path("sky63"))
    expected = oracle.window_filter(image, (63, 63), mode="reflect", statistic=kind, rank=rank, omit_nan=True,
                                    method="histogram")
    plan = plan_for(image, kind, (63, 63), "reflect", rank)
    plain, _ = device_run(plan, image)
    assert_bits(plain, expected, f"C5 {kind} plain")
    out, (tiles, fallback) = device_run(plan, image, weights, bracket.sort_len(meta, kind))
    assert_bits(out, expected, f"C5 {kind} bracket")
    assert tiles > 0 and fallback < tiles


@pytest.mark.parametrize("kind,rank", [("median", 0), ("rank", 154), ("rank", 806)])
def test_c2_trained_bracket_matches_oracle(kind: str, rank: int) -> None:
    image = synthetic.sky_gradient_crop((1024, 1024), seed=107)
    weights, meta = bracket.load(bracket.model_path("sky31_f64"))
    expected = oracle.window_filter(image, (31, 31), mode="nearest", statistic=kind, rank=rank, omit_nan=True,
                                    method="histogram")
    plan = plan_for(image, kind, (31, 31), "nearest", rank)
    out, (tiles, fallback) = device_run(plan, image, weights, bracket.sort_len(meta, kind))
    assert_bits(out, expected, f"C2 {kind} {rank} bracket")
    assert tiles > 0 and fallback < tiles


def test_c5_public_rank_propagates() -> None:
    image = synthetic.sky_quadratic_crop((256, 320), seed=108)
    expected = oracle.window_filter(image, (63, 63), mode="reflect", statistic="rank", rank=397, method="histogram")
    assert_bits(nf.rank_filter(image, 397, 63), expected, "C5 public rank")


# ---- the reference's gold scene -------------------------------------------


@pytest.mark.parametrize("name", ["control_flat", "control_ramp"])
def test_gold_controls(name: str) -> None:
    truth = GOLD[name]
    mask = GOLD["control_mask"]
    corrupted = truth.copy()
    corrupted[mask] = np.nan
    np.testing.assert_array_equal(nf.fixpix(corrupted, mask, method="median"), truth)
    np.testing.assert_array_equal(nf.median_filter(truth)[1:-1, 1:-1], truth[1:-1, 1:-1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ["sparse", "stress"])
def test_gold_repair(dtype, case: str) -> None:
    array = GOLD[f"observed_{case}"].astype(dtype)
    mask = GOLD[f"mask_{case}"]
    got = nf.fixpix(array, mask, method="median", size=3, mode="constant", min_count=3, ignore_inf=True)
    assert_bits(got, repair_reference(array, mask, 3), f"gold repair {case}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("size", [3, 9, 31])
def test_gold_stress_filters(dtype, size: int) -> None:
    """The stress image's defects (hot and dead pixels, NaN, +-inf, a full
    column, runs and a block) through nanmedian and the propagating median."""
    array = GOLD["observed_stress"].astype(dtype)
    for function, omit in (("nanmedian_filter", True), ("median_filter", False)):
        expected = oracle.window_filter(array, (size, size), omit_nan=omit)
        assert_bits(getattr(nf, function)(array, size), expected, f"gold {function} {size}")


def test_gold_literal_windows() -> None:
    for index, (r0, r1, c0, c1) in enumerate(RECORD["clean_windows"]):
        for array, omit in ((GOLD[f"clean_window_{index}"], False),
                            (np.ascontiguousarray(GOLD["observed_stress"][r0:r1, c0:c1]), True)):
            want = _cases.brute_force(array, (3, 3), origin=(0, 0), mode=("reflect", "reflect"), cval=0.0,
                                      omit_nan=omit, ignore_inf=False, min_count=1)
            got = nf.nanmedian_filter(array) if omit else nf.median_filter(array)
            assert_bits(got, want, f"gold window {index}")


@pytest.mark.parametrize("function,kind,extra", [("nanmedian_filter", "median", {}),
                                                 ("nanpercentile_filter", "percentile", {"percentile": 10.0})])
def test_public_call_with_the_shipped_bracket_matches_oracle(function: str, kind: str, extra: dict) -> None:
    """The drop-in call on a 63x63 float32 window of at least 1024^2 pixels
    runs the shipped learned bracket (cached plan) and still gives the
    oracle's bits; a second call with other data reuses the plan."""
    _core.clear_plan_cache()
    for seed in (109, 110):
        image = synthetic.sky_quadratic_crop((1024, 1088), seed=seed)
        if kind == "median":
            got = nf.nanmedian_filter(image, 63)
        else:
            got = nf.nanpercentile_filter(image, extra["percentile"], 63)
        expected = oracle.window_filter(image, (63, 63), mode="reflect", statistic=kind, q=extra.get("percentile", 0),
                                        omit_nan=True, method="histogram")
        assert_bits(got, expected, f"public {function} seed {seed}")
    plans = [plan for entries in _core._PLANS._idle.values() for plan, _ in entries]
    assert len(plans) == 1 and '"bracket": true' in plans[0].describe(), [p.describe() for p in plans]


def test_public_fixpix_reuses_plans_and_matches_oracle() -> None:
    _core.clear_plan_cache()
    rng = np.random.default_rng(111)
    for seed in (112, 113):
        image = synthetic.sky_small((300, 420), seed=seed)
        mask = rng.random(image.shape) < 0.02
        got = nf.fixpix(image, mask, method="median", size=5)
        want = oracle.fixpix_median(image, mask, (5, 5), mode="reflect")
        assert_bits(got, want, f"fixpix seed {seed}")
    # A mask given as an integer array takes the converting path.
    got = nf.fixpix(image, mask.astype(np.int32) * 7, method="median", size=5)
    assert_bits(got, want, "fixpix integer mask")


def test_public_fixpix_63x63_runs_the_shipped_bracket_and_matches_oracle() -> None:
    """fixpix(method="median", size=63) on a float32 image of at least 2^20
    pixels gives its plan the shipped bracket model; the automatic path then
    estimates the dense form (the whole-array launch, bracket included) below
    the sparse one, and the result is still the oracle's bits."""
    _core.clear_plan_cache()
    image = synthetic.sky_quadratic_crop((1024, 1088), seed=114)
    mask = np.random.default_rng(115).random(image.shape) < 0.02
    got = nf.fixpix(image, mask, method="median", size=63)
    want = oracle.fixpix_median(image, mask, (63, 63), mode="reflect")
    assert_bits(got, want, "fixpix 63x63 bracket")
    plans = [plan for entries in _core._PLANS._idle.values() for plan, _ in entries]
    assert len(plans) == 1 and '"bracket": true' in plans[0].describe(), [p.describe() for p in plans]
