"""GPU: the learned bracket never changes a result.

Every bracketed launch is compared bit for bit with the plain slide kernel
(and, at oracle-friendly sizes, with the CPU oracle). Hand-set models pin the
fallback logic: a zero-width bracket must send (nearly) every tile to the
plain kernel, an unbounded bracket none, and a NaN model degenerates to an
unbounded bracket. The training kernels are pinned by a closed form: with
constant features the fitted outputs are the tau quantiles of the targets.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from ndfiltrs_amd import _core, bracket, synthetic
from oracle import oracle
from tests import _cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


def constant_model(lo: float, hi: float) -> np.ndarray:
    """A model whose outputs are the standardised offsets (lo, hi) everywhere."""
    weights = np.zeros(bracket.PARAMS, dtype=np.float32)
    weights[-2:] = (lo, hi)
    return weights


def make_plan(image, kind, size, parameter=0, mode="reflect", footprint=None, omit=True, ignore_inf=False,
              min_count=1):
    ndim = image.ndim
    flat = None
    if footprint is not None:
        size = list(footprint.shape)
        flat = footprint.ravel().tolist()
    return _core.Plan(image.dtype, image.shape, kind, size, flat, (0,) * ndim, (mode,) * ndim, 0.0, omit,
                      ignore_inf, min_count, parameter)


def run(plan, image, weights=None, sort_len=4096):
    """One launch into a poisoned output buffer: an output the launch does
    not write keeps the poison (checked here), never an earlier result."""
    d_in = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    d_in.upload(image)
    d_out.fill(_cases.poison_bits(image.dtype), image.itemsize)
    plan.set_bracket(weights, sort_len) if weights is not None else plan.set_bracket(None)
    plan.apply_device(d_in.address, d_out.address)
    plan.last_error()
    out = d_out.download(np.empty_like(image))
    report = plan.bracket_report()
    d_in.free()
    d_out.free()
    _cases.assert_no_poison(out, "bracket" if weights is not None else "plain")
    return out, report


def assert_same_bits(a, b):
    view = np.uint32 if a.dtype == np.float32 else np.uint64
    mismatch = np.flatnonzero(a.view(view).ravel() != b.view(view).ravel())
    assert mismatch.size == 0, f"{mismatch.size} outputs differ, first at flat index {mismatch[0]}"


CASES = [
    # (name, image factory, kind, size, parameter, mode)
    ("c5_median", lambda: synthetic.sky_quadratic((384, 512), seed=11, sources=12), "median", (63, 63), 0, "reflect"),
    ("c5_rank10", lambda: synthetic.sky_quadratic((384, 512), seed=12, sources=12), "rank", (63, 63), 397, "reflect"),
    ("c2_f64", lambda: synthetic.sky_gradient_f64((300, 260), seed=13), "median", (31, 31), 0, "nearest"),
    ("c1_small", lambda: synthetic.sky_small((200, 330), seed=14), "median", (5, 5), 0, "reflect"),
    ("wrap_rank", lambda: synthetic.sky_small((150, 150), seed=15), "rank", (9, 7), 50, "wrap"),
    ("constant_mode", lambda: synthetic.sky_small((130, 140), seed=16), "median", (15, 15), 0, "constant"),
]


@pytest.mark.parametrize("lo,hi", [(-1.0, 1.0), (-0.2, 0.2), (0.0, 0.0), (-3.0, 3.0), (-1e30, 1e30)])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_bracket_bits_equal_plain(case, lo, hi) -> None:
    name, factory, kind, size, parameter, mode = case
    image = factory()
    plan = make_plan(image, kind, size, parameter, mode)
    plain, _ = run(plan, image)
    bracketed, (tiles, fallback) = run(plan, image, constant_model(lo, hi))
    assert_same_bits(bracketed, plain)
    assert tiles > 0 and 0 <= fallback <= tiles
    if lo == hi == 0.0:
        assert fallback >= 0.9 * tiles, (fallback, tiles)
    if hi >= 1e30 and image.dtype == np.float32:
        # Unbounded bracket: only capacity (more samples than the sort) can fail.
        if bracket_support(plan, size) <= 4096:
            assert fallback == 0


def bracket_support(plan, size) -> int:
    """Samples in the support of one bracket tile (64 columns x tile rows)."""
    info = json.loads(plan.describe())
    return (int(info["bracket_tile_rows"]) + size[0] - 1) * (64 + size[1] - 1)


def test_bracket_matches_oracle_small() -> None:
    image = synthetic.sky_quadratic((192, 256), seed=21, sources=6)
    plan = make_plan(image, "median", (63, 63))
    out, (tiles, fallback) = run(plan, image, constant_model(-0.5, 0.5))
    expected = oracle.window_filter(image, size=(63, 63), mode="reflect", statistic="median", omit_nan=True,
                                    method="histogram")
    assert_same_bits(out, expected)
    assert 0 <= fallback <= tiles


@pytest.mark.parametrize("kind,parameter", [("median", 0), ("percentile", 37.5), ("rank", 300)])
def test_split_passes_match_oracle(kind, parameter) -> None:
    """A +-3 sigma bracket holds every sample of a 31x31 support (94 x 62 >
    4096 sort entries): each tile splits its range at the key midpoint and
    runs two passes; median and percentile pairs whose two elements fall on
    either side of the midpoint are combined across the passes. No tile may
    fall back, and every output equals the oracle."""
    image = synthetic.sky_small((192, 256), seed=23)
    plan = make_plan(image, kind, (31, 31), parameter)
    out, (tiles, fallback) = run(plan, image, constant_model(-3.0, 3.0))
    assert bracket_support(plan, (31, 31)) > 4096
    assert fallback == 0, (fallback, tiles)
    expected = oracle.window_filter(image, size=(31, 31), mode="reflect", statistic=kind, omit_nan=True,
                                    rank=parameter if kind == "rank" else 0,
                                    q=parameter if kind == "percentile" else 50.0, method="direct")
    assert_same_bits(out, expected)


@pytest.mark.parametrize("lo,hi", [(0.3, 0.6), (-0.6, -0.3)])
def test_extension_passes_match_oracle(lo, hi) -> None:
    """A bracket entirely above (or below) the medians of a flat N(1000, 10)
    sky: every target lies outside it, so each tile adds adjacent ranges
    (each doubling the key span) until the targets are covered; no tile may
    fall back and every output equals the oracle."""
    rng = np.random.default_rng(24)
    image = rng.normal(1000.0, 10.0, size=(160, 192)).astype(np.float32)
    image[rng.random(image.shape) < 0.02] = np.nan
    plan = make_plan(image, "median", (15, 15))
    out, (tiles, fallback) = run(plan, image, constant_model(lo, hi))
    assert fallback == 0, (fallback, tiles)
    expected = oracle.window_filter(image, size=(15, 15), mode="reflect", statistic="median", omit_nan=True,
                                    method="direct")
    assert_same_bits(out, expected)


EDGE_STATS = [
    # (dtype, kind, parameter, window, (lo, hi))
    (np.float32, "rank", 0, (15, 15), (0.3, 0.6)),           # lowest rank, bracket above it: extensions down
    (np.float32, "rank", 224, (15, 15), (-0.6, -0.3)),       # highest rank, bracket below it: extensions up
    (np.float32, "percentile", 0.0, (15, 15), (-3.0, 3.0)),
    (np.float32, "percentile", 100.0, (15, 15), (-3.0, 3.0)),
    (np.float32, "percentile", 62.5, (31, 31), (-3.0, 3.0)),  # split passes, interpolated pairs
    (np.float64, "median", 0, (31, 31), (-3.0, 3.0)),        # float64 split passes
    (np.float64, "median", 0, (15, 15), (0.3, 0.6)),         # float64 extensions
    (np.float64, "rank", 100, (31, 31), (-0.05, 0.05)),
]


@pytest.mark.parametrize("dtype,kind,parameter,size,bounds", EDGE_STATS,
                         ids=[f"{np.dtype(c[0]).name}-{c[1]}{c[2]}-{c[3][0]}-{c[4][0]}" for c in EDGE_STATS])
def test_passes_on_edge_statistics_match_plain_and_oracle(dtype, kind, parameter, size, bounds) -> None:
    """Extreme ranks and percentiles, float64 and even-count pairs through
    split and extension passes: bit-equal to the plain kernel and the oracle
    (whatever the number of tiles the passes leave to the plain kernel)."""
    rng = np.random.default_rng(31)
    image = rng.normal(1000.0, 10.0, size=(130, 150)).astype(dtype)
    image[rng.random(image.shape) < 0.03] = np.nan
    image[::17, ::13] = np.inf
    plan = make_plan(image, kind, size, parameter)
    plain, _ = run(plan, image)
    out, (tiles, fallback) = run(plan, image, constant_model(*bounds))
    assert_same_bits(out, plain)
    assert 0 <= fallback <= tiles
    expected = oracle.window_filter(image, size=size, mode="reflect", statistic=kind, omit_nan=True,
                                    rank=parameter if kind == "rank" else 0,
                                    q=parameter if kind == "percentile" else 50.0, method="direct")
    assert_same_bits(out, expected)


@pytest.mark.parametrize("cval", [float("nan"), 1e6, 1000.0])
def test_constant_mode_cval_through_passes(cval) -> None:
    """cval is a sample of the border windows (NaN is omitted, 1e6 sits above
    every bracket, 1000 inside it)."""
    rng = np.random.default_rng(32)
    image = rng.normal(1000.0, 10.0, size=(120, 140)).astype(np.float32)
    image[rng.random(image.shape) < 0.02] = np.nan
    plan = _core.Plan(image.dtype, image.shape, "median", (21, 21), None, (0, 0), ("constant",) * 2, cval, True,
                      False, 1, 0)
    plain, _ = run(plan, image)
    out, _ = run(plan, image, constant_model(-0.2, 0.2))
    assert_same_bits(out, plain)
    expected = oracle.window_filter(image, size=(21, 21), mode="constant", cval=cval, statistic="median",
                                    omit_nan=True, method="direct")
    assert_same_bits(out, expected)


RANDOM_SEEDS = list(range(24))


@pytest.mark.parametrize("seed", RANDOM_SEEDS)
def test_random_configurations_bits_equal_plain(seed) -> None:
    """Seeded random geometry, boundary mode, dtype, statistic, policy, sort
    capacity and constant bracket (narrow, offset or wide, so tiles take
    single, split and extension passes): bracketed launches equal the plain
    kernel bit for bit."""
    rng = np.random.default_rng(1000 + seed)
    dtype = np.float32 if rng.random() < 0.6 else np.float64
    ky, kx = int(rng.integers(3, 40)), int(rng.integers(3, 65))
    shape = (int(rng.integers(ky + 20, 200)), int(rng.integers(70, 260)))
    image = rng.normal(1000.0, 10.0, size=shape)
    image += np.linspace(0.0, 50.0 * rng.random(), shape[1])[None, :]
    image[rng.random(shape) < rng.choice([0.0, 0.02, 0.2])] = np.nan
    if rng.random() < 0.5:
        image[rng.random(shape) < 0.005] = np.inf
    image = image.astype(dtype)
    mode = str(rng.choice(["reflect", "constant", "nearest", "mirror", "wrap"]))
    kind = str(rng.choice(["median", "rank", "percentile"]))
    parameter = 0
    if kind == "rank":
        parameter = int(rng.integers(0, ky * kx))
    elif kind == "percentile":
        parameter = float(rng.choice([0.0, 10.0, 37.5, 50.0, 90.0, 100.0]))
    omit = rng.random() < 0.85
    # Propagate has no ignore_inf or min_count (the reference rejects them).
    ignore_inf = omit and bool(rng.random() < 0.5)
    min_count = int(rng.integers(1, max(2, ky * kx // 2))) if omit else 1
    plan = _core.Plan(image.dtype, image.shape, kind, (ky, kx), None, (0, 0), (mode, mode),
                      float(rng.choice([0.0, 1000.0, np.nan])), omit, ignore_inf, min_count, parameter)
    if json.loads(plan.describe())["bracket_tile_rows"] == 0:
        pytest.skip("no bracket geometry for this window")
    plain, _ = run(plan, image)
    centre = float(rng.normal(0.0, 0.5))
    half = float(rng.choice([0.02, 0.3, 1.0, 4.0]))
    sort_len = int(rng.choice([1024, 2048, 4096]))
    out, (tiles, fallback) = run(plan, image, constant_model(centre - half, centre + half), sort_len)
    print(f"seed {seed}: {dtype.__name__} {shape} {ky}x{kx} {mode} {kind}({parameter}) omit={omit} "
          f"min_count={min_count} sort {sort_len} bracket {centre:+.2f}+-{half}: fallback {fallback}/{tiles}",
          flush=True)
    assert_same_bits(out, plain)


def test_nan_model_is_an_unbounded_bracket() -> None:
    image = synthetic.sky_small((160, 192), seed=22)
    plan = make_plan(image, "median", (11, 11))
    plain, _ = run(plan, image)
    weights = constant_model(-1.0, 1.0)
    weights[-1] = np.nan
    out, (tiles, fallback) = run(plan, image, weights)
    assert_same_bits(out, plain)
    assert fallback == 0


@pytest.mark.parametrize("sort_len", [1024, 2048, 8192])
def test_sort_lengths(sort_len) -> None:
    image = synthetic.sky_quadratic((256, 320), seed=23, sources=8)
    plan = make_plan(image, "median", (31, 31))
    plain, _ = run(plan, image)
    out, (tiles, fallback) = run(plan, image, constant_model(-0.8, 0.8), sort_len)
    assert_same_bits(out, plain)


def test_propagate_policy() -> None:
    image = synthetic.sky_small((180, 200), seed=24)
    image[::7, ::5] = np.inf
    for kind, parameter in (("median", 0), ("rank", 60)):
        plan = make_plan(image, kind, (9, 13), parameter, omit=False)
        plain, _ = run(plan, image)
        out, _ = run(plan, image, constant_model(-0.7, 0.7))
        assert_same_bits(out, plain)


def test_ignore_inf_and_mirror() -> None:
    image = synthetic.sky_small((170, 230), seed=29)
    image[::11, ::3] = np.inf
    image[5::13, ::4] = -np.inf
    plan = make_plan(image, "median", (17, 15), mode="mirror", ignore_inf=True)
    plain, _ = run(plan, image)
    out, _ = run(plan, image, constant_model(-0.9, 0.9))
    assert_same_bits(out, plain)


def test_footprint_is_rejected() -> None:
    image = synthetic.sky_small((64, 64), seed=30)
    fp = np.ones((5, 5), dtype=bool)
    fp[0, 0] = False
    plan = make_plan(image, "median", None, footprint=fp)
    with pytest.raises(ValueError, match="rectangular window"):
        plan.set_bracket(constant_model(-1.0, 1.0))


def test_all_nan_tiles_and_min_count() -> None:
    image = synthetic.sky_small((200, 256), seed=25)
    image[:, 64:140] = np.nan
    plan = make_plan(image, "median", (21, 21), min_count=200)
    plain, _ = run(plan, image)
    out, _ = run(plan, image, constant_model(-1.0, 1.0))
    assert_same_bits(out, plain)


def test_trained_model_if_present() -> None:
    if not os.path.exists(bracket.DEFAULT_MODEL):
        pytest.skip("no trained bracket model in the tree")
    weights, _ = bracket.load()
    image = synthetic.sky_quadratic((512, 640), seed=26, sources=20)
    for kind, parameter in (("median", 0), ("rank", 397)):
        plan = make_plan(image, kind, (63, 63), parameter)
        plain, _ = run(plan, image)
        out, (tiles, fallback) = run(plan, image, weights)
        assert_same_bits(out, plain)
        assert fallback <= tiles


def test_training_recovers_target_quantiles() -> None:
    """Constant features: the pinball optimum is the tau quantile of each target."""
    rng = np.random.default_rng(27)
    count = 20000
    features = np.zeros((count, bracket.FEATURES), dtype=np.float32)
    features[:, 3] = 0.5
    lower = rng.normal(-2.0, 0.5, size=count).astype(np.float32)
    upper = rng.normal(2.0, 0.25, size=count).astype(np.float32)
    targets = np.stack([lower, upper], axis=1)
    weight = np.ones(count, dtype=np.float32)
    weight[:100] = 0.0  # ignored samples
    weights, losses = bracket.fit(bracket.init_weights(1), features, targets, weight, 1500, 1e-2, 0.05, 0.95)
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    pred = bracket.forward(weights, features[:1])[0]
    q_lo = np.quantile(lower[100:], 0.05)
    q_hi = np.quantile(upper[100:], 0.95)
    assert abs(pred[0] - q_lo) < 0.05, (pred, q_lo)
    assert abs(pred[1] - q_hi) < 0.05, (pred, q_hi)


def test_training_data_shapes_and_labels() -> None:
    image = synthetic.sky_quadratic((256, 256), seed=28, sources=4)
    plan = make_plan(image, "median", (63, 63))
    d_in = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    d_in.upload(image)
    plan.apply_device(d_in.address, d_out.address)
    features, targets, weight = plan.bracket_training_data(d_in.address, d_out.address)
    out = d_out.download(np.empty_like(image))
    d_in.free()
    d_out.free()
    assert features.shape[1] == 6 and targets.shape == (features.shape[0], 2)
    assert np.all(np.isfinite(features)) and np.all(weight == 1.0)
    assert np.all(targets[:, 0] <= targets[:, 1])
    # The tile labels bracket the plain outputs of their tile by construction:
    # check the global extremes (standardisation is monotone per tile).
    assert np.isfinite(out).any()
