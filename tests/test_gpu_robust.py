"""GPU parity of the percentile, minimum and maximum statistics and of the
grey-scale morphology built on them.

Every engine that can run a geometry (extrema, network, slide, generic;
forced in turn) gives the oracle's bits on special-value inputs for percentiles 0 to
100 and for the extrema; the public calls give the reference's literal
answers; morphology matches SciPy on finite data, as the reference's tests
check (tests/test_morphology.py:74-147).
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
from scipy import ndimage

import ndfiltrs_amd as nf
from ndfiltrs_amd import _core
from oracle import oracle
from tests import _cases
from tests.test_robust_oracle import check_literal, literal_cases

pytestmark = pytest.mark.gpu

ENGINES = ("extrema", "network", "slide", "generic")


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


def assert_bits(actual, expected, context: str) -> None:
    _cases.assert_no_poison(actual, context)
    a, e = _cases.bits_of(actual), _cases.bits_of(expected)
    bad = np.flatnonzero(a.ravel() != e.ravel())
    assert bad.size == 0, f"{context}: {bad.size} outputs differ, first at flat index {bad[0]}: " \
                          f"got {a.ravel()[bad[0]]:x} expected {e.ravel()[bad[0]]:x}"


def plan_outputs(values, kind, size, parameter, footprint=None, origin=None, mode="reflect", cval=0.0, omit=True,
                 ignore_inf=False, min_count=1):
    """(engine, output) for every engine that can run this geometry."""
    ndim = values.ndim
    flat = None
    if footprint is not None:
        size = list(footprint.shape)
        flat = footprint.ravel().tolist()
    origin = origin or (0,) * ndim
    modes = (mode,) * ndim if isinstance(mode, str) else tuple(mode)
    for engine in ("auto",) + ENGINES:
        plan = _core.Plan(values.dtype, values.shape, kind, size, flat, origin, modes, cval, omit, ignore_inf,
                          min_count, parameter)
        if engine != "auto":
            try:
                plan.set_engine(engine)
            except ValueError:
                continue
        yield engine, plan.apply(values)


GEOMETRIES = [
    ((150,), (8,), (-4,)),
    ((70, 66), (3, 3), (0, 0)),
    ((70, 66), (6, 5), (-3, 2)),
    ((9, 40, 37), (3, 2, 5), (-1, 0, 2)),
    ((70, 66), (11, 11), (2, -3)),
    ((130, 140), (31, 25), (0, 0)),
]
Q_VALUES = (0.0, 12.5, 50.0, 100.0 / 3, 87.5, 100.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize(("shape", "size", "origin"), GEOMETRIES)
def test_percentile_engines_match_oracle(dtype, shape, size, origin) -> None:
    values = _cases.special_sample(shape, dtype, seed=17 * len(shape) + size[0])
    for q, (mode, omit, ignore_inf) in itertools.product(Q_VALUES, [("reflect", True, False), ("wrap", False, False),
                                                                   ("constant", True, True), ("mirror", True, False)]):
        expected = oracle.window_filter(values, size, origin=origin, statistic="percentile", q=q, omit_nan=omit,
                                        ignore_inf=ignore_inf, mode=mode, cval=0.25)
        for engine, out in plan_outputs(values, "percentile", size, q, origin=origin, mode=mode, cval=0.25,
                                        omit=omit, ignore_inf=ignore_inf):
            assert_bits(out, expected, f"percentile {q} {mode} omit={omit} {engine}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize(("shape", "size", "origin"), GEOMETRIES)
def test_extrema_engines_match_oracle(dtype, shape, size, origin) -> None:
    values = _cases.special_sample(shape, dtype, seed=23 * len(shape) + size[0])
    volume = int(np.prod(size))
    for kind, (mode, omit, ignore_inf, min_count) in itertools.product(
            ("min", "max"), [("reflect", False, False, 1), ("constant", True, False, 1), ("wrap", True, True, 2),
                             ("nearest", True, False, volume), ("mirror", True, True, 1)]):
        expected = oracle.window_filter(values, size, origin=origin, statistic=kind, omit_nan=omit,
                                        ignore_inf=ignore_inf, min_count=min_count, mode=mode, cval=-0.0)
        for engine, out in plan_outputs(values, kind, size, 0, origin=origin, mode=mode, cval=-0.0, omit=omit,
                                        ignore_inf=ignore_inf, min_count=min_count):
            assert_bits(out, expected, f"{kind} {mode} omit={omit} min_count={min_count} {engine}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_footprint_percentiles_and_extrema(dtype) -> None:
    values = _cases.special_sample((57, 83), dtype, seed=5, fraction=0.08)
    yy, xx = np.mgrid[-6:7, -6:7]
    ring = ((yy * yy + xx * xx) >= 16) & ((yy * yy + xx * xx) <= 36)
    cross = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool)
    for footprint, origin in ((ring, (1, 0)), (cross, (0, -1))):
        for q in (0.0, 20.0, 50.0, 100.0):
            expected = oracle.window_filter(values, footprint.shape, footprint=footprint, origin=origin,
                                            statistic="percentile", q=q, omit_nan=True, mode="wrap")
            for engine, out in plan_outputs(values, "percentile", None, q, footprint=footprint, origin=origin,
                                            mode="wrap"):
                assert_bits(out, expected, f"footprint percentile {q} {engine}")
        for kind in ("min", "max"):
            expected = oracle.window_filter(values, footprint.shape, footprint=footprint, origin=origin,
                                            statistic=kind, mode="mirror")
            for engine, out in plan_outputs(values, kind, None, 0, footprint=footprint, origin=origin,
                                            mode="mirror", omit=False):
                assert_bits(out, expected, f"footprint {kind} {engine}")


def public_pct(values, q, size, omit, mode, **kwargs):
    values = np.asarray(values, dtype=np.float64) if not isinstance(values, np.ndarray) else values
    function = nf.nanpercentile_filter if omit else nf.percentile_filter
    return function(values, q, size or values.shape, mode=mode, **kwargs)


@pytest.mark.parametrize("case", literal_cases(), ids=[c[0] for c in literal_cases()])
def test_public_percentile_literals(case) -> None:
    _, call, want = case
    check_literal(call(public_pct), want)


def test_public_percentile_mixed_modes() -> None:
    values = np.array([[[1.0, 2.0, 3.0]], [[10.0, 20.0, 30.0]]])
    got = nf.percentile_filter(values, 25, (3, 1, 1), mode=("constant", "nearest", "wrap"), cval=5)
    np.testing.assert_array_equal(got, [[[3, 3.5, 4]], [[3, 3.5, 4]]])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode", _cases.MODES)
@pytest.mark.parametrize("q", [0.0, 12.5, 50.0, 87.5, 100.0])
def test_percentile_matches_scipy_generic_filter(dtype, mode: str, q: float) -> None:
    array = np.array([[3, -4, 8], [1, 12, -2]], dtype=dtype)
    expected = ndimage.generic_filter(array, lambda v: np.percentile(v, q, method="linear"), size=(3, 5), mode=mode,
                                      cval=1.25)
    np.testing.assert_array_equal(nf.percentile_filter(array, q, (3, 5), mode=mode, cval=1.25), expected)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_public_extrema_literals(dtype) -> None:
    for function in (nf.min_filter, nf.nanmin_filter):
        assert np.signbit(function(np.array([0.0, -0.0, 1.0], dtype), 3, mode="wrap")).all()
        np.testing.assert_array_equal(function(np.array([-np.inf, 2, np.inf], dtype), 3, mode="wrap"), -np.inf)
        assert np.signbit(function(np.array([0.0], dtype), mode="constant", cval=-0.0)).all()
    for function in (nf.max_filter, nf.nanmax_filter):
        assert not np.signbit(function(np.array([-0.0, 0.0, -1.0], dtype), 3, mode="wrap")).any()
        np.testing.assert_array_equal(function(np.array([-np.inf, 2, np.inf], dtype), 3, mode="wrap"), np.inf)
    for function in (nf.nanmin_filter, nf.nanmax_filter):
        for invalid in (np.nan, np.inf, -np.inf):
            check_literal(function(np.full(3, invalid, dtype), ignore_inf=True), math.nan)
        check_literal(function(np.array([np.nan], dtype), mode="constant", cval=7, min_count=2), 7.0)
        check_literal(function(np.array([np.nan], dtype), mode="constant", cval=7, min_count=3), math.nan)
        check_literal(function(np.array([7.0], dtype), min_count=3), 7.0)


MORPHOLOGY = [
    (nf.min_filter, ndimage.grey_erosion),
    (nf.grey_dilation, ndimage.grey_dilation),
    (nf.grey_opening, ndimage.grey_opening),
    (nf.grey_closing, ndimage.grey_closing),
    (nf.white_tophat, ndimage.white_tophat),
    (nf.black_tophat, ndimage.black_tophat),
    (nf.morphological_gradient, ndimage.morphological_gradient),
    (nf.morphological_laplace, ndimage.morphological_laplace),
]
WINDOWS = {
    "1d-size-odd": ((17,), {"size": 5}),
    "1d-footprint-even": ((17,), {"footprint": np.array([1, 0, 1, 1])}),
    "2d-size-even": ((9, 10), {"size": (2, 4)}),
    "2d-footprint-odd": ((9, 10), {"footprint": np.array([[0, 1, 0], [1, 1, 1], [0, 0, 1]])}),
    "3d-footprint-mixed": ((5, 6, 7), {"footprint": np.array([[[1, 0, 1], [0, 1, 1]], [[1, 1, 0], [0, 0, 1]]])}),
}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode", _cases.MODES)
@pytest.mark.parametrize("name", WINDOWS)
def test_morphology_matches_scipy(dtype, mode: str, name: str) -> None:
    shape, window = WINDOWS[name]
    values = np.random.default_rng(len(name)).normal(size=shape).astype(dtype)
    lengths = window["footprint"].shape if "footprint" in window else \
        tuple(np.broadcast_to(window["size"], (len(shape),)).tolist())
    origin = tuple((n - 1) // 2 if i % 2 else -(n // 2) for i, n in enumerate(lengths))
    kwargs = {**window, "origin": origin, "mode": mode, "cval": -0.75}
    for function, reference in MORPHOLOGY:
        actual = function(values, **kwargs)
        expected = reference(values, **kwargs)
        assert actual.dtype == dtype
        if function is nf.morphological_laplace:
            # SciPy forms (d + e) - x - x; the drop-in forms d + e - 2x.
            tolerance = 1e-6 if dtype == np.float32 else 1e-14
            np.testing.assert_allclose(actual, expected, rtol=tolerance, atol=tolerance)
        else:
            np.testing.assert_array_equal(actual, expected, err_msg=function.__name__)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["median", "percentile", "min", "max"])
def test_bracket_serves_every_statistic(dtype, kind: str) -> None:
    """The learned bracket gives the plain kernel's bits for percentiles and
    extrema too (a hand-set model; the fallback keeps them exact)."""
    from ndfiltrs_amd import bracket, synthetic

    image = synthetic.sky_quadratic((256, 320), seed=41, sources=8).astype(dtype)
    parameter = 30.0 if kind == "percentile" else 0
    plan = _core.Plan(image.dtype, image.shape, kind, (31, 31), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1,
                      parameter)
    weights = np.zeros(bracket.PARAMS, dtype=np.float32)
    weights[-2:] = (-1.0, 1.0)
    d_in = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    try:
        d_in.upload(image)
        outputs = []
        for model in (None, weights):
            d_out.fill(_cases.poison_bits(dtype), image.itemsize)
            plan.set_bracket(model)
            plan.apply_device(d_in.address, d_out.address)
            plan.last_error()
            outputs.append(d_out.download(np.empty_like(image)))
    finally:
        d_in.free()
        d_out.free()
    statistic = {"median": "median", "percentile": "percentile", "min": "min", "max": "max"}[kind]
    expected = oracle.window_filter(image, (31, 31), statistic=statistic, q=30.0, omit_nan=True, mode="reflect")
    assert_bits(outputs[0], expected, f"{kind} plain")
    assert_bits(outputs[1], expected, f"{kind} bracket")


EXTREMA_CASES = [
    ((1000,), (7,), (0,), np.float32),
    ((257, 300), (63, 63), (0, 0), np.float32),
    ((300, 257), (64, 2), (-32, 0), np.float64),
    ((130, 140), (1, 129), (0, 64), np.float32),
    ((6, 50, 70), (3, 7, 7), (1, 0, -3), np.float32),
    ((5, 40, 33), (5, 4, 1), (-2, 1, 0), np.float64),
    ((2, 3, 700), (2, 3, 5), (0, 1, 2), np.float32),
    # Images narrower than 65 columns take 64-column tiles; a tile above half
    # of the LDS runs 16 waves, which then share the y pass at 4 columns (16
    # rows) per wave (the float64 63x63 case), smaller ones 8 waves.
    ((200, 60), (15, 15), (3, -2), np.float32),
    ((90, 50), (9, 9), (0, 0), np.float64),
    ((150, 60), (63, 63), (2, -5), np.float64),
    ((4, 70, 40), (2, 9, 5), (0, 1, -1), np.float32),
]


@pytest.mark.parametrize(("shape", "size", "origin", "dtype"), EXTREMA_CASES)
def test_extrema_engine_matches_the_separable_oracle(shape, size, origin, dtype) -> None:
    values = _cases.special_sample(shape, dtype, seed=int(np.prod(size)) % 1000, fraction=0.02)
    for kind, mode, omit, ignore_inf, cval in itertools.product(("min", "max"), _cases.MODES, (False, True),
                                                               (False, True), (np.nan, 1.5)):
        if ignore_inf and not omit:
            continue
        expected = oracle.window_filter(values, size, origin=origin, statistic=kind, omit_nan=omit,
                                        ignore_inf=ignore_inf, mode=mode, cval=cval, method="separable")
        plan = _core.Plan(values.dtype, values.shape, kind, size, None, origin, (mode,) * values.ndim, cval, omit,
                          ignore_inf, 1, 0)
        assert '"engine": "extrema"' in plan.describe(), plan.describe()
        assert_bits(plan.apply(values), expected, f"{kind} {mode} omit={omit} ignore_inf={ignore_inf} cval={cval}")
