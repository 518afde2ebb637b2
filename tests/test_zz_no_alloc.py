"""GPU: the no-allocation launch (ndf_plan_apply_device_static) gives the
bits of the ordinary device launch (ndf_plan_apply_device) on every engine,
including the learned bracket, on two calls in a row with the second on a
smaller input; and it rejects a plan whose tables and scratch exceed its
static device array."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from ndfiltrs_amd import _core, _lib, bracket, synthetic
from tests import _cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


def apply_static(plan: _core.Plan, d_input: int, d_output: int) -> None:
    err = ctypes.create_string_buffer(256)
    status = _lib.lib.ndf_plan_apply_device_static(plan._handle, d_input, d_output, plan.len, None, err, len(err))
    if status != _lib.OK:
        _core._raise(status, err.value.decode())


def make_plan(image, kind, size, mode="reflect", parameter=0, footprint=None, engine="auto", ignore_inf=False):
    ndim = image.ndim
    flat = None
    if footprint is not None:
        size = list(footprint.shape)
        flat = footprint.ravel().tolist()
    plan = _core.Plan(image.dtype, image.shape, kind, size, flat, (0,) * ndim, (mode,) * ndim, 0.0, True,
                      ignore_inf, 1, parameter)
    if engine != "auto":
        plan.set_engine(engine)
    return plan


def both_launches(plan, image, weights=None) -> tuple[np.ndarray, np.ndarray]:
    """(ordinary launch output, no-allocation launch output), each into a
    poisoned buffer."""
    d_in = _core.DeviceArray(image.nbytes)
    d_ref = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    try:
        d_in.upload(image)
        poison = _cases.poison_bits(image.dtype)
        d_ref.fill(poison, image.itemsize)
        d_out.fill(poison, image.itemsize)
        if weights is not None:
            plan.set_bracket(weights)
        plan.apply_device(d_in.address, d_ref.address)
        plan.last_error()
        apply_static(plan, d_in.address, d_out.address)
        _core.synchronize()
        return d_ref.download(np.empty_like(image)), d_out.download(np.empty_like(image))
    finally:
        d_in.free()
        d_ref.free()
        d_out.free()


def ring() -> np.ndarray:
    yy, xx = np.mgrid[-6:7, -6:7]
    r2 = yy * yy + xx * xx
    return (r2 >= 16) & (r2 <= 36)


CASES = [
    # name, (large image, small image), plan arguments, bracket model
    ("bracket_63x63_median", (lambda: synthetic.sky_quadratic_crop((1024, 1024), seed=201),
                              lambda: synthetic.sky_quadratic_crop((384, 640), seed=202)),
     dict(kind="median", size=(63, 63)), "sky63"),
    ("bracket_31x31_f64_rank", (lambda: synthetic.sky_gradient_crop((768, 640), seed=203),
                                lambda: synthetic.sky_gradient_crop((300, 260), seed=204)),
     dict(kind="rank", size=(31, 31), mode="nearest", parameter=154), "sky31_f64"),
    ("slide_63x63_rank", (lambda: synthetic.sky_quadratic_crop((512, 512), seed=205),
                          lambda: synthetic.sky_quadratic_crop((200, 130), seed=206)),
     dict(kind="rank", size=(63, 63), parameter=397), None),
    ("stack_5x5", (lambda: synthetic.sky_small((512, 512), seed=207),
                   lambda: synthetic.sky_small((100, 150), seed=208)),
     dict(kind="median", size=(5, 5)), None),
    ("network_5x5", (lambda: synthetic.sky_small((512, 512), seed=217),
                     lambda: synthetic.sky_small((100, 150), seed=218)),
     dict(kind="median", size=(5, 5), engine="network"), None),
    ("stack_cube", (lambda: synthetic.cube((6, 48, 80), seed=219),
                    lambda: synthetic.cube((3, 20, 24), seed=220)),
     dict(kind="median", size=(3, 7, 7), mode="wrap", ignore_inf=True), None),
    ("network_annulus", (lambda: synthetic.sky_annulus((384, 384), seed=209),
                         lambda: synthetic.sky_annulus((64, 96), seed=210)),
     dict(kind="median", size=None, mode="wrap", footprint=ring(), ignore_inf=True), None),
    ("generic_cube", (lambda: synthetic.cube((6, 48, 80), seed=211),
                      lambda: synthetic.cube((3, 20, 24), seed=212)),
     dict(kind="median", size=(3, 7, 7), mode="wrap", engine="generic", ignore_inf=True), None),
    ("slide_cube", (lambda: synthetic.cube((6, 48, 80), seed=213),
                    lambda: synthetic.cube((3, 20, 24), seed=214)),
     dict(kind="median", size=(3, 7, 7), mode="mirror", engine="slide"), None),
    ("all_nan_rank", (lambda: synthetic.sky_small((64, 64), seed=215),
                      lambda: synthetic.sky_small((16, 8), seed=216)),
     dict(kind="rank", size=(3, 3), parameter=9), None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_static_launch_equals_device_launch(case) -> None:
    name, factories, kwargs, model = case
    weights = bracket.load(bracket.model_path(model))[0] if model else None
    for call, factory in enumerate(factories):
        image = factory()
        print(f"{name}: call {call}, shape {image.shape}", flush=True)
        plan = make_plan(image, **kwargs)
        reference, static = both_launches(plan, image, weights)
        _cases.assert_no_poison(reference, f"{name} device launch")
        _cases.assert_no_poison(static, f"{name} static launch")
        assert np.array_equal(_cases.bits_of(static), _cases.bits_of(reference)), f"{name} call {call}"
        if weights is not None:
            tiles, fallback = plan.bracket_report()
            assert tiles > 0 and 0 <= fallback < tiles


def test_static_launch_rejects_plans_beyond_its_scratch() -> None:
    # 16384^2 float32 with the bracket: 25 MiB of block moments per 8192^2,
    # so about 100 MiB here, above the 32 MiB array.
    shape = (16384, 16384)
    plan = _core.Plan(np.float32, shape, "median", (63, 63), None, (0, 0), ("reflect", "reflect"), 0.0, True, False,
                      1, 0)
    plan.set_bracket(bracket.load(bracket.model_path("sky63"))[0])
    nbytes = shape[0] * shape[1] * 4
    d_in = _core.DeviceArray(nbytes)
    d_out = _core.DeviceArray(nbytes)
    try:
        with pytest.raises(ValueError, match="no-allocation launch needs .* bytes of device scratch"):
            apply_static(plan, d_in.address, d_out.address)
    finally:
        d_in.free()
        d_out.free()
