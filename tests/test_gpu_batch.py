"""GPU: batched images (per-image window and mode) against single-image plans.

Every batch result is compared bit for bit with one ndf_plan per image (and,
for a few images, with the CPU oracle); the captured graph is replayed with
eager work and input changes in between, and every replay is checked.
"""

from __future__ import annotations

import numpy as np
import pytest

from ndfiltrs_amd import _core, batch, synthetic
from oracle import oracle
from tests import _cases

pytestmark = pytest.mark.gpu

MODES = ("reflect", "constant", "nearest", "mirror", "wrap")


@pytest.fixture(scope="module", autouse=True)
def device() -> None:
    assert _core.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"


def make_images(count: int, seed: int, dtype=np.float32, low: int = 9, high: int = 300):
    rng = np.random.default_rng(seed)
    images, sizes, modes = [], [], []
    for i in range(count):
        rows, cols = (int(v) for v in rng.integers(low, high, size=2))
        image = rng.normal(100.0, 5.0, size=(rows, cols))
        image[rng.random((rows, cols)) < 0.02] = np.nan
        image[rng.random((rows, cols)) < 0.01] = np.inf
        images.append(image.astype(dtype))
        sizes.append([int(rng.choice(np.arange(1, 16, 2))), int(rng.choice(np.arange(1, 16, 2)))])
        modes.append([MODES[int(rng.integers(0, 5))], MODES[int(rng.integers(0, 5))]])
    return images, sizes, modes


def single(image, size, mode, kind="median", omit=True, ignore_inf=False, parameter=0, cval=0.0):
    plan = _core.Plan(image.dtype, image.shape, kind, size, None, (0, 0), mode, cval, omit, ignore_inf, 1,
                      parameter)
    return plan.apply(image)


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("sort_len,rows,graph", [(2048, 16, True), (2048, 4, False), (4096, 64, True),
                                                 (16384, 8, True)])
def test_batch_matches_single_plans(sort_len, rows, graph) -> None:
    images, sizes, modes = make_images(24, seed=sort_len + rows)
    b = batch.Batch(np.float32, [i.shape for i in images], sizes, modes, ignore_inf=True, cval=3.5)
    b.configure(sort_len, rows, graph)
    outs = b.apply(images)
    for image, size, mode, out in zip(images, sizes, modes, outs):
        expected = single(image, size, mode, ignore_inf=True, cval=3.5)
        np.testing.assert_array_equal(bits(out), bits(expected))


def test_batch_matches_oracle() -> None:
    images, sizes, modes = make_images(8, seed=3, low=5, high=80)
    b = batch.Batch(np.float32, [i.shape for i in images], sizes, modes)
    outs = b.apply(images)
    for image, size, mode, out in zip(images, sizes, modes, outs):
        expected = oracle.window_filter(image, size, mode=mode, omit_nan=True)
        np.testing.assert_array_equal(bits(out), bits(expected))


def test_float64_rank_propagate_and_all_nan_images() -> None:
    images, sizes, modes = make_images(12, seed=4, dtype=np.float64)
    # rank 40 is >= the tap count of small windows: those images are all NaN.
    b = batch.Batch(np.float64, [i.shape for i in images], sizes, modes, kind="rank", omit_nan=False,
                    parameter=40)
    outs = b.apply(images)
    for image, size, mode, out in zip(images, sizes, modes, outs):
        expected = single(image, size, mode, kind="rank", omit=False, parameter=40)
        np.testing.assert_array_equal(bits(out), bits(expected))


def test_empty_image_in_batch() -> None:
    images = [np.ones((0, 7), np.float32), np.arange(30, dtype=np.float32).reshape(5, 6)]
    b = batch.Batch(np.float32, [i.shape for i in images], [3, 3], ["reflect", "wrap"])
    outs = b.apply(images)
    assert outs[0].shape == (0, 7)
    np.testing.assert_array_equal(bits(outs[1]), bits(single(images[1], [3, 3], ["wrap", "wrap"])))


def test_graph_replays_follow_input_changes() -> None:
    images, sizes, modes = make_images(16, seed=5)
    b = batch.Batch(np.float32, [i.shape for i in images], sizes, modes)
    b.configure(2048, 16, True)
    flat = b.pack(images)
    d_in = _core.DeviceArray(flat.nbytes)
    d_out = _core.DeviceArray(flat.nbytes)
    rng = np.random.default_rng(6)
    try:
        for replay in range(4):
            # New contents in the same buffers, eager single-image work in
            # between: every replay must see the new input.
            flat = flat.copy()
            flat[rng.integers(0, flat.size, size=flat.size // 50)] = rng.normal(100.0, 50.0, size=flat.size // 50)
            d_in.upload(flat)
            single(images[0], sizes[0], modes[0])
            d_out.fill(_cases.POISON32, 4)
            b.apply_device(d_in.address, d_out.address)
            raw = d_out.download(np.empty_like(flat))
            _cases.assert_no_poison(raw, f"replay {replay}")
            got = b.unpack(raw)
            for i in (0, 7, 15):
                expected = single(b.unpack(flat)[i], sizes[i], modes[i])
                np.testing.assert_array_equal(bits(got[i]), bits(expected), err_msg=f"replay {replay} image {i}")
    finally:
        d_in.free()
        d_out.free()


def test_autotune_keeps_results() -> None:
    images, sizes, modes = synthetic.cutouts(count=40, seed=7)
    b = batch.Batch(np.float32, [i.shape for i in images], sizes, modes)
    flat = b.pack(images)
    d_in = _core.DeviceArray(flat.nbytes)
    d_out = _core.DeviceArray(flat.nbytes)
    try:
        d_in.upload(flat)
        d_out.fill(_cases.POISON32, 4)
        b.apply_device(d_in.address, d_out.address)
        before = d_out.download(np.empty_like(flat))
        _cases.assert_no_poison(before, "before autotune")
        info = b.autotune(d_in.address, d_out.address, repeats=1)
        assert info["tuned"] and info["ok"]
        d_out.fill(_cases.POISON32, 4)
        b.apply_device(d_in.address, d_out.address)
        after = d_out.download(np.empty_like(flat))
        _cases.assert_no_poison(after, "after autotune")
        np.testing.assert_array_equal(bits(after), bits(before))
        for i in (0, 13, 39):
            expected = single(images[i], [sizes[i]] * 2, [modes[i]] * 2)
            np.testing.assert_array_equal(bits(b.unpack(after)[i]), bits(expected))
    finally:
        d_in.free()
        d_out.free()


def test_configuration_too_small_for_a_window_is_rejected() -> None:
    image = np.ones((40, 40), np.float32)
    b = batch.Batch(np.float32, [image.shape], [15], ["reflect"])
    b.configure(1024, 4, True)  # 18 support rows x 78 columns > 1024
    with pytest.raises(ValueError, match="does not fit a 1024-entry tile sort"):
        b.apply([image])
