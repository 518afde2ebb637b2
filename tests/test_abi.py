"""The C-ABI library loads and exports every entry point of include/ndfgpu.h.

Runs without a GPU: symbol lookup and plan construction only (no kernel
launches). Also checks the engine and tile choice for the benchmark shapes.
"""

from __future__ import annotations

import ctypes
import json
import os
import re

import numpy as np
import pytest

from ndfiltrs_amd import _core, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ndfgpu.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ndf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points() -> None:
    names = declared_functions()
    for required in ("ndf_window_filter", "ndf_plan_create", "ndf_plan_apply", "ndf_plan_apply_device",
                     "ndf_plan_destroy", "ndf_get_num_threads"):
        assert required in names


def test_library_exports_every_declared_symbol() -> None:
    lib = ctypes.CDLL(_lib.LIBRARY)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_library_is_built_for_gfx950_only() -> None:
    data = open(_lib.LIBRARY, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx90a", b"gfx942", b"gfx1100"):
        assert b"amdgcn-amd-amdhsa--" + other not in data


def _plan(dtype, shape, size, **kw) -> dict:
    ndim = len(shape)
    plan = _core.Plan(
        dtype, shape, kw.get("kind", "median"), size, kw.get("footprint"), kw.get("origin", [0] * ndim),
        kw.get("mode", ["reflect"] * ndim), 0.0, kw.get("omit", True), kw.get("ignore_inf", False), 1,
        kw.get("parameter", 0),
    )
    return json.loads(plan.describe())


def test_extrema_plans_take_the_extrema_engine() -> None:
    for kind in ("min", "max"):
        info = _plan(np.float32, (8192, 8192), (63, 63), kind=kind)
        assert info["engine"] == "extrema" and info["lds_bytes"] <= 163840, info
        info = _plan(np.float64, (128, 1024, 1024), (3, 7, 7), kind=kind)
        assert info["engine"] == "extrema", info
    # A retained count above one needs a selection engine.
    plan = _core.Plan(np.float32, (512, 512), "min", (5, 5), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 3, 0)
    assert json.loads(plan.describe())["engine"] != "extrema"


def test_engine_choice_for_the_benchmark_shapes() -> None:
    c5 = _plan(np.float32, (8192, 8192), (63, 63))
    assert c5["engine"] == "slide" and c5["tile"][1] == 64 and c5["lds_bytes"] <= 163840
    c2 = _plan(np.float64, (4096, 4096), (31, 31), mode=["nearest"] * 2)
    assert c2["engine"] == "slide" and c2["lds_bytes"] <= 163840
    yy, xx = np.mgrid[-6:7, -6:7]
    ring = (((yy**2 + xx**2) >= 16) & ((yy**2 + xx**2) <= 36)).ravel().tolist()
    c3 = _plan(np.float32, (8192, 8192), (13, 13), footprint=ring, origin=[1, 0], mode=["wrap"] * 2)
    assert c3["engine"] == "network" and c3["count"] == 68 and c3["network_wires"] == 72
    cube = _plan(np.float32, (128, 1024, 1024), (3, 7, 7), mode=["wrap"] * 3)
    assert cube["engine"] == "stack" and cube["count"] == 147 and cube["slices"] == [3, 7, 7]
    assert cube["run"] == 128 and cube["units"] == 1024 * 16
    c1 = _plan(np.float32, (2048, 2048), (5, 5))
    assert c1["engine"] == "stack" and c1["slices"] == [5, 1, 5] and c1["run"] == 10 and c1["units"] == 32 * 205
    five = _plan(np.float64, (2, 2, 2, 2, 3), (3, 3, 3, 3, 5))
    assert five["engine"] == "generic"
    fill = _plan(np.float32, (10, 10), (3, 3), kind="rank", parameter=9)
    assert fill["engine"] == "fill"


def test_plan_rejects_length_mismatch_without_a_device() -> None:
    plan = _core.Plan(np.float32, (4, 4), "median", (3, 3), None, (0, 0), ("reflect",) * 2, 0.0, False, False, 1, 0)
    err = ctypes.create_string_buffer(128)
    buf = np.zeros(15, dtype=np.float32)
    status = _lib.lib.ndf_plan_apply(plan._handle, buf.ctypes.data, buf.ctypes.data, 15, err, len(err))
    assert status == _lib.ERR_LENGTH_MISMATCH
    assert err.value.decode() == "input length does not match the configured shape"


@pytest.mark.parametrize("engine", ["generic", "slide", "network", "stack", "auto"])
def test_engines_can_be_forced(engine: str) -> None:
    plan = _core.Plan(np.float32, (40, 40), "median", (5, 5), None, (0, 0), ("reflect",) * 2, 0.0, False, False, 1, 0)
    plan.set_engine(engine)
    described = json.loads(plan.describe())["engine"]
    assert described == ("stack" if engine == "auto" else engine)


def test_network_engine_limits() -> None:
    big = _core.Plan(np.float32, (40, 40), "median", (13, 13), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1,
                     0)
    assert json.loads(big.describe())["engine"] == "slide"
    with pytest.raises(ValueError, match="network engine needs"):
        big.set_engine("network")
    f64 = _core.Plan(np.float64, (40, 40), "median", (9, 9), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1, 0)
    assert json.loads(f64.describe())["engine"] == "slide"  # 81 taps > 64 for float64
    f32 = _core.Plan(np.float32, (40, 40), "median", (9, 9), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1, 0)
    assert json.loads(f32.describe())["engine"] == "stack"  # a 9x9 box has slice networks
    f32.set_engine("network")
    assert json.loads(f32.describe())["network_wires"] == 96
    for size, wires in (((11, 11), 128), ((12, 12), 160)):
        plan = _core.Plan(np.float32, (40, 40), "median", size, None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1, 0)
        assert json.loads(plan.describe())["network_wires"] == wires


def test_bracket_validation_without_a_device() -> None:
    plan = _core.Plan(np.float32, (256, 256), "median", (63, 63), None, (0, 0), ("reflect",) * 2, 0.0, True,
                      False, 1, 0)
    weights = np.zeros(1346, dtype=np.float32)
    with pytest.raises(ValueError, match="1346 parameters"):
        plan.set_bracket(weights[:-1])
    with pytest.raises(ValueError, match="sort length"):
        plan.set_bracket(weights, 3000)
    plan.set_bracket(weights, 4096)
    plan.set_bracket(None)
    line = _core.Plan(np.float32, (4096,), "median", (63,), None, (0,), ("reflect",), 0.0, True, False, 1, 0)
    with pytest.raises(ValueError, match="2-D image"):
        line.set_bracket(weights)
    cube = _core.Plan(np.float32, (8, 64, 64), "median", (3, 5, 5), None, (0, 0, 0), ("reflect",) * 3, 0.0, True,
                      False, 1, 0)
    with pytest.raises(ValueError, match="2-D image"):
        cube.set_bracket(weights)


def test_overlapping_device_buffers_are_rejected_without_a_device() -> None:
    """The filters need disjoint input and output buffers; the check runs
    before any device call (the addresses here are never dereferenced)."""
    plan = _core.Plan(np.float32, (64, 64), "median", (5, 5), None, (0, 0), ("reflect",) * 2, 0.0, True, False, 1, 0)
    base = 1 << 40
    for d_in, d_out in ((base, base), (base, base + 4 * 64 * 63), (base + 4, base)):
        with pytest.raises(ValueError, match="must not overlap"):
            plan.apply_device(d_in, d_out)


def test_batch_validation_without_a_device() -> None:
    from ndfiltrs_amd import batch

    b = batch.Batch(np.float32, [(10, 12), (0, 4), (30, 7)], [3, [5, 1], 15], ["reflect", "wrap", ["mirror", "nearest"]])
    assert b.len == 10 * 12 + 30 * 7
    assert list(b.offsets) == [0, 120, 120, 330]
    with pytest.raises(ValueError, match="batch image 1: window sizes must be positive"):
        batch.Batch(np.float32, [(10, 12), (5, 5)], [3, 0], ["reflect", "reflect"])
    with pytest.raises(ValueError, match="unsupported boundary mode"):
        batch.Batch(np.float32, [(10, 12)], [3], ["bogus"])
    with pytest.raises(ValueError, match="batch images must be 2-D"):
        batch.Batch(np.float32, [(10,)], [3], ["reflect"])
    with pytest.raises(ValueError, match="sort length"):
        b.configure(3000, 16)
    with pytest.raises(ValueError, match="tile rows"):
        b.configure(2048, 6)
    with pytest.raises(ValueError, match="input length"):
        b.pack([np.zeros((10, 12), np.float32), np.zeros((0, 4), np.float32), np.zeros((30, 8), np.float32)])
