"""Benchmark: NaN-omitting median 63x63 on an 8192^2 float32 sky (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one nanmedian_filter pass (size 63, reflect) over the
device-resident 8192x8192 float32 image (sky N(1000,10) + quadratic
background + 2000 sources + 2% NaN). Also measured and reported beside it:
the 31x31 float64 nanmedian on 4096^2 (nearest), bit-checked against the CPU
oracle. Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 39.32e12  # 256 CU x 64 lanes x 2.4 GHz, 32-bit lane ops/s
HBM_PEAK = 8.0e12


def log(message: str) -> None:
    print(message, file=sys.stderr, flush=True)


def check_strip(nf_out: np.ndarray, image: np.ndarray, size: int, mode: str, rows: int) -> bool:
    """Bit-compare the first `rows` output rows with the oracle (histogram
    method) run on the rows those outputs read."""
    from oracle import oracle

    half = size // 2
    sub = np.ascontiguousarray(image[: rows + half])
    expected = oracle.window_filter(sub, (size, size), mode=mode, omit_nan=True, method="histogram")
    view = f"u{nf_out.itemsize}"
    return bool(np.array_equal(nf_out[:rows].view(view), expected[:rows].view(view)))


def measure(image: np.ndarray, size: int, mode: str, steps: int, warmup: int, check_rows: int) -> dict:
    from ndfiltrs_amd import _core

    plan = _core.Plan(image.dtype, image.shape, "median", (size, size), None, (0, 0), (mode, mode), 0.0, True,
                      False, 1, 0)
    d_in = _core.DeviceArray(image.nbytes)
    d_out = _core.DeviceArray(image.nbytes)
    d_in.upload(image)
    for _ in range(warmup):
        plan.apply_device(d_in.address, d_out.address)
    _core.synchronize()
    start = time.perf_counter()
    for _ in range(steps):
        plan.apply_device(d_in.address, d_out.address)
    _core.synchronize()
    wall_ms = (time.perf_counter() - start) * 1e3 / steps
    kernel_ms = plan.time_device(d_in.address, d_out.address, 1, steps)
    out = d_out.download(np.empty_like(image))
    exact = check_strip(out, image, size, mode, check_rows)
    pixels = image.size
    return {
        "plan": json.loads(plan.describe()),
        "ms_per_step": wall_ms,
        "kernel_ms": kernel_ms,
        "mpix_s": pixels / (wall_ms * 1e-3) / 1e6,
        "kernel_mpix_s": pixels / (kernel_ms * 1e-3) / 1e6,
        "bit_exact_rows_checked": check_rows,
        "bit_exact": exact,
    }


def cpu_baseline(image: np.ndarray, size: int, mode: str, rows: int) -> dict:
    from oracle import oracle

    threads = oracle.default_threads()
    sample = np.ascontiguousarray(image[:rows])
    start = time.perf_counter()
    oracle.window_filter(sample, (size, size), mode=mode, omit_nan=True, method="histogram", threads=threads)
    seconds = time.perf_counter() - start
    return {
        "value": sample.size / seconds / 1e6,
        "unit": "Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {rows} rows x {image.shape[1]} of the same image, nanmedian {size}x{size} {mode}, "
                  f"oracle block-rank-histogram restatement, {seconds:.2f} s",
    }


def load_traffic() -> dict:
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(path):
        with open(path) as handle:
            return json.load(handle)
    return {}


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the filter shards by row strips "
              "with a halo, described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        return 2
    from ndfiltrs_amd import _core, synthetic

    if _core.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    log("generating 8192x8192 float32 sky (63x63 workload)")
    c5 = synthetic.sky_quadratic()
    log("measuring nanmedian 63x63")
    main_result = measure(c5, 63, "reflect", args.steps, args.warmup, check_rows=256)
    log(json.dumps(main_result))

    log("generating 4096x4096 float64 sky (31x31 workload)")
    c2 = synthetic.sky_gradient_f64()
    log("measuring nanmedian 31x31 float64")
    c2_result = measure(c2, 31, "nearest", args.steps, args.warmup, check_rows=512)
    log(json.dumps(c2_result))

    pixels = c5.size
    count = 63 * 63
    t_kernel = main_result["kernel_ms"] * 1e-3
    valu_achieved = pixels * count / t_kernel / 1e12
    traffic = load_traffic().get("k_slide_f32_63x63_8192")
    roofline = {
        "bound": "valu",
        "achieved": valu_achieved,
        "peak": VALU_PEAK / 1e12,
        "unit": "Tlane-op/s",
        "frac": valu_achieved * 1e12 / VALU_PEAK,
        "traffic": traffic,
        "kernel": "k_slide<float>",
        "algorithmic_units": "63*63 32-bit key compares per output x 67108864 outputs per launch",
        "hbm_view": {
            "achieved": pixels * 8 / t_kernel / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": pixels * 8 / t_kernel / HBM_PEAK,
        },
    }
    baseline = None
    if not args.no_cpu:
        log("timing the CPU baseline")
        baseline = cpu_baseline(c5, 63, "reflect", rows=1024)
    c2_t = c2_result["kernel_ms"] * 1e-3
    line = {
        "metric": "Mpix/s nanmedian 63x63 (plain GPU kernel, bit-exact)",
        "value": main_result["mpix_s"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_result["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sky N(1000,10) + quadratic background + 2000 sources + 2% NaN)",
        "config": {
            "workload": "nanmedian_filter size=63 mode=reflect on 8192x8192 float32, device-resident",
            "engine": main_result["plan"],
            "bit_exact": main_result["bit_exact"],
        },
        "roofline": roofline,
        "cpu_baseline": baseline,
        "nanmedian_31x31_f64": {
            "workload": "nanmedian_filter size=31 mode=nearest on 4096x4096 float64",
            "mpix_s": c2_result["mpix_s"],
            "kernel_ms": c2_result["kernel_ms"],
            "bit_exact": c2_result["bit_exact"],
            "valu_frac": c2.size * 961 * 2 / c2_t / VALU_PEAK,
            "engine": c2_result["plan"],
        },
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
