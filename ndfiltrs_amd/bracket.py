"""Learned bracket proposal: training, model files and use with a plan.

The model is a 6 -> 32 -> 32 -> 2 ReLU MLP (fp32, 1346 parameters) that maps
box-sum features of a bracket tile's support (robust level and noise scale
from 8x8 block means and standard deviations, plane slope, retained and
bright-block fractions, target quantile) to a standardised value bracket for
the tile's window results. Training labels are the smallest and largest
result of each tile computed by the plain GPU kernel on seeded images; the
HIP kernels ``k_mlp_grad`` / ``k_mlp_adam`` fit pinball losses (quantile
``tau_lo`` of the lower end, ``tau_hi`` of the upper end). A bracket never
changes results: tiles whose targets leave it are recomputed by the plain
kernel.

    python -m ndfiltrs_amd.bracket train --workload sky63
    python -m ndfiltrs_amd.bracket train --workload sky31_f64
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _core, _lib, synthetic

PARAMS = 1346
FEATURES = 6
HIDDEN = 32
HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "models")

# Training workloads: the image law (seeded crops of the benchmark field),
# window, boundary mode, dtype and the statistics labelled per image.
WORKLOADS = {
    "sky63": {
        "model": "bracket_sky_63x63.json",
        "dtype": np.float32,
        "generator": synthetic.sky_quadratic_crop,
        "side": 2048,
        "size": 63,
        "mode": "reflect",
        "ranks": [397],
        "tau": (0.003, 0.997),
        # Sort capacity of the bracket kernel (inside samples per pass), per
        # statistic: the 10 % rank's narrow ranges fit 2048 entries, which
        # run six workgroups per CU instead of four.
        "sort_len": {"median": 4096, "rank": 2048, "percentile": 2048},
        "law": "2048^2 crops (seeded offsets) of the 8192^2 field: sky N(1000,10) + quadratic background up to "
               "300 + 2000 sources per 8192^2 + 2% NaN (float32)",
    },
    "sky31_f64": {
        "model": "bracket_sky_31x31_f64.json",
        "dtype": np.float64,
        "generator": synthetic.sky_gradient_crop,
        "side": 1024,
        "size": 31,
        "mode": "nearest",
        "ranks": [154, 806],
        # Tighter pinball quantiles: the 31x31 statistics vary more across a
        # tile, so the 0.3 % tails gave 3.4-sigma brackets holding 40 % of a
        # tile's samples; 2 % tails halve the width (fallback about 1 %).
        "tau": (0.02, 0.98),
        # A 2048-entry sort fits three float64 workgroups per CU; a tile whose
        # range holds more samples splits it into extra passes, which the
        # ranks' narrower ranges rarely need and the median's often would.
        "sort_len": {"median": 4096, "rank": 2048, "percentile": 4096},
        "law": "1024^2 crops (seeded offsets) of the 4096^2 field: sky N(1000,10) + x-gradient 0..200 + 500 "
               "sources per 4096^2 + 0.1% hot pixels U(1e4,1e5) + 2% NaN + the field's 3 NaN columns (float64)",
    },
}
DEFAULT_MODEL = os.path.join(MODELS, WORKLOADS["sky63"]["model"])


def model_path(workload: str) -> str:
    return os.path.join(MODELS, WORKLOADS[workload]["model"])


def init_weights(seed: int) -> np.ndarray:
    """He-initialised weights; the output bias starts at a +-1 sigma bracket."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / FEATURES), size=(FEATURES, HIDDEN))
    w2 = rng.normal(0.0, np.sqrt(2.0 / HIDDEN), size=(HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, 0.1 * np.sqrt(1.0 / HIDDEN), size=(HIDDEN, 2))
    parts = [w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3.ravel(), np.array([-1.0, 1.0])]
    weights = np.concatenate(parts).astype(np.float32)
    assert weights.size == PARAMS
    return weights


def forward(weights: np.ndarray, features: np.ndarray) -> np.ndarray:
    """NumPy forward pass (evaluation only; inference on the device uses HIP)."""
    w = np.asarray(weights, dtype=np.float32)
    o = 0
    w1 = w[o:o + FEATURES * HIDDEN].reshape(FEATURES, HIDDEN)
    o += FEATURES * HIDDEN
    b1 = w[o:o + HIDDEN]
    o += HIDDEN
    w2 = w[o:o + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN)
    o += HIDDEN * HIDDEN
    b2 = w[o:o + HIDDEN]
    o += HIDDEN
    w3 = w[o:o + HIDDEN * 2].reshape(HIDDEN, 2)
    o += HIDDEN * 2
    b3 = w[o:o + 2]
    h1 = np.maximum(features @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    out = h2 @ w3 + b3
    return np.sort(out, axis=1)


def training_image(seed: int, shape=(2048, 2048), workload: str = "sky63") -> np.ndarray:
    """A seeded crop of a benchmark field's law."""
    return WORKLOADS[workload]["generator"](shape=shape, seed=seed)


def statistics_of(workload: str) -> list[tuple[str, int]]:
    return [("median", 0)] + [("rank", r) for r in WORKLOADS[workload]["ranks"]]


def collect(images, size: int, statistics, mode: str = "reflect"):
    """Features/targets/weights of every bracket tile for every (image, statistic)."""
    feats, targs, wts = [], [], []
    for image in images:
        d_in = _core.DeviceArray(image.nbytes)
        d_out = _core.DeviceArray(image.nbytes)
        d_in.upload(image)
        for kind, parameter in statistics:
            plan = _core.Plan(
                image.dtype, image.shape, kind, (size, size), None, (0, 0), (mode, mode), 0.0, True, False, 1,
                parameter,
            )
            plan.apply_device(d_in.address, d_out.address)
            f, t, w = plan.bracket_training_data(d_in.address, d_out.address)
            feats.append(f)
            targs.append(t)
            wts.append(w)
        d_in.free()
        d_out.free()
    return np.concatenate(feats), np.concatenate(targs), np.concatenate(wts)


def fit(weights: np.ndarray, features, targets, weight, steps: int, lr: float, tau_lo: float, tau_hi: float):
    """Full-batch Adam with the HIP training kernels; returns (weights, losses)."""
    w = np.ascontiguousarray(weights, dtype=np.float32).copy()
    x = np.ascontiguousarray(features, dtype=np.float32)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    wt = np.ascontiguousarray(weight, dtype=np.float32)
    losses = np.zeros(steps, dtype=np.float32)
    f32 = ctypes.POINTER(ctypes.c_float)
    err = ctypes.create_string_buffer(256)
    status = _lib.lib.ndf_mlp_train(
        w.ctypes.data_as(f32), w.size, x.ctypes.data_as(f32), t.ctypes.data_as(f32), wt.ctypes.data_as(f32),
        x.shape[0], steps, lr, tau_lo, tau_hi, losses.ctypes.data_as(f32), err, len(err),
    )
    if status != _lib.OK:
        raise RuntimeError(err.value.decode())
    return w, losses


def coverage(weights, features, targets, weight) -> dict:
    pred = forward(weights, features)
    valid = weight > 0
    inside = (pred[:, 0] <= targets[:, 0]) & (pred[:, 1] >= targets[:, 1])
    width = pred[:, 1] - pred[:, 0]
    return {
        "tiles": int(valid.sum()),
        "covered": float(inside[valid].mean()) if valid.any() else 0.0,
        "mean_width_sigma": float(width[valid].mean()) if valid.any() else 0.0,
    }


def save(path: str, weights: np.ndarray, meta: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as handle:
        json.dump({"format": "ndfiltrs-amd-bracket-mlp/1", "meta": meta,
                   "weights": [float(v) for v in np.asarray(weights, dtype=np.float32)]}, handle)
        handle.write("\n")


def sort_len(meta: dict, kind: str = "median") -> int:
    """The bracket kernel's sort capacity recorded with a model for one
    statistic ("median", "rank" or "percentile"; default 4096)."""
    value = meta.get("sort_len", 4096)
    return int(value.get(kind.split("_")[0], 4096) if isinstance(value, dict) else value)


def load(path: str = DEFAULT_MODEL) -> tuple[np.ndarray, dict]:
    with open(path) as handle:
        data = json.load(handle)
    weights = np.asarray(data["weights"], dtype=np.float32)
    if weights.size != PARAMS:
        raise ValueError("bracket model must have 1346 parameters")
    return weights, data.get("meta", {})


def train_main(argv=None) -> int:
    parser = argparse.ArgumentParser(description="train the learned-bracket MLP")
    parser.add_argument("--workload", choices=sorted(WORKLOADS), default="sky63")
    parser.add_argument("--out", default=None)
    parser.add_argument("--images", type=int, default=64)
    parser.add_argument("--side", type=int, default=None)
    parser.add_argument("--steps", type=int, default=3000)
    parser.add_argument("--lr", type=float, default=3e-3)
    parser.add_argument("--tau-lo", type=float, default=None, help="lower pinball quantile (workload default)")
    parser.add_argument("--tau-hi", type=float, default=None, help="upper pinball quantile (workload default)")
    parser.add_argument("--seed", type=int, default=20261203)
    args = parser.parse_args(argv)
    spec = WORKLOADS[args.workload]
    if args.tau_lo is None:
        args.tau_lo = spec["tau"][0]
    if args.tau_hi is None:
        args.tau_hi = spec["tau"][1]
    side = args.side or spec["side"]
    out = args.out or model_path(args.workload)
    statistics = statistics_of(args.workload)
    start = time.time()
    seeds = [args.seed + 1000 + i for i in range(args.images)]
    held = max(args.images // 8, 1)
    train_seeds, hold_seeds = seeds[: args.images - held], seeds[args.images - held:]

    def images(seed_list):
        for seed in seed_list:
            yield training_image(seed, (side, side), args.workload)

    x, t, w = collect(images(train_seeds), spec["size"], statistics, spec["mode"])
    xh, th, wh = collect(images(hold_seeds), spec["size"], statistics, spec["mode"])
    print(f"collected {x.shape[0]} training and {xh.shape[0]} held-out tiles in {time.time() - start:.1f} s",
          file=sys.stderr, flush=True)
    weights = init_weights(args.seed)
    weights, losses = fit(weights, x, t, w, args.steps, args.lr, args.tau_lo, args.tau_hi)
    probe = _core.Plan(spec["dtype"], (side, side), "median", (spec["size"], spec["size"]), None, (0, 0),
                       (spec["mode"], spec["mode"]), 0.0, True, False, 1, 0)
    meta = {
        "workload": args.workload,
        "seed": args.seed,
        "images": args.images,
        "held_out_images": held,
        "image_side": side,
        "window": [spec["size"], spec["size"]],
        "mode": spec["mode"],
        "statistics": [list(s) for s in statistics],
        "law": spec["law"],
        "bracket_tile": [int(json.loads(probe.describe())["bracket_tile_rows"]), 64],
        "steps": args.steps,
        "learning_rate": args.lr,
        "tau": [args.tau_lo, args.tau_hi],
        "sort_len": spec["sort_len"],
        "final_loss": float(losses[-1]),
        "train": coverage(weights, x, t, w),
        "held_out": coverage(weights, xh, th, wh),
        "seconds": time.time() - start,
    }
    save(out, weights, meta)
    print(json.dumps(meta), flush=True)
    return 0


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "train":
        sys.exit(train_main(sys.argv[2:]))
    print(__doc__)
