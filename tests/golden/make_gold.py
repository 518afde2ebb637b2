"""Extract the arrays the parity tests use from the reference's gold scene.

Reads ``tests/data/gold_v1/scene.npz`` of the reference checkout (a data
fixture, loaded with ``allow_pickle=False``), checks every array against the
SHA-256 recorded in the reference's ``manifest.json``, and writes the subset
to ``tests/golden/gold_v1.npz`` together with those hashes
(``tests/golden/gold_v1.json``), so the tests can prove the subset is the
reference's data bit for bit. Run in a checkout that has the reference at
``/root/reference``:

    python tests/golden/make_gold.py
"""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = "/root/reference/tests/data/gold_v1"
# The observed images with their masks (the repair and filter parity inputs),
# the exact controls, and the clean image's three literal test windows
# (tests/test_gold_image.py:162-170 of the reference).
ARRAYS = ("observed_sparse", "mask_sparse", "observed_stress", "mask_stress", "control_flat", "control_ramp",
          "control_mask")
CLEAN_WINDOWS = ((slice(44, 53), slice(53, 62)), (slice(117, 126), slice(155, 164)), (slice(173, 186), slice(38, 51)))


def main() -> None:
    with open(os.path.join(SOURCE, "manifest.json")) as handle:
        manifest = json.load(handle)
    with np.load(os.path.join(SOURCE, "scene.npz"), allow_pickle=False) as archive:
        scene = {name: archive[name] for name in archive.files}
    hashes = {}
    for name, entry in manifest["arrays"].items():
        digest = hashlib.sha256(scene[name].tobytes(order="C")).hexdigest()
        if digest != entry["sha256"]:
            raise SystemExit(f"{name}: SHA-256 {digest} differs from the manifest")
        hashes[name] = digest
    subset = {name: scene[name] for name in ARRAYS}
    for index, (rows, cols) in enumerate(CLEAN_WINDOWS):
        subset[f"clean_window_{index}"] = np.ascontiguousarray(scene["clean"][rows, cols])
    np.savez_compressed(os.path.join(HERE, "gold_v1.npz"), **subset)
    record = {
        "source": "tests/data/gold_v1/scene.npz (scene_id ndfiltrs-gold-v1)",
        "sha256": {name: hashes[name] for name in ARRAYS},
        "clean_windows": [[r.start, r.stop, c.start, c.stop] for r, c in CLEAN_WINDOWS],
        "clean_sha256": hashes["clean"],
    }
    with open(os.path.join(HERE, "gold_v1.json"), "w") as handle:
        json.dump(record, handle, indent=1)
        handle.write("\n")
    print(f"wrote {len(subset)} arrays")


if __name__ == "__main__":
    main()
