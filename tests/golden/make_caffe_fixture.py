#!/usr/bin/env python3
"""Writes tests/golden/caffe_models.npz: the layer tables of the two trained
nets the reference ships (model_4258.caffemodel, model_4750.caffemodel),
read from their protobuf bytes by leela_amd/caffe_weights.py.

Both are 22-plane policy nets: Convolution, ReLU, BatchNorm per layer and a
last 3x3 Convolution to one channel with no activation (then Flatten and
the loss). Per model `m` and conv `i`: m/i/w (OIHW float32), m/i/b,
m/i/act (1 = ReLU, 2 = none), and, where a BatchNorm follows,
m/i/bn_mean, m/i/bn_var and m/i/bn_scale (the stored blobs). The values
are the files' own float32 values.

Run: python tests/golden/make_caffe_fixture.py [REFERENCE_DIR]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from leela_amd import caffe_weights as cw  # noqa: E402

MODELS = ("model_4258", "model_4750")


def main(ref: Path) -> int:
    arrays = {}
    for m in MODELS:
        table = cw.layer_table(cw.read_caffemodel(str(ref / f"{m}.caffemodel")))
        arrays[f"{m}/n"] = np.int32(len(table))
        for i, L in enumerate(table):
            assert L["kind"] == "conv", (m, L["name"])
            arrays[f"{m}/{i}/w"] = L["w"].astype(np.float32)
            arrays[f"{m}/{i}/b"] = L["b"].astype(np.float32)
            arrays[f"{m}/{i}/act"] = np.int32(L["act"])
            if L["bn"] is not None:
                mean, var, scale = L["bn"]
                arrays[f"{m}/{i}/bn_mean"] = mean.astype(np.float32)
                arrays[f"{m}/{i}/bn_var"] = var.astype(np.float32)
                arrays[f"{m}/{i}/bn_scale"] = np.float32(scale)
    out = Path(__file__).resolve().parent / "caffe_models.npz"
    np.savez_compressed(out, **arrays)
    print(f"wrote {out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(Path(sys.argv[1] if len(sys.argv) > 1
                               else "/root/reference")))
