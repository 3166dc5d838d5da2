"""Writes tests/golden/reference_kats.json: the known-answer vectors of the
reference's own tests for this path, as data (inputs and expected outputs).

  dot / norm  apps/tests/backends/cuda/cuda_tool_linear_system.cu:26-71
  RLE         apps/tests/common/run_length_encode.cpp:6-34

The values are the literals those tests assert; this script only serialises them.
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))

kats = {
    "dot_norm_f32": [
        {"x": [3.0, 4.0, 12.0], "y": [2.0, -1.0, 0.5], "dot": 8.0, "norm_x": 13.0},
        {"x": [2.0] * 513, "y": [-0.25] * 513, "dot": -256.5, "norm_x": 2.0 * math.sqrt(513.0), "norm_eps": 1e-5},
        {"x": [6.0, 8.0], "norm_x": 10.0},
    ],
    "norm_f64_scaled": {"x": [1.0e200, 3.0, 1.0e-200], "norm": 1.0e200, "eps": 1e-14},
    "empty": {"dot": 0.0, "norm": 0.0},
    "dot_shape_mismatch": {"x": [1.0], "y": [1.0, 2.0]},
    "norm_nonfinite": [{"x": ["inf", 1.0], "expect": "inf"}, {"x": ["nan", 1.0], "expect": "nan"}],
    "rle": [
        {"input": [1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3], "unique": [1, 2, 3], "counts": [4, 3, 4]},
        {"input": [1, 2, 3, 4, 5], "unique": [1, 2, 3, 4, 5], "counts": [1, 1, 1, 1, 1]},
    ],
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
