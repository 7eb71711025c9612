"""Writes reference_known_answers.json: the expected values that the
reference's own tests hold for this path, transcribed from them.

    tests/test_spatial.cpp:52-70    cell hash of four points, against
                                    cpuCellHash (tests/TestUtils.cpp:30-42),
                                    restated here with numpy FP32
    tests/test_spatial.cpp:102-145  neighbour counts (isolated, pair, 3^3)
    tests/test_pbf.cpp:47-65        free fall after dt = 0.01
    tests/test_pbf.cpp:102-106      isolated-particle lambda (with the
                                    engine's epsilon 800, Engine.cpp:35)
    tests/test_scan.cpp:61-72       all-zeros and all-ones scans
No reference code is imported or run; the values are computed below from
the formulas the reference tests state.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cpu_cell_hash(p, origin=(-3.0, -3.0, -3.0), cell=0.1, dims=(60, 90, 60)):
    inv = np.float32(1.0) / np.float32(cell)
    c = [int(math.floor(np.float32(np.float32(p[a]) - np.float32(origin[a])) * inv)) for a in range(3)]
    c = [min(max(c[a], 0), dims[a] - 1) for a in range(3)]
    return c[0] + c[1] * dims[0] + c[2] * dims[0] * dims[1]


def main():
    pts = [[0.0, 0.0, 0.0, 1.0], [2.5, 4.0, -1.5, 1.0], [-2.9, 5.5, 2.5, 1.0], [1.0, -2.0, -2.5, 1.0]]
    s = float(np.float32(0.1) * np.float32(0.99))
    lattice = [[x * s, y * s, z * s, 1.0] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
    h = 0.1
    w0 = 315.0 / (64.0 * math.pi * h ** 9) * h ** 6
    out = {
        "hash": {"points": pts, "keys": [cpu_cell_hash(p) for p in pts]},
        "neighbors": [
            {"points": [[0, 0, 0, 1], [1.5, 1.5, 1.5, 1], [-1.5, -1.5, -1.5, 1]], "expect": {"0": 0, "1": 0, "2": 0}},
            {"points": [[0, 0, 0, 1], [0.05, 0, 0, 1]], "expect": {"0": 1, "1": 1}},
            {"points": lattice, "expect": {"13": 6}},
        ],
        "free_fall": {"x": [0, 1, 0, 1], "dt": 0.01, "y": 0.99902, "rel": 1e-3},
        "isolated_lambda": {"h": h, "rest_density": 8000.0, "epsilon": 800.0,
                            "lambda": -(w0 / 8000.0 - 1.0) / 800.0},
        "scan": {"zeros_100": [0] * 100, "ones_500": list(range(500))},
    }
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
