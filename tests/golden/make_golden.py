"""Writes the small golden fixtures in this directory from the C oracle.

Inputs are seeded and hand-picked (the reference's own test Labs and
pixels where it has them); outputs are the oracle's.  They pin the oracle
against regressions and give the GPU tests oracle-free expected values.
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32).tolist()


def main():
    oracle.load()
    rng = np.random.default_rng(2024)
    out = {}

    # placement: 6x5 grid, 12 materials, usage 2 -> 24 < 30 cells forces resets
    lab = np.round(rng.uniform(-30, 30, (12, 3)), 2).astype(np.float32)
    cl = np.round(rng.uniform(-30, 30, (30, 3)), 2).astype(np.float32)
    r = oracle.place(lab, cl, 6, 5, 2, 0.25)
    out["place_6x5"] = {"lab": lab.tolist(), "cell_lab": cl.tolist(), "grid_w": 6, "grid_h": 5,
                        "max_usage": 2, "weight": 0.25, "grid": r["grid"].tolist(),
                        "stage": r["stage"].tolist(), "usage": r["usage"].tolist(),
                        "score_bits": bits(r["score"]), "stats": r["stats"]}

    # swap deltas on a 4x4 grid with two empty cells (reference test Labs + more)
    lab2 = np.array([[50, 0, 0], [60, 10, 10], [40, -10, -10], [55, 5, 5], [0, -100, -100],
                     [100, 100, 100], [25, -50, 50]], np.float32)
    g = rng.integers(0, 7, 16).astype(np.int32)
    g[[3, 9]] = -1
    pairs = [[int(a), int(b), int(c), int(d)] for a, b, c, d in rng.integers(0, 4, (40, 4))]
    out["swap_4x4"] = {"lab": lab2.tolist(), "grid": g.tolist(), "pairs": pairs,
                       "delta_bits": [int(np.float32(oracle.swap_delta(lab2, g, 4, 4, *p)).view(np.uint32))
                                      for p in pairs],
                       "total_cost_bits": int(np.float32(oracle.total_cost(lab2, g, 4, 4)).view(np.uint32))}

    # SA: 5x4 grid, seeded counter RNG, trace
    lab3 = np.round(rng.uniform(-40, 40, (9, 3)), 1).astype(np.float32)
    g3 = rng.integers(0, 9, 20).astype(np.int32)
    a = oracle.anneal(lab3, g3, 5, 4, 300, 3.0, 0.995, seed=77, trace=True)
    out["anneal_5x4"] = {"lab": lab3.tolist(), "grid": g3.tolist(), "iterations": 300, "t0": 3.0,
                         "decay": 0.995, "seed": 77, "final_grid": a["grid"].tolist(),
                         "delta_bits": bits(a["trace_delta"]),
                         "accept": a["trace_accept"].tolist(),
                         "final_cost_bits": int(np.float32(a["result"]["final_cost"]).view(np.uint32))}

    # K nearest for 8 query Labs against 200 materials
    lab4 = np.round(rng.uniform(-50, 80, (200, 3)), 1).astype(np.float32)
    q = np.round(rng.uniform(-50, 80, (8, 3)), 1).astype(np.float32)
    idx, d = oracle.nearest_n(lab4, q, 100)
    out["topk_200"] = {"lab": lab4.tolist(), "query": q.tolist(), "idx": idx.tolist(),
                       "sqdist_bits": bits(d)}

    # Lab, average RGB and adjustments on the reference tests' pixels
    px = [[255, 0, 0], [128, 128, 128], [0, 255, 0], [0, 0, 255], [128, 64, 192], [100, 95, 105],
          [255, 100, 50], [200, 150, 150], [50, 50, 50], [255, 255, 255], [0, 0, 0]]
    out["lab_pixels"] = {"rgb": px, "lab_bits": [bits(oracle.rgb8_to_lab(*p)) for p in px]}
    adjs = [[0.0, 1.0, 0.0, 1.0], [0.2, 1.0, 0.0, 1.0], [0.0, 1.5, 0.0, 1.0], [0.0, 1.0, 120.0, 1.0],
            [0.0, 1.0, 0.0, 2.0], [0.0, 1.0, 0.0, 0.5], [0.5, 1.5, 60.0, 1.5]]
    out["adjust_pixels"] = {"adj": adjs, "rgb": px,
                            "out": [[oracle.adjust_pixel(np.array(a_, np.float32), p).tolist() for p in px]
                                    for a_ in adjs]}
    pairs2 = [([128, 128, 128], [128, 128, 128]), ([50, 50, 50], [200, 200, 200]),
              ([255, 0, 0], [0, 255, 0]), ([255, 0, 0], [200, 150, 150]),
              ([255, 0, 0], [255, 0, 255]), ([100, 95, 105], [105, 95, 100])]
    out["optimal_adjustment"] = {"pairs": pairs2, "strength": 1.0,
                                 "adj_bits": [bits(oracle.optimal_adjustment(t, g_, 1.0)) for t, g_ in pairs2]}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
