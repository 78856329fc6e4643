#!/usr/bin/env python3
"""Regenerates the committed fixtures in tests/golden/ (run in the build
container; the GPU tests only read the JSON files).

philox_rocrand.json  Philox4x32-10 words computed by rocRAND's host-callable
                     engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h),
                     an implementation independent of this project's.
psa_tiny.json        inputs and outputs of the CPU oracle's PSA restatement on
                     small hand-picked and seeded weight vectors (N <= 64): the
                     double instantiation with the lower-bound split of the
                     reference GPU kernels ("f64") and with the reference CPU
                     split ("f64_cpu_split"), and the float instantiation with
                     the reference CPU split, unscaled and with the 1e-4 scale.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def rocrand_vectors():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "rr")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe,
                        os.path.join(HERE, "rocrand_philox_vectors.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def psa_cases():
    cases = []
    hand = [
        ([1.0, 1.0, 1.0, 1.0], 2),
        ([4.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0], 4),
        ([0.5, 3.0, 0.25, 0.25, 2.0, 0.0, 1.0, 1.0, 0.0, 5.0], 3),
        ([1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0], 2),
        ([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9], 1),
    ]
    for w, K in hand:
        cases.append((np.asarray(w, np.float32), K, "hand"))
    for n, K, seed in [(17, 4, 0), (33, 8, 1), (64, 16, 2), (64, 2, 3), (50, 49, 4)]:
        cases.append((O.gen_uniform01(seed, n), K, f"uniform seed {seed}"))
    cases.append((O.gen_zipf(2, 1.1, 64), 16, "zipf 1.1 seed 2"))
    cases.append((O.gen_lognormal(1, 0.0, 2.0, 48), 6, "lognormal(0,2) seed 1"))
    out = []
    for w, K, label in cases:
        p64, a64, d64 = O.psa_alias_table_f64(w, K, split_rule=O.SPLIT_LOWER_BOUND)
        p64c, a64c, _ = O.psa_alias_table_f64(w, K, split_rule=O.SPLIT_REFERENCE_CPU)
        p32, a32, d32 = O.psa_alias_table_f32(w, K, reference_scale=False)
        p32s, a32s, _ = O.psa_alias_table_f32(w, K, reference_scale=True)
        out.append(dict(
            label=label, K=K, weights=[float(x) for x in w],
            f64=dict(mean=d64["mean"], H=d64["H"], idx_cpu=d64["idx_cpu"].tolist(),
                     splits=[[int(s["i"]), int(s["j"]), float(s["spill"])] for s in d64["splits"]],
                     p=[float(x) for x in p64], a=a64.tolist()),
            f64_cpu_split=dict(p=[float(x) for x in p64c], a=a64c.tolist()),
            f32=dict(mean=d32["mean"], H=d32["H"], p=[float(x) for x in p32], a=a32.tolist()),
            f32_scaled=dict(p=[float(x) for x in p32s], a=a32s.tolist()),
        ))
    return out


def main():
    O.build()
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        json.dump(rocrand_vectors(), f, indent=1)
    with open(os.path.join(HERE, "psa_tiny.json"), "w") as f:
        json.dump(psa_cases(), f, indent=1)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
