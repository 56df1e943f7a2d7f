"""Oracle outputs for the full-size single-cylinder workload (1,024 hull +
256 interior waterplane panels, infinite depth, irregular-frequency removal),
written to tests/golden/c1_oracle.npz for tests/test_gpu_configs.py.

The oracle (oracle/hm_oracle.c) takes ~6 s per frequency on one host core, so
its A, B and F are computed here once, at three of the workload's sixteen
frequencies (the first, a middle one and the last), instead of in every GPU
run.

    python tests/golden/make_config_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hams-mrel_amd"))
import meshes  # noqa: E402
from oracle import Oracle  # noqa: E402

C1_FREQS = (0, 7, 15)


def main():
    O = Oracle()
    p = meshes.c1_single_cylinder()
    out = {"idx": np.array(C1_FREQS), "omegas": p.omegas[list(C1_FREQS)]}
    res = [O.solve_frequency(p, float(p.omegas[i])) for i in C1_FREQS]
    assert all(r["info"] == 0 for r in res)
    for k in ("amas", "bdmp", "exfc"):
        out[k] = np.array([r[k] for r in res])
    np.savez(os.path.join(HERE, "c1_oracle.npz"), **out)
    print("c1_oracle.npz:", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
