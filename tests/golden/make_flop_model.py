"""Algorithmic work per panel pair, by regime and series length, for the roofline
figures of bench.py (writes tests/golden/flop_model.json).

The count comes from the oracle's operation counter (oracle/hm_oracle.c:
additions, multiplications, divisions and square roots count 1, each libm
transcendental call counts HMO_TRANS_FLOPS = 20), over a random sample of the
workload's own pairs at a spread of its frequencies: one Green-function
evaluation per pair plus the Rankine panel integral for near pairs, exactly the
work of one thread of the assembly kernels.

    python tests/golden/make_flop_model.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hams-mrel_amd"))
import meshes  # noqa: E402
from oracle import Oracle  # noqa: E402

F32 = lambda x: float(np.float32(x))  # noqa: E731


def classify(rr, h):
    rh = np.abs(rr / h)
    reg = np.where(rh < 0.0005, 0, np.where(rh < 0.05, 1, np.where(rh < 0.5, 2, 3)))
    nt = np.zeros_like(reg)
    nt = np.where(reg == 2, (F32(-88.89) * rh + F32(54.45)).astype(int), nt)
    nt = np.where(reg == 1, (F32(-1010.10) * rh + F32(100.50)).astype(int), nt)
    return reg, nt


def model(prob, O, per_bucket=60, nfreq=3, seed=1):
    rng = np.random.default_rng(seed)
    n, nt_ = prob.n, prob.n + prob.ni
    fp = np.concatenate([prob.xyz_p, prob.ixyz_p]) if prob.ni else prob.xyz_p
    i = rng.integers(0, nt_, 400000)
    j = rng.integers(0, n, 400000)
    src, fld = prob.xyz_p[j], fp[i]
    rr = np.hypot(src[:, 0] - fld[:, 0], src[:, 1] - fld[:, 1])
    flag = np.linalg.norm(src - fld, axis=1) <= 50.0 * prob.pnsz[j]
    ws = prob.omegas[np.linspace(0, len(prob.omegas) - 1, nfreq).astype(int)]
    out = {}
    if prob.h < 0:
        keys = np.zeros(len(i), dtype=int)
    else:
        reg, ntt = classify(rr, prob.h)
        keys = reg * 128 + ntt
    for key in np.unique(keys):
        sel = np.nonzero(keys == key)[0][:per_bucket]
        tot = 0.0
        for w in ws:
            wvn = O.dispersion(w, prob.h, g_disp=prob.g_disp) if prob.h > 0 else None
            for k in sel:
                O.flops_reset()
                if flag[k]:
                    vv = prob.xyz[prob.ncon[j[k], :prob.ncn[j[k]]] - 1]
                    O.sglintbd(vv, fld[k])
                if prob.h > 0:
                    O.fingreen3d(src[k], fld[k], w * w / prob.g, wvn, prob.h, int(flag[k]))
                else:
                    O.infgreen3d(src[k], fld[k], w * w / prob.g, int(flag[k]))
                tot += O.flops()
        out[int(key)] = tot / (len(sel) * len(ws))
    return out


def main():
    O = Oracle()
    table = {}
    for name, fn in (("c1", meshes.c1_single_cylinder), ("c2", meshes.c2_three_cylinders),
                     ("c4", meshes.c4_eight_cylinders), ("c5", meshes.c5_sixteen_cylinders)):
        p = fn()
        table[name] = {"deep": bool(p.h < 0), "flops_per_pair": model(p, O)}
        print(name, len(table[name]["flops_per_pair"]), "buckets", flush=True)
    with open(os.path.join(HERE, "flop_model.json"), "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
