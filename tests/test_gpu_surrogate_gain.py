"""What the learned surrogate gains at equal accuracy.

A 2 x 2 array of the 8-cylinder workload's bodies (r = 4 m, T = 6 m, 16 m
spacing, 2048 panels each, h = 40 m): most pairs are in the EIGENE regime
(0.05 <= R/h < 0.5, up to 50 K0/K1 terms and three epsilon limits each), the
part the network replaces. The network is trained here (16M samples, 60
epochs, about 30 s on one MI355X), then each frequency is solved with the plain
series and with the network at the cell bound 1.5e-3.

Bars: added mass, damping and excitation within 1e-4 of the plain FP64 solve
(the accuracy the accelerated paths are held to), and assembly at least 1.3 x
faster with the network (measured on the 8-cylinder workload: 1.9-2.1 x at
<= 4.4e-5; here the printed ratio is the measurement).
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def test_surrogate_gain_at_equal_accuracy():
    bodies = [meshes.cylinder(4.0, 6.0, 64, 24, 8, center=(16.0 * i, 16.0 * j)) for j in range(2) for i in range(2)]
    prob = meshes.Problem(bodies, h=40.0, omegas=[0.611, 1.632], beta=np.deg2rad([0.0, 90.0]))
    c = hm_bem.Context()
    try:
        c.setup(prob)
        st = c.surrogate_train(nsamples=16_000_000, epochs=60, batch=4096, lr=2e-3, seed=20261204)
        print("trained", st, flush=True)
        hist = c.pair_histogram().sum(axis=1)
        print("pairs per regime (Linton, Pidcock, EIGENE, EIGEN)", hist, flush=True)
        c.set_surrogate(False)
        c.solve_frequency(prob.omegas[0], hm_bem.HM_FP64)            # warm-up
        for w in prob.omegas:
            c.set_surrogate(False)
            c.set_timing(True)
            ref = c.solve_frequency(w, hm_bem.HM_FP64)
            t_plain = c.timing()["assembly"][0]
            c.set_timing(False)
            c.set_surrogate(True, 1.5e-3)
            c.set_timing(True)
            got = c.solve_frequency(w, hm_bem.HM_FP64)
            tm = c.timing()
            c.set_timing(False)
            k = c.surrogate_counts()
            errs = {key: relerr(got[key], ref[key]) for key in KEYS}
            ratio = t_plain / tm["assembly"][0]
            print(f"omega {w}: assembly plain {t_plain:.1f} ms, with network {tm['assembly'][0]:.1f} ms "
                  f"(network {tm['asm_surrogate'][0]:.1f} ms, series {tm['asm_eigene'][0]:.1f} ms) = {ratio:.2f} x; "
                  f"pairs {k}; errors {errs}", flush=True)
            assert k["network"] > 0.5 * hist[2], k
            for key in KEYS:
                assert errs[key] < 1e-4, (w, key, errs)
            assert ratio >= 1.3, (w, ratio, t_plain, tm["assembly"][0])
    finally:
        c.close()
