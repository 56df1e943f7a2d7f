"""The learned surrogate's infinite-depth branch: a network of Re GF and Re GFh
(the Havelock function of INFGREEN3D, InfGreen_Appr.f90:104-183, and its KR
derivative, less their singular terms) of (log KR, sin KR, cos KR, K(z+zeta)),
trained on the GPU from the plain Havelock function, driving the deep-water
assembly of the C1 cylinder (1024 hull + 256 waterplane panels, infinite
depth, irregular-frequency removal).

Bars:
- tolerance 0 (no cell qualifies): the surrogate kernel's fallback is the plain
  INFGREEN3D, so A/B/F equal the plain path's bit for bit;
- network against the plain Havelock function at fresh points: each point's
  error within its cell's bound (1.5 x the largest validation error of the
  cell) for at least 99 % of the points in cells at or below 2e-3 (the bound is
  a measured maximum, not a proof);
- cell tolerance 2e-3 (about half of C1's pairs): A/B/F within 1e-4 of the
  plain path at three frequencies spanning 0.1-2.0 rad/s.
What it gains is measured and printed, not asserted: the plain Havelock fits
(polynomials and a few transcendentals per pair) are cheaper than the network,
so the deep-water assembly with the network takes longer (about 2.5x at C1).
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


@pytest.fixture(scope="module")
def trained():
    p = meshes.c1_single_cylinder()
    c = hm_bem.Context()
    c.setup(p)
    st = c.surrogate_train_deep(nsamples=16_000_000, epochs=60, batch=4096, lr=2e-3, seed=20261205)
    print("deep-water surrogate training", st, flush=True)
    yield p, c, st
    c.close()


def test_deep_network_against_the_plain_havelock_function(trained):
    _, c, st = trained
    assert st["val_rmse"] < 5e-3, st
    rng = np.random.default_rng(3)
    kr = np.exp(rng.uniform(np.log(1e-3), np.log(50.0), 20000))
    kv = -20.0 * rng.uniform(0.0, 1.0, 20000) ** 2
    net, plain, bound = c.surrogate_eval_deep(kr, kv)
    d = np.hypot(kr, kv)
    m0 = np.abs(plain[:, 0]) + 1.0 / d
    m1 = np.abs(plain[:, 1]) + np.abs(plain[:, 0]) + 2.0 / d + 1.0 / d ** 2
    e = np.maximum(np.abs(net[:, 0] - plain[:, 0]) / m0, np.abs(net[:, 1] - plain[:, 1]) / m1)
    sel = bound <= 2e-3
    print(f"errors: median {np.median(e):.2e}, 90 % {np.quantile(e, 0.9):.2e}; {sel.mean():.1%} of the points in "
          f"cells <= 2e-3", flush=True)
    assert sel.sum() > 100
    assert np.mean(e[sel] <= bound[sel]) >= 0.99


def test_deep_network_in_the_c1_assembly(trained):
    p, c, _ = trained
    ws = [float(p.omegas[i]) for i in (0, 7, 15)]
    c.set_surrogate_deep(False)
    for w in ws:
        c.solve_frequency(w, hm_bem.HM_FP64)      # warm-up
    c.set_timing(True)
    ref = {w: c.solve_frequency(w, hm_bem.HM_FP64) for w in ws}
    t_plain = c.timing()["assembly"]
    c.set_timing(False)
    # no cell qualifies: the fallback is the plain kernel
    c.set_surrogate_deep(True, 0.0)
    got = c.solve_frequency(ws[1], hm_bem.HM_FP64)
    assert c.surrogate_deep_count() == 0
    for k in KEYS:
        assert np.array_equal(got[k], ref[ws[1]][k])
    c.set_surrogate_deep(True, 2e-3)
    c.set_timing(True)
    for w in ws:
        got = c.solve_frequency(w, hm_bem.HM_FP64)
        errs = {k: relerr(got[k], ref[w][k]) for k in KEYS}
        print(f"C1 w={w:.3f} deep-water network at 2e-3: errors {errs}", flush=True)
        for k in KEYS:
            assert errs[k] <= 1e-4, (w, k, errs)
    t_net = c.timing()["assembly"]
    c.set_timing(False)
    taken = c.surrogate_deep_count()
    total = len(ws) * p.n * (p.n + p.ni)
    print(f"network took {taken} of {total} pairs ({taken / total:.1%}); assembly per system: plain "
          f"{t_plain[0] / t_plain[1]:.3f} ms, with the network {t_net[0] / t_net[1]:.3f} ms "
          f"(gain {t_plain[0] / t_plain[1] / (t_net[0] / t_net[1]):.2f}x)", flush=True)
    assert taken > 0
    c.set_surrogate_deep(False)
