"""Zero- and infinite-frequency limits (SYBO = 1) on the GPU against the oracle's
same chain (itself pinned by the hemisphere closed forms in
test_oracle_limits.py).

Bar: 1e-10 relative for the FP64 path, as for ordinary frequencies. In finite
depth the reference's infinite-frequency incident wave is cosh(1e20 (z+h)) /
sinh(1e20 h) = inf/inf (PatclVelct.f90:110-111), so its excitation is NaN; the
GPU path must give NaN in exactly the same entries and match elsewhere.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr
from oracle import Oracle

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _cases():
    return [
        ("hemisphere deep", meshes.Problem([meshes.hemisphere(1.0, 24, 6)], h=-1.0, omegas=[1.0],
                                           beta=np.array([0.0, 1.0]))),
        ("cylinder pair h=31.7", meshes.Problem([meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(0.0, 0.0)),
                                                 meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0, 0.0))],
                                                h=31.7, omegas=[1.0], beta=np.array([0.0]))),
    ]


def _compare(got, ref, bar, what):
    for k in KEYS:
        g, r = np.asarray(got[k]), np.asarray(ref[k])
        assert np.array_equal(np.isnan(g), np.isnan(r)), (what, k, "NaN pattern")
        ok = ~np.isnan(r)
        if ok.any():
            e = relerr(g[ok], r[ok])
            assert e < bar, (what, k, e)


@pytest.mark.parametrize("which", [0, 1], ids=["zero", "infinite"])
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_limit_fp64_matches_oracle(case, which):
    name, prob = case
    ref = Oracle().solve_limit(prob, which)
    c = hm_bem.Context()
    try:
        c.setup(prob)
        got = c.solve_limit(which, hm_bem.HM_FP64)
        _compare(got, ref, 1e-10, (name, which))
        if which == 0:
            assert np.max(np.abs(got["bdmp"])) == 0.0 and np.max(np.abs(got["exfc"])) == 0.0
        # an ordinary frequency afterwards is unaffected by the limit state:
        # bit-identical to a fresh context
        w = 0.9
        after = c.solve_frequency(w, hm_bem.HM_FP64)
        fresh = hm_bem.Context()
        try:
            fresh.setup(prob)
            ref_w = fresh.solve_frequency(w, hm_bem.HM_FP64)
        finally:
            fresh.close()
        for k in KEYS:
            assert np.array_equal(after[k], ref_w[k]), (name, k)
    finally:
        c.close()


def test_zero_frequency_hmatrix():
    prob = meshes.Problem([meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0 * i, 12.0 * j))
                           for j in range(2) for i in range(2)], h=60.0, omegas=[1.0], beta=np.array([0.0]))
    dense, hm = hm_bem.Context(), hm_bem.Context()
    try:
        dense.setup(prob)
        hm.setup(prob)
        hm.set_hmatrix(True, eps=1e-5, eps_s=1e-6, kmax=50)
        ref = dense.solve_limit(0, hm_bem.HM_FP64)
        got = hm.solve_limit(0, hm_bem.HM_GMRES)
        assert relerr(got["amas"], ref["amas"]) < 1e-4
        assert np.max(np.abs(got["bdmp"])) == 0.0 and np.max(np.abs(got["exfc"])) == 0.0
    finally:
        dense.close()
        hm.close()


def test_limit_argument_errors():
    c = hm_bem.Context()
    try:
        with pytest.raises(hm_bem.HMError, match="before hm_set_depth"):
            c.set_frequency_limit(0)
        c.setup(_cases()[0][1])
        with pytest.raises(hm_bem.HMError, match="which must be"):
            c.set_frequency_limit(2)
    finally:
        c.close()
