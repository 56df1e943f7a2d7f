"""Several frequencies in flight (hm_set_workers): hm_run_frequencies on clones
of the context, each frequency in a host thread of its own, against the same
frequencies one after another on the context alone.

Bar: equal bit for bit. Each frequency's outputs come from the same kernels on
the same inputs whichever context computes them (the classification's pair
order may differ between contexts, but every pair is evaluated on its own, and
every reduction has a fixed order).
"""
import numpy as np
import pytest

import hm_bem
import meshes

pytestmark = pytest.mark.gpu


def _array(h, omegas):
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0 * i, 12.0 * j)) for j in range(2) for i in range(2)]
    return meshes.Problem(bs, h=h, omegas=list(omegas), beta=np.array([0.0, np.pi / 2]))


def _run(p, workers, hmatrix, precision):
    c = hm_bem.Context()
    try:
        c.setup(p)
        if hmatrix:
            c.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
            c.set_gmres(restart=50, tol=1e-6)
        c.set_workers(workers)
        first = c.run_frequencies(p.omegas, precision)
        # a second call reuses the clones
        second = c.run_frequencies(p.omegas[::-1], precision)
        return first, second
    finally:
        c.close()


@pytest.mark.parametrize("h,hmatrix", [(60.0, True), (-1.0, True), (60.0, False)], ids=["hm-finite", "hm-deep", "dense"])
def test_workers_match_one_context(h, hmatrix):
    p = _array(h, (0.4, 0.9, 1.3, 1.7, 2.2))
    prec = hm_bem.HM_GMRES if hmatrix else hm_bem.HM_FP64
    ref1, ref2 = _run(p, 1, hmatrix, prec)
    for nw in (2, 3):
        got1, got2 = _run(p, nw, hmatrix, prec)
        for a, b in zip(got1 + got2, ref1 + ref2):
            assert np.array_equal(a, b), (h, hmatrix, nw, float(np.max(np.abs(a - b))))
    # the reversed list gives each frequency's outputs in its own slot
    for a, b in zip(ref1, ref2):
        assert np.array_equal(a, b[::-1])
