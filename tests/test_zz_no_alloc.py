"""hm_fingreen3d_batch_noalloc (the finite-depth Green function batch staged in
the library's fixed device array, no device allocation) against
hm_fingreen3d_batch through the C-ABI: bit for bit the same GRN on two calls in
a row, the second with a smaller input (the array keeps the first call's
contents, which the second must not read), across all four series regimes;
n above the array's 262144 pairs is refused with a clear error and leaves the
array usable."""
import ctypes

import numpy as np
import pytest

import hm_bem

pytestmark = pytest.mark.gpu
MAX_POINTS = 262144


def _pairs(rng, n, h):
    # horizontal distances across the LINTON / PIDCOCK / EIGENE / EIGEN ranges
    rr = h * np.exp(rng.uniform(np.log(1e-4), np.log(2.0), n))
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    src = np.column_stack([rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), -rng.uniform(0.0, 0.9 * h, n)])
    fld = np.column_stack([src[:, 0] + rr * np.cos(th), src[:, 1] + rr * np.sin(th), -rng.uniform(0.0, 0.9 * h, n)])
    flag = rng.integers(0, 2, n).astype(np.int32)
    return src, fld, flag


def _noalloc(c, src, fld, v, wvn, h, flag):
    n = len(src)
    cols = [np.ascontiguousarray(a) for a in (src[:, 0], fld[:, 0], src[:, 1], fld[:, 1], src[:, 2], fld[:, 2])]
    wvn = np.ascontiguousarray(wvn, dtype=np.float64)
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    out = np.zeros((n, 4), dtype=np.complex128)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = c.L.hm_fingreen3d_batch_noalloc(c.ptr, ctypes.c_int(n), *[p(a) for a in cols], ctypes.c_double(v), p(wvn),
                                        ctypes.c_int(len(wvn)), ctypes.c_double(h), p(fl), p(out))
    return rc, out


def test_noalloc_matches_the_allocating_entry_point():
    rng = np.random.default_rng(20261204)
    h, w = 40.0, 0.9
    v = w * w / 9.81
    wvn = hm_bem.dispersion(w, h, 9.81)
    c = hm_bem.Context()
    try:
        for n in (40000, 3001):                      # the second call smaller than the first
            src, fld, flag = _pairs(rng, n, h)
            print(f"no-alloc batch: n={n}", flush=True)
            want = c.fingreen3d(src, fld, v, wvn, h, flag)
            rc, got = _noalloc(c, src, fld, v, wvn, h, flag)
            assert rc == 0, c.L.hm_error_string(c.ptr)
            assert np.array_equal(got, want), np.max(np.abs(got - want))
        src, fld, flag = _pairs(rng, 8, h)
        big = np.zeros((MAX_POINTS + 1, 3))
        rc, _ = _noalloc(c, big, big, v, wvn, h, np.zeros(MAX_POINTS + 1, dtype=np.int32))
        assert rc == -102 and b"exceeds" in c.L.hm_error_string(c.ptr)
        rc, got = _noalloc(c, src, fld, v, wvn, h, flag)
        assert rc == 0 and np.array_equal(got, c.fingreen3d(src, fld, v, wvn, h, flag))
    finally:
        c.close()
