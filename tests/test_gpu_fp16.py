"""FP16 MFMA application of the block preconditioner (H-matrix form with the
explicit block inverses, hm_set_precond_fp16).

Bars:
- the complex GEMM (hm_cgemm_f16_check: operands rounded to FP16 on the device,
  FP32 accumulation): every product of two FP16 values is exact in FP32, so the
  only error is the FP32 summation of the 2k real terms of each entry, at most
  2k 2^-24 times the sum of their magnitudes (the classical bound for recursive
  summation). A transposed operand, a wrong sign of the imaginary product or a
  bfloat16 rounding breaks it by orders of magnitude at the small k, and the
  distance to the unrounded product is FP16's (a few 1e-4, not 1e-2);
- the solution with the FP16 preconditioner against the FP32 one: GMRES stops
  on the FP64 residual in both (||B - A X|| <= tol ||B||, tol = 1e-6), so
  A/B/F agree to ~cond * tol (bar 2e-5, as for the other GMRES forms). The
  FP16 rounding of the inverses (~1e-3 of D M^-1 - I) limits what one restart
  cycle reaches to about that, so at tol = 1e-6 the FP16 form may need one
  cycle more than the FP32 one, never two more.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _rounded(x):
    return x.real.astype(np.float16).astype(np.float64) + 1j * x.imag.astype(np.float16).astype(np.float64)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (70, 33, 45), (128, 98, 300), (64, 64, 2048), (200, 7, 17)])
def test_cgemm_f16_matches_rounded_product(m, n, k):
    rng = np.random.default_rng(m * 10007 + n * 101 + k)
    a = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(np.complex64)
    b = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex64)
    c = hm_bem.Context()
    try:
        got = c.cgemm_f16(a, b).astype(np.complex128)
    finally:
        c.close()
    ar, br = _rounded(a), _rounded(b)
    ref = ar @ br
    mag = (np.abs(ar.real) + np.abs(ar.imag)) @ (np.abs(br.real) + np.abs(br.imag))
    bound = 2 * k * 2.0 ** -24 * mag + 1e-30
    worst = float(np.max(np.maximum(np.abs(got.real - ref.real), np.abs(got.imag - ref.imag)) / bound))
    print(f"m={m} n={n} k={k}: worst error / summation bound {worst:.3f}", flush=True)
    assert worst <= 1.0
    # the rounding to FP16 itself: a few 1e-4 of the product
    exact = a.astype(np.complex128) @ b.astype(np.complex128)
    assert relerr(got, exact) < 2e-3


def _array(nx, ny, h, spacing=12.0):
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(spacing * i, spacing * j)) for j in range(ny) for i in range(nx)]
    return meshes.Problem(bs, h=h, omegas=[0.5, 1.6], beta=np.array([0.0, np.pi / 2]))


@pytest.mark.parametrize("h", [60.0, -1.0], ids=["finite", "deep"])
def test_fp16_preconditioner_against_fp32(h):
    p = _array(2, 2, h)
    ctx = {}
    try:
        for half in (True, False):
            c = hm_bem.Context()
            ctx[half] = c
            c.setup(p)
            c.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
            c.set_gmres(restart=50, tol=1e-6, precond=-1)
            c.set_precond_fp16(half)
        for w in p.omegas:
            got = ctx[True].solve_frequency(w, hm_bem.HM_GMRES)
            s16 = ctx[True].gmres_stats()
            ref = ctx[False].solve_frequency(w, hm_bem.HM_GMRES)
            s32 = ctx[False].gmres_stats()
            errs = {k: relerr(got[k], ref[k]) for k in KEYS}
            print(f"h={h} w={w}: fp16 {s16} fp32 {s32} errors {errs}", flush=True)
            assert s16["converged"] and s16["rel_residual"] <= 1e-6, s16
            assert s16["cycles"] <= s32["cycles"] + 1, (s16, s32)
            assert max(errs.values()) < 2e-5, errs
    finally:
        for c in ctx.values():
            c.close()
