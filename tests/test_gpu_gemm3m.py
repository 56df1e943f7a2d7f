"""FP64 complex GEMM with three real products per complex one
(hm_cgemm_3m_check), the form the H-matrix residuals apply the diagonal
blocks with.

Bar: the FP64 rounding bound of the three-product form. The real part is
P1 - P2 with P1 = sum Ar Br and P2 = sum Ai Bi, the imaginary part
P3 - P1 - P2 with P3 = sum (Ar + Ai)(Br + Bi); each sum of k terms is within
k u of the sum of its terms' magnitudes (u = 2^-53), the two operand sums add
2u, the subtractions u each. So every entry is within 2 (k + 4) u times
sum_k (|Ar| + |Ai|)(|Br| + |Bi|) of the exact product. A wrong sign, a dropped
product or a transposed operand misses it by ~1e15 on random operands; the
four-product form meets it too, so it pins the arithmetic, not the form.
"""
import numpy as np
import pytest

import hm_bem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("trans", [False, True], ids=["a", "a^T"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (70, 33, 45), (128, 98, 300), (200, 7, 17), (64, 64, 2048),
                                   (2048, 130, 2048)])
def test_cgemm_3m_within_rounding_bound(m, n, k, trans):
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    a = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    b = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    c = hm_bem.Context()
    try:
        # a^T is passed as its k x m transpose, stored column-major
        got = c.cgemm_3m(np.ascontiguousarray(a.T), b, trans_a=True) if trans else c.cgemm_3m(a, b)
    finally:
        c.close()
    ref = a @ b
    mag = (np.abs(a.real) + np.abs(a.imag)) @ (np.abs(b.real) + np.abs(b.imag))
    bound = 2 * (k + 4) * 2.0 ** -53 * mag
    err = np.maximum(np.abs(got.real - ref.real), np.abs(got.imag - ref.imag))
    worst = float(np.max(err / bound))
    print(f"m={m} n={n} k={k} trans={trans}: worst error / rounding bound {worst:.4f}", flush=True)
    assert worst <= 1.0
