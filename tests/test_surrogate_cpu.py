"""The CPU reference of the surrogate network (oracle/mlp_ref.py), pinned by
central differences and by BF16 rounding cases; no GPU."""
import numpy as np

import mlp_ref as M


def test_bf16_rounding_cases():
    one = np.float32(1.0)
    ulp = 2.0 ** -7                       # BF16 spacing at 1.0
    cases = [(1.0, 1.0), (1.0 + ulp, 1.0 + ulp),
             (1.0 + ulp / 2, 1.0),         # tie -> even (1.0)
             (1.0 + 3 * ulp / 2, 1.0 + 2 * ulp),   # tie -> even (1 + 2 ulp)
             (1.0 + ulp / 2 + 2.0 ** -20, 1.0 + ulp),
             (-3.140625, -3.140625), (0.0, 0.0)]
    for x, want in cases:
        assert M.bf16(np.float32(x)) == want, (x, M.bf16(np.float32(x)), want)
    assert M.bf16(one) == 1.0


def test_pack_roundtrip():
    p = np.random.default_rng(1).normal(size=M.N_PARAMS)
    w, b = M.unpack(p)
    assert np.array_equal(M.pack(w, b), p)


def test_grad_against_central_differences():
    rng = np.random.default_rng(7)
    p = rng.normal(size=M.N_PARAMS) * 0.2
    x = rng.uniform(-1, 1, (16, 4))
    ys = np.array([0.7, 1.3, 2.0])
    u = rng.normal(size=(16, 3)) * ys
    g = M.grad(p, x, u, ys, quant=False)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 40, replace=False),
                          [M.OFF_W0, M.OFF_B0, M.OFF_W1 + 5, M.OFF_W4 + 3, M.OFF_B4 + 2]])
    for k in idx:
        e = np.zeros(M.N_PARAMS)
        e[k] = 1e-6
        fd = (M.loss(p + e, x, u, ys, False) - M.loss(p - e, x, u, ys, False)) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), (k, fd, g[k])
