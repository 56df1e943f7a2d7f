"""CPU reference of the surrogate network (TEST INFRASTRUCTURE ONLY).

The network of the library's learned surrogate (hams-mrel_amd/csrc/hm_surrogate.hip):
4 -> 64 -> 64 -> 64 -> 64 -> 3, SiLU after each hidden layer, parameters packed
as W0 (64x4, row-major out x in), b0 (64), then three times W (64x64), b (64),
then W4 (3x64), b4 (3). With quant=True the hidden and output weights and the
hidden activations are rounded to BF16 (round to nearest, ties to even), the
first layer, the biases and all accumulation stay in higher precision, as in
the GPU kernels. Loss = mean over batch and outputs of (y_hat - u / yscale)^2.

Everything here is float64, so it serves as the reference for the GPU's FP32
forward/backward kernels; `grad` is itself pinned by central differences of
`loss` (tests/test_surrogate_cpu.py).
"""
import numpy as np

H = 64
N_IN = 4
N_OUT = 3
OFF_W0 = 0
OFF_B0 = OFF_W0 + H * N_IN
OFF_W1 = OFF_B0 + H
HIDDEN_STRIDE = H * H + H
OFF_W4 = OFF_W1 + 3 * HIDDEN_STRIDE
OFF_B4 = OFF_W4 + N_OUT * H
N_PARAMS = OFF_B4 + N_OUT          # 12995


def bf16(x):
    """round float values to the nearest BF16 value (ties to even), as float64"""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w = [p[OFF_W0:OFF_B0].reshape(H, N_IN)]
    b = [p[OFF_B0:OFF_W1]]
    for layer in range(3):
        o = OFF_W1 + layer * HIDDEN_STRIDE
        w.append(p[o:o + H * H].reshape(H, H))
        b.append(p[o + H * H:o + HIDDEN_STRIDE])
    w.append(p[OFF_W4:OFF_B4].reshape(N_OUT, H))
    b.append(p[OFF_B4:N_PARAMS])
    return w, b


def pack(w, b):
    out = np.zeros(N_PARAMS)
    out[OFF_W0:OFF_B0] = w[0].ravel()
    out[OFF_B0:OFF_W1] = b[0]
    for layer in range(3):
        o = OFF_W1 + layer * HIDDEN_STRIDE
        out[o:o + H * H] = w[layer + 1].ravel()
        out[o + H * H:o + HIDDEN_STRIDE] = b[layer + 1]
    out[OFF_W4:OFF_B4] = w[4].ravel()
    out[OFF_B4:N_PARAMS] = b[4]
    return out


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, x, quant=True):
    """returns outputs (B x 3) and the per-layer (z, a) for the backward pass"""
    w, b = unpack(p)
    if quant:
        w = [w[0]] + [bf16(m) for m in w[1:]]
    a = np.asarray(x, dtype=np.float64)
    cache = []
    for layer in range(4):
        z = a @ w[layer].T + b[layer]
        act = z * _sig(z)
        if quant:
            act = bf16(act)
        cache.append((a, z))
        a = act
    y = a @ w[4].T + b[4]
    cache.append((a, None))
    return y, cache, w


def loss(p, x, u, yscale, quant=True):
    y, _, _ = forward(p, x, quant)
    e = y - np.asarray(u, dtype=np.float64) / np.asarray(yscale, dtype=np.float64)
    return float(np.mean(e * e))


def grad(p, x, u, yscale, quant=True):
    """d loss / d p (straight-through for the BF16 rounding of activations)"""
    y, cache, w = forward(p, x, quant)
    bsz = y.shape[0]
    e = y - np.asarray(u, dtype=np.float64) / np.asarray(yscale, dtype=np.float64)
    dy = 2.0 * e / (N_OUT * bsz)
    gw = [None] * 5
    gb = [None] * 5
    a4 = cache[4][0]
    gw[4] = dy.T @ a4
    gb[4] = dy.sum(0)
    da = dy @ w[4]
    for layer in range(3, -1, -1):
        a_in, z = cache[layer]
        s = _sig(z)
        dz = da * (s + z * s * (1.0 - s))
        gw[layer] = dz.T @ a_in
        gb[layer] = dz.sum(0)
        da = dz @ w[layer]
    return pack(gw, gb)
