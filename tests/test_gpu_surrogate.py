"""The learned surrogate on the GPU: forward/backward kernels against the CPU
reference (oracle/mlp_ref.py), training, the error map, and assembly with the
network against the plain series."""
import numpy as np
import pytest

import hm_bem
import meshes
import mlp_ref as M
from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    c = hm_bem.Context()
    st0 = c.surrogate_train(nsamples=8192, epochs=0, batch=1024, seed=11)
    blob0 = c.surrogate_get()
    st = c.surrogate_train(nsamples=1_000_000, epochs=16, batch=4096, lr=3e-3, seed=12)
    print("untrained", st0, "trained", st, flush=True)
    yield c, st0, blob0, st
    c.close()


def test_gradients_match_cpu_reference(trained):
    c, _, blob0, _ = trained
    g = hm_bem.Context()
    g.surrogate_set(blob0)
    p = blob0[:M.N_PARAMS].astype(np.float64)
    ys = blob0[M.N_PARAMS:M.N_PARAMS + 3].astype(np.float64)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (512, 4)).astype(np.float32)
    u = (rng.normal(size=(512, 3)) * ys).astype(np.float32)
    # FP32 network: the GPU's FP32 kernels against the float64 reference
    gg, lg = g.surrogate_grad(x, u, quant=False)
    gc = M.grad(p, x.astype(np.float64), u.astype(np.float64), ys, quant=False)
    lc = M.loss(p, x.astype(np.float64), u.astype(np.float64), ys, quant=False)
    assert abs(lg - lc) <= 1e-5 * abs(lc)
    assert relerr(gg, gc) < 1e-4
    # BF16 network: rounding boundaries may fall differently for a few activations
    gq, lq = g.surrogate_grad(x, u, quant=True)
    gcq = M.grad(p, x.astype(np.float64), u.astype(np.float64), ys, quant=True)
    assert abs(lq - M.loss(p, x.astype(np.float64), u.astype(np.float64), ys, quant=True)) <= 1e-3 * abs(lq)
    assert relerr(gq, gcq) < 1e-2
    g.close()


def test_mfma_inference_matches_cpu_reference(trained):
    """The inference kernels run the BF16 network as MFMA tiles with FP32
    accumulation; the reference is the float64 network with the same BF16
    rounding of weights and activations. Accumulation order and the odd
    activation falling on the other side of a BF16 rounding boundary make them
    differ at ~1e-3 of the output scale; a wrong tile layout would make them
    differ at O(1)."""
    c, _, _, _ = trained
    blob = c.surrogate_get()
    p = blob[:M.N_PARAMS].astype(np.float64)
    ys = blob[M.N_PARAMS:M.N_PARAMS + 3].astype(np.float64)
    rng = np.random.default_rng(17)
    x = rng.uniform(-1, 1, (1000, 4)).astype(np.float32)       # not a multiple of the 64-pair wave tile
    u = c.surrogate_predict(x).astype(np.float64)
    y_ref, _, _ = M.forward(p, x.astype(np.float64), quant=True)
    u_ref = y_ref * ys
    scale = np.max(np.abs(u_ref), axis=0)
    err = np.abs(u - u_ref) / scale
    print("MFMA inference vs reference: max", err.max(axis=0), "median", np.median(err, axis=0), flush=True)
    assert np.all(err.max(axis=0) < 1e-2), err.max(axis=0)
    assert np.all(np.median(err, axis=0) < 1e-3), np.median(err, axis=0)


def test_training_reduces_error(trained):
    _, st0, _, st = trained
    assert st["val_rmse"] < 0.1 * st0["val_rmse"], (st0, st)
    assert st["cells_below_1e-3"] > st0["cells_below_1e-3"]
    assert st["train_loss"] < 1e-2


def test_network_against_plain_series(trained):
    c, _, _, _ = trained
    rng = np.random.default_rng(5)
    n = 4000
    rh = np.exp(rng.uniform(np.log(0.05), np.log(0.5), n))
    zf, zp = -rng.uniform(0, 0.25, n), -rng.uniform(0, 0.25, n)
    kh = np.exp(rng.uniform(np.log(0.01), np.log(50.0), n))
    tn, tp, bound = c.surrogate_eval(rh, zf + zp, zf - zp, kh)
    scale = np.max(np.abs(tp), axis=0)
    err = np.abs(tn - tp) / scale
    assert np.median(err) < 1e-2
    assert np.all(np.isfinite(bound))


def _quad():
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0 * i, 12.0 * j)) for j in range(2) for i in range(2)]
    return meshes.Problem(bs, h=40.0, omegas=[0.6, 1.4], beta=np.array([0.0, np.pi / 2]))


def test_assembly_with_surrogate(trained):
    c, _, _, _ = trained
    blob = c.surrogate_get()
    p = _quad()
    ctx = hm_bem.Context()
    ctx.setup(p)
    ctx.surrogate_set(blob)
    for w in p.omegas:
        ctx.set_surrogate(False)
        ref = ctx.solve_frequency(w, hm_bem.HM_FP64)
        ctx.set_frequency(w)
        ctx.assemble()
        a0, s0, _ = ctx.get_system()
        # tolerance 0: every pair falls back to the plain kernels -> identical matrices
        ctx.set_surrogate(True, 0.0)
        ctx.set_frequency(w)
        ctx.assemble()
        a1, s1, _ = ctx.get_system()
        assert np.array_equal(a1, a0) and np.array_equal(s1, s0)
        assert ctx.surrogate_counts()["network"] == 0
        # tolerance 3e-3: the network takes the cells it was validated on
        ctx.set_surrogate(True, 3e-3)
        ctx.set_frequency(w)
        ctx.assemble()
        a2, s2, _ = ctx.get_system()
        k = ctx.surrogate_counts()
        assert k["network"] > 0, k
        assert relerr(a2, a0) < 3e-3 and relerr(s2, s0) < 3e-3, (relerr(a2, a0), relerr(s2, s0), k)
        got = ctx.solve_frequency(w, hm_bem.HM_FP64)
        for key in ("amas", "bdmp", "exfc"):
            assert relerr(got[key], ref[key]) < 2e-3, (w, key, relerr(got[key], ref[key]), k)
        print("omega", w, "counts", k, "A err", relerr(a2, a0), "forces err",
              {key: relerr(got[key], ref[key]) for key in ("amas", "bdmp", "exfc")}, flush=True)
    ctx.close()


def test_hmatrix_with_surrogate(trained):
    """H-matrix form with the network on: it serves the EIGENE pairs of the
    diagonal blocks through the same error map and partition as the dense
    assembly; the cross approximation keeps the plain series (network values
    switched on and off cell by cell inside a block are not low rank: the
    blocks then stop at the rank cap). Forces against the plain FP64 dense solve
    at the bar of the dense assembly test above (2e-3 for this briefly trained
    network); where the network takes pairs the result must move."""
    c, _, _, _ = trained
    blob = c.surrogate_get()
    p = _quad()
    dense = hm_bem.Context()
    hmx = hm_bem.Context()
    try:
        dense.setup(p)
        hmx.setup(p)
        hmx.surrogate_set(blob)
        hmx.set_hmatrix(True, eps=1e-5, eps_s=1e-6, kmax=50)
        hmx.set_gmres(restart=50, tol=1e-8, maxiter=500)
        for w in p.omegas:
            ref = dense.solve_frequency(w, hm_bem.HM_FP64)
            hmx.set_surrogate(False)
            plain = hmx.solve_frequency(w, hm_bem.HM_GMRES)
            hmx.set_surrogate(True, 3e-3)
            got = hmx.solve_frequency(w, hm_bem.HM_GMRES)
            k = hmx.surrogate_counts()
            st = hmx.hmatrix_stats()
            errs = {key: relerr(got[key], ref[key]) for key in ("amas", "bdmp", "exfc")}
            moved = max(relerr(got[key], plain[key]) for key in ("amas", "bdmp", "exfc"))
            print("omega", w, "counts", k, "errors", errs, "network moved results by", moved, "hmatrix", st, flush=True)
            assert st["unconverged"] == 0, st
            if k["network"] > 0:
                assert moved > 0.0
            for key in errs:
                assert errs[key] < 2e-3, (w, key, errs)
    finally:
        dense.close()
        hmx.close()


def test_surrogate_argument_errors():
    c = hm_bem.Context()
    with pytest.raises(hm_bem.HMError) as e:
        c.set_surrogate(True, 1e-3)
    assert e.value.code == -200
    with pytest.raises(hm_bem.HMError) as e:
        c.surrogate_train(nsamples=100)
    assert e.value.code == -102
    c.close()
