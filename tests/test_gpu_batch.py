"""Batches of independent problems (hm_batch_*) against single-problem solves.

The batch runs the same device code per pair, per LU panel and per force
reduction as a single-problem context (one kernel body, launched over the pair
lists of all members or with blockIdx.z = member), so with the same series
variant its FP64 results are compared with the single-problem FP64 results at
1e-13 (the two differ only where the compiler fuses a multiply-add differently
in the two kernels). HIP graph replays are compared bit for bit with the same
chain issued eagerly, three frequencies in a row with eager work in between.
FP32 LU + FP64 refinement is held to 1e-10 of FP64 (it refines until the last
correction is below 1e-13 of the solution). Members with interior waterplane
points (irregular-frequency removal) are held to the same bars. GMRES members
stop on ||B - A X|| <= tol ||B|| per column (tol 1e-8 here), so A/B/F agree with
FP64 to ~cond * tol: bar 1e-6, graph replays and eager chains bit for bit.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _members():
    beta = np.array([0.0, np.pi / 3])
    two = [meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(0.0, 0.0)),
           meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0, 0.0))]
    return [
        meshes.Problem([meshes.cylinder(3.0, 4.0, 12, 3, 2)], h=-1.0, omegas=[1.0], beta=beta),
        meshes.Problem([meshes.box(6.0, 4.0, 3.0, 6, 4, 3)], h=40.0, omegas=[1.0], beta=beta[:1]),
        meshes.Problem([meshes.cylinder(5.0, 5.0, 16, 4, 3)], h=80.0, omegas=[1.0], beta=beta),
        meshes.Problem([meshes.cylinder(2.0, 3.0, 20, 5, 2)], h=-1.0, omegas=[1.0], beta=beta),
        meshes.Problem(two, h=60.0, omegas=[1.0], beta=beta),
    ]


def _single(prob, w, variant):
    c = hm_bem.Context()
    try:
        c.setup(prob, frames=False)          # the batch takes global coordinates
        c.set_series_variant(variant)
        return c.solve_frequency(w, hm_bem.HM_FP64)
    finally:
        c.close()


def test_batch_fp64_matches_single_problems():
    probs = _members()
    b = hm_bem.Batch(probs)
    try:
        b.prepare(hm_bem.HM_FP64, autotune=True, use_graph=True)
        info = b.info()
        print("batch info", info, "sizes", [p.n for p in probs], flush=True)
        assert info["members"] == len(probs) and info["nmax"] == max(p.n for p in probs)
        assert info["series_variant"] in (0, 1, 2) and info["grid_blocks"] > 0
        tu = b.tuning()
        print("batch tuning", tu, flush=True)
        # the K chunk changes how the GEMM tiles stage K through LDS, not the
        # order in which the MFMAs accumulate it: results stay as the single
        # contexts' (K chunk 16)
        assert tu["gemm_kchunk"] in (8, 16, 32) and min(tu["lu_ms"]) > 0
        ws = np.array([0.6, 0.9, 1.3, 1.7, 2.2])          # one frequency per member
        for w in (1.1, ws):
            got = b.solve(w)
            for k, prob in enumerate(probs):
                wk = float(np.broadcast_to(w, (len(probs),))[k])
                ref = _single(prob, wk, info["series_variant"])
                for key in KEYS:
                    e = relerr(got[k][key], ref[key])
                    assert e < 1e-13, (k, wk, key, e)
    finally:
        b.close()


def test_batch_graph_replays_match_eager():
    probs = _members()
    g = hm_bem.Batch(probs)
    e = hm_bem.Batch(probs)
    try:
        g.prepare(hm_bem.HM_FP64, autotune=False, use_graph=True)
        e.prepare(hm_bem.HM_FP64, autotune=False, use_graph=False)
        for w in (0.8, 1.5, 0.8):
            ref = e.solve(w)
            got = g.solve(w)
            for k in range(len(probs)):
                for key in KEYS:
                    assert np.array_equal(got[k][key], ref[k][key]), (w, k, key)
    finally:
        g.close()
        e.close()


def test_batch_fp32_refinement():
    """FP32 LU + refinement, autotuned (series variant and the K chunk of the
    grouped GEMM tiles), against single FP64 contexts with the same series
    variant"""
    probs = _members()
    b32 = hm_bem.Batch(probs)
    try:
        b32.prepare(hm_bem.HM_FP32_IR, autotune=True, use_graph=True, nrefine=3)
        tu = b32.tuning()
        print("fp32 batch tuning", tu, flush=True)
        assert tu["gemm_kchunk"] in (8, 16, 32) and min(tu["lu_ms"]) > 0
        for w in (0.7, 1.9):
            got = b32.solve(w)
            info = b32.info()
            assert info["last_correction"] < 1e-13, info
            for k, prob in enumerate(probs):
                ref = _single(prob, w, tu["series_variant"])
                for key in KEYS:
                    err = relerr(got[k][key], ref[key])
                    assert err < 1e-10, (w, k, key, err, info)
    finally:
        b32.close()


def test_batch_argument_errors():
    probs = _members()[:2]
    b = hm_bem.Batch(probs)
    try:
        with pytest.raises(hm_bem.HMError, match="before hm_batch_prepare"):
            b.solve(1.0)
        with pytest.raises(hm_bem.HMError, match="precision"):
            b.prepare(7)
        b.prepare(hm_bem.HM_FP64, autotune=False, use_graph=False)
        with pytest.raises(hm_bem.HMError, match="must be > 0"):
            b.solve([1.0, 0.0])
        with pytest.raises(hm_bem.HMError, match="after hm_batch_prepare"):
            rc = b.L.hm_batch_set_constants(b.ptr, hm_bem._d(1000.0), hm_bem._d(9.81), hm_bem._d(9.81),
                                            hm_bem._d(1.0), hm_bem._d(0.0), hm_bem._d(0.0))
            b._chk(rc, "hm_batch_set_constants")
    finally:
        b.close()
    b = hm_bem.Batch(probs)
    try:
        rc = b.L.hm_batch_set_waterplane(b.ptr, hm_bem._i(7), hm_bem._i(0), None)
        with pytest.raises(hm_bem.HMError, match="no member"):
            b._chk(rc, "hm_batch_set_waterplane")
    finally:
        b.close()


def _waterplane_members():
    """members with interior waterplane points (irregular-frequency removal) among plain ones"""
    beta = np.array([0.0, np.pi / 4])
    return [
        meshes.Problem([meshes.cylinder(3.0, 4.0, 12, 3, 2, waterplane_nr=2)], h=-1.0, omegas=[1.0], beta=beta),
        meshes.Problem([meshes.cylinder(4.0, 3.0, 16, 3, 2)], h=50.0, omegas=[1.0], beta=beta),
        meshes.Problem([meshes.cylinder(5.0, 5.0, 16, 4, 3, waterplane_nr=3)], h=70.0, omegas=[1.0], beta=beta[:1]),
        meshes.Problem([meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(0.0, 0.0), waterplane_nr=2),
                        meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0, 0.0), waterplane_nr=2)],
                       h=60.0, omegas=[1.0], beta=beta),
    ]


def test_batch_waterplane_members():
    """members with waterplane points solve through the normal equations C =
    A^T A, D = A^T B in the batch as in a single context (the same kernels:
    1e-13 in FP64), and FP32 LU + refinement on C stays within 1e-10 of FP64"""
    probs = _waterplane_members()
    assert [p.ni > 0 for p in probs] == [True, False, True, True]
    b = hm_bem.Batch(probs)
    b32 = hm_bem.Batch(probs)
    try:
        b.prepare(hm_bem.HM_FP64, autotune=False, use_graph=True)
        b32.prepare(hm_bem.HM_FP32_IR, autotune=False, use_graph=True, nrefine=3)
        print("waterplane batch", b.info(), [(p.n, p.ni, p.h) for p in probs], flush=True)
        for w in (0.6, 1.7):
            got = b.solve(w)
            got32 = b32.solve(w)
            for k, prob in enumerate(probs):
                ref = _single(prob, w, 1)
                for key in KEYS:
                    e = relerr(got[k][key], ref[key])
                    e32 = relerr(got32[k][key], got[k][key])
                    assert e < 1e-13, (k, w, key, e)
                    assert e32 < 1e-10, (k, w, key, e32, b32.info())
    finally:
        b.close()
        b32.close()


def test_batch_gmres_members():
    probs = [p for p in _members()]
    ws = [0.7, 1.3]
    ref = hm_bem.Batch(probs)
    gm = hm_bem.Batch(probs)
    eager = hm_bem.Batch(probs)
    try:
        ref.prepare(hm_bem.HM_FP64, autotune=False, use_graph=False)
        for b, graph in ((gm, True), (eager, False)):
            b.set_gmres(restart=50, tol=1e-8, maxiter=400)
            b.prepare(hm_bem.HM_GMRES, autotune=False, use_graph=graph)
        for w in ws:
            want = ref.solve(w)
            got = gm.solve(w)
            st = gm.gmres_stats()
            again = eager.solve(w)
            print(f"w={w}: gmres {st}", flush=True)
            assert st["all_converged"] and st["worst_rel_residual"] <= 1e-8, st
            for k, (a, b2, c) in enumerate(zip(got, want, again)):
                for key in KEYS:
                    assert relerr(a[key], b2[key]) < 1e-6, (w, k, key, relerr(a[key], b2[key]))
                    assert np.array_equal(a[key], c[key]), (w, k, key)
    finally:
        ref.close()
        gm.close()
        eager.close()


def test_batch_gmres_refuses_waterplane_members():
    p = meshes.Problem([meshes.cylinder(3.0, 4.0, 12, 3, 2, waterplane_nr=2)], h=-1.0, omegas=[1.0],
                       beta=np.array([0.0]))
    b = hm_bem.Batch([p])
    try:
        with pytest.raises(hm_bem.HMError, match="waterplane"):
            b.prepare(hm_bem.HM_GMRES, autotune=False, use_graph=False)
        with pytest.raises(hm_bem.HMError, match="restart"):
            b.set_gmres(restart=0)
    finally:
        b.close()
