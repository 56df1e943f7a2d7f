"""Every named workload at its full size on the GPU.

- single cylinder (1,024 hull + 256 interior waterplane panels, infinite
  depth, irregular-frequency removal): the plain FP64 path against the
  oracle's A, B, F (tests/golden/c1_oracle.npz, made by
  tests/golden/make_config_fixtures.py) at 1e-10, at the first, a middle and
  the last of its 16 frequencies; FP32 LU + FP64 refinement too;
- three cylinders, h = 50 m (4,608 panels, 18 modes, 8 headings): FP32 LU on
  MFMA + FP64 refinement against the FP64 LU, every coefficient reported, the
  whole set within 1e-10 (refinement stops at a last correction below 1e-13);
- 64 random problems (seeded draw): 8 members, one in eight, solved together by
  the batch path against one context each at 1e-13 (same series variant);
- eight cylinders, h = 40 m (16,384 panels): the learned surrogate in the dense
  assembly (default training, cell bound 1e-3) against the plain series,
  A/B/F within 1e-4 at two frequencies, with the share of pairs it took;
- sixteen cylinders, h = 60 m (32,768 panels, 96 modes): the H-matrix form as
  the benchmark runs it (mirrored diagonal blocks, ACA 1e-6 / 1e-7, GMRES(50)
  tol 1e-6) against the dense FP64 LU at four frequencies spanning 0.2-2.5
  rad/s, each output within 1e-4 normwise.
"""
import os

import numpy as np
import pytest

import hm_bem
import meshes
from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _ctx(p, frames=True):
    c = hm_bem.Context()
    c.setup(p, frames=frames)
    return c


def test_c1_full_size_against_the_oracle():
    p = meshes.c1_single_cylinder()
    assert (p.n, p.ni) == (1024, 256)
    ref = np.load(os.path.join(GOLDEN, "c1_oracle.npz"))
    c = _ctx(p)
    try:
        for t, i in enumerate(ref["idx"]):
            w = float(p.omegas[i])
            assert w == float(ref["omegas"][t])
            for prec in (hm_bem.HM_FP64, hm_bem.HM_FP32_IR):
                got = c.solve_frequency(w, prec)
                errs = {k: relerr(got[k], ref[k][t]) for k in KEYS}
                print(f"C1 w={w:.4f} precision {prec}: {errs}", flush=True)
                for k in KEYS:
                    assert errs[k] <= 1e-10, (w, prec, k, errs)
    finally:
        c.close()


def test_c2_fp32_refinement_full_mesh():
    p = meshes.c2_three_cylinders()
    assert (p.n, p.m, p.nbeta) == (4608, 18, 8)
    c = _ctx(p)
    try:
        worst = {}
        for i in (0, 20, 39):
            w = float(p.omegas[i])
            ref = c.solve_frequency(w, hm_bem.HM_FP64)
            got = c.solve_frequency(w, hm_bem.HM_FP32_IR)
            for k in KEYS:
                # per coefficient, relative to the output's largest entry
                per = np.abs(got[k] - ref[k]) / np.max(np.abs(ref[k]))
                worst[k] = max(worst.get(k, 0.0), float(per.max()))
                print(f"C2 w={w:.4f} {k}: largest coefficient error {per.max():.2e} at "
                      f"{np.unravel_index(per.argmax(), per.shape)}, median {np.median(per):.2e}", flush=True)
        for k in KEYS:
            assert worst[k] <= 1e-10, worst
    finally:
        c.close()


def test_c3_batch_members_against_single_contexts():
    probs = meshes.c3_random_problems()
    members = probs[::8]
    assert len(members) == 8
    b = hm_bem.Batch(members)
    try:
        b.prepare(hm_bem.HM_FP64, autotune=True, use_graph=True)
        info = b.info()
        print("C3 batch", info, [m.n for m in members], [m.h for m in members], flush=True)
        ws = [float(members[0].omegas[k]) for k in (3, 20)]
        for w in ws:
            got = b.solve(w)
            for k, prob in enumerate(members):
                c = _ctx(prob, frames=False)          # the batch takes global coordinates
                try:
                    c.set_series_variant(info["series_variant"])
                    ref = c.solve_frequency(w, hm_bem.HM_FP64)
                finally:
                    c.close()
                for key in KEYS:
                    e = relerr(got[k][key], ref[key])
                    assert e <= 1e-13, (w, k, prob.name, key, e)
    finally:
        b.close()


def test_c4_surrogate_full_mesh():
    p = meshes.c4_eight_cylinders()
    assert p.n == 16384
    c = _ctx(p)
    try:
        st = c.surrogate_train(nsamples=16_000_000, epochs=60, batch=4096, lr=2e-3, seed=20261204, quant=True)
        print("C4 surrogate training", st, flush=True)
        for i in (12, 36):
            w = float(p.omegas[i])
            c.set_surrogate(False)
            ref = c.solve_frequency(w, hm_bem.HM_FP64)
            c.set_surrogate(True, 1e-3)
            got = c.solve_frequency(w, hm_bem.HM_FP64)
            counts = c.surrogate_counts()
            errs = {k: relerr(got[k], ref[k]) for k in KEYS}
            print(f"C4 w={w:.4f} surrogate pairs {counts} errors {errs}", flush=True)
            assert counts["network"] > 0
            for k in KEYS:
                assert errs[k] <= 1e-4, (w, k, errs)
    finally:
        c.close()


def test_c5_hmatrix_against_dense_fp64():
    p = meshes.c5_sixteen_cylinders()
    assert (p.n, p.m) == (32768, 96)
    c = _ctx(p)
    try:
        idx = (0, 21, 42, 63)
        ref = {}
        for i in idx:
            ref[i] = c.solve_frequency(float(p.omegas[i]), hm_bem.HM_FP64)
        c.set_gmres(restart=50, tol=1e-6, precond=-1)
        c.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
        for i in idx:
            w = float(p.omegas[i])
            got = c.solve_frequency(w, hm_bem.HM_GMRES)
            gs, hs = c.gmres_stats(), c.hmatrix_stats()
            errs = {k: relerr(got[k], ref[i][k]) for k in KEYS}
            print(f"C5 w={w:.4f} gmres {gs} ranks {hs['mean_rank']:.1f}/{hs['max_rank']} errors {errs}", flush=True)
            assert gs["converged"]
            for k in KEYS:
                assert errs[k] <= 1e-4, (w, k, errs)
    finally:
        c.close()
