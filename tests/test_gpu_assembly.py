"""The per-frequency chain on the GPU (assembly -> LU -> solve -> forces) against
the oracle, through the C-ABI.

Two error measures per output array (added mass, damping, excitation):
  normwise   max|gpu - ref| / max|ref|
  entrywise  max |gpu - ref| / max(|ref|, 1e-6 max|ref|)   (sub-noise entries absolute)
The bar is max(target, 20 x band), band = the largest change of the oracle's own
result when every mesh node moves by one ulp (the reference's conditioning).
Deep water: target 1e-10 normwise; the measured band is ~1e-12 normwise and
~1e-9 entrywise. Finite depth: the reference's epsilon-accelerated Green
function series (R/h in [5e-4, 0.5)) make its own outputs move by ~1e-5
normwise under that perturbation, which bounds any reimplementation; the bar
never exceeds 1e-4 (the mixed-precision target).
"""
import os

import numpy as np
import pytest

import hm_bem
import meshes
from conftest import GOLDEN, entrywise_relerr, relerr

KEYS = ("amas", "bdmp", "exfc")

pytestmark = pytest.mark.gpu


def _deep_cylinder():
    b = meshes.cylinder(5.0, 5.0, 24, 4, 2, waterplane_nr=2)
    return meshes.Problem([b], h=-1.0, omegas=[0.4, 1.5], beta=np.array([0.0, 0.7]))


def _finite_pair():
    bs = [meshes.cylinder(3.0, 4.0, 20, 4, 2, center=(0.0, 0.0)),
          meshes.cylinder(3.0, 4.0, 20, 4, 2, center=(12.0, 0.0))]
    return meshes.Problem(bs, h=30.0, omegas=[0.4, 1.3], beta=np.array([0.0, 1.0]))


def _perturbed(prob, direction):
    """the same problem with every mesh node moved by one ulp"""
    bodies = []
    for b in prob.bodies:
        nb = meshes.Body(xyz=np.nextafter(b.xyz, direction), ncn=b.ncn, ncon=b.ncon, xr=b.xr,
                         ixyz=None if b.ixyz is None else np.nextafter(b.ixyz, direction), incn=b.incn,
                         incon=b.incon)
        bodies.append(nb)
    p = meshes.Problem(bodies, h=prob.h, omegas=prob.omegas, beta=prob.beta, rho=prob.rho, g=prob.g,
                       g_disp=prob.g_disp)
    return p


def _oracle_band(oracle, prob, w, ref, matrices=False):
    """largest normwise / entrywise change of the oracle under +-1 ulp node moves"""
    band = {}
    for d in (np.inf, -np.inf):
        r = oracle.solve_frequency(_perturbed(prob, d), w, want_matrices=matrices)
        for k in KEYS + (("amat", "smat") if matrices else ()):
            nw, ew = band.get(k, (0.0, 0.0))
            band[k] = (max(nw, relerr(r[k], ref[k])), max(ew, entrywise_relerr(r[k], ref[k])))
    return band


def _check(got, ref, band, target_nw, target_ew, cap=1e-4):
    bad = []
    for k in KEYS:
        nw = relerr(got[k], ref[k])
        ew = entrywise_relerr(got[k], ref[k])
        tol_nw = min(max(target_nw, 20.0 * band[k][0]), cap)
        tol_ew = max(target_ew, 20.0 * band[k][1])
        if not (nw <= tol_nw and ew <= tol_ew):
            bad.append((k, nw, tol_nw, ew, tol_ew))
    return bad


@pytest.fixture(scope="module")
def deep():
    p = _deep_cylinder()
    c = hm_bem.Context()
    c.setup(p)
    yield p, c
    c.close()


@pytest.fixture(scope="module")
def finite():
    p = _finite_pair()
    c = hm_bem.Context()
    c.setup(p)
    yield p, c
    c.close()


def test_deep_irr_system_matrices(deep, oracle):
    p, c = deep
    for w in p.omegas:
        r = oracle.solve_frequency(p, w, want_matrices=True)
        c.set_frequency(w)
        c.assemble()
        a, s, rhs = c.get_system()
        assert relerr(a, r["amat"]) < 1e-12
        assert relerr(s, r["smat"]) < 1e-12
        assert relerr(rhs, r["rhs"]) < 1e-12


@pytest.mark.parametrize("precision", [hm_bem.HM_FP64, hm_bem.HM_FP32_IR])
def test_deep_irr_forces(deep, oracle, precision):
    p, c = deep
    for w in p.omegas:
        r = oracle.solve_frequency(p, w)
        band = _oracle_band(oracle, p, w, r)
        g = c.solve_frequency(w, precision)
        bad = _check(g, r, band, 1e-10, 1e-8)
        assert not bad, (w, precision, bad)
        if precision == hm_bem.HM_FP32_IR:
            assert 1 <= g["iters"] < 30


def test_finite_system_and_forces(finite, oracle):
    p, c = finite
    for w in p.omegas:
        r = oracle.solve_frequency(p, w, want_matrices=True)
        band = _oracle_band(oracle, p, w, r, matrices=True)
        c.set_frequency(w)
        c.assemble()
        a, s, rhs = c.get_system()
        assert relerr(a, r["amat"]) <= max(1e-12, 20.0 * band["amat"][0])
        assert relerr(s, r["smat"]) <= max(1e-12, 20.0 * band["smat"][0])
        for prec in (hm_bem.HM_FP64, hm_bem.HM_FP32_IR):
            g = c.solve_frequency(w, prec)
            bad = _check(g, r, band, 1e-10, 1e-8)
            assert not bad, (w, prec, bad)


def test_fp32_refinement_matches_fp64(finite, deep):
    # the mixed-precision solve reproduces the FP64 solve of the same GPU matrices
    for p, c in (finite, deep):
        a = c.solve_frequency(0.9, hm_bem.HM_FP64)
        b = c.solve_frequency(0.9, hm_bem.HM_FP32_IR)
        for k in KEYS:
            assert relerr(b[k], a[k]) < 1e-11
            assert entrywise_relerr(b[k], a[k]) < 1e-8


def test_potentials_solve_the_system(deep):
    # MXPOT satisfies A^T A x = A^T b of the assembled least-squares system
    p, c = deep
    c.set_frequency(1.0)
    c.assemble()
    a, _, rhs = c.get_system()
    c.factor(hm_bem.HM_FP64)
    c.solve()
    x = c.get_potentials()
    lhs = a.T @ (a @ x)
    r = a.T @ rhs
    assert relerr(lhs, r) < 1e-11


def test_config1_expected_outputs_gpu():
    """the reference's own integration case, test/data/config1 (3 x 234 panels,
    h = 70 m), at the 5 printed digits of its expected_output files"""
    p = meshes.reference_config1_from(os.path.join(GOLDEN, "config1_input"))
    c = hm_bem.Context()
    c.setup(p)
    amas, bdmp, exfc = c.run_frequencies(p.omegas, hm_bem.HM_FP64)
    for f in range(3):
        for k in range(6):
            ea = np.loadtxt(os.path.join(GOLDEN, "config1_expected", f"OAMASS{k + 1}.txt"))[9 * f:9 * f + 9]
            eb = np.loadtxt(os.path.join(GOLDEN, "config1_expected", f"ODAMPING{k + 1}.txt"))[9 * f:9 * f + 9]
            q = 0
            for m in range(3):
                for n in range(3):
                    assert np.max(np.abs(amas[f, 6 * m + k, 6 * n:6 * n + 6] - ea[q, 2:])) <= 5e-5 * np.max(np.abs(amas[f])) + 1e-6
                    assert np.max(np.abs(bdmp[f, 6 * m + k, 6 * n:6 * n + 6] - eb[q, 2:])) <= 5e-5 * np.max(np.abs(bdmp[f])) + 1e-6
                    q += 1
            ef = np.loadtxt(os.path.join(GOLDEN, "config1_expected", f"OEXFOR{k + 1}.txt"))[9 * f:9 * f + 9]
            got = np.array([exfc[f, 6 * m + k, hh] for hh in range(3) for m in range(3)])
            scale = np.max(np.abs(exfc[f]))
            assert np.max(np.abs(got.real - ef[:, 2])) <= 2e-5 * scale
            assert np.max(np.abs(got.imag - ef[:, 3])) <= 2e-5 * scale
    c.close()


def test_call_order_errors():
    c = hm_bem.Context()
    with pytest.raises(hm_bem.HMError) as e:
        c.assemble()
    assert e.value.code == -200
    with pytest.raises(hm_bem.HMError) as e:
        c.set_frequency(1.0)
    assert e.value.code == -200
    c.close()


def test_rejects_bad_geometry():
    p = _deep_cylinder()
    c = hm_bem.Context()
    bad = np.array(p.ncon, copy=True)
    bad[3, 2] = 10 ** 6          # vertex beyond the node list
    p.ncon = bad
    with pytest.raises(hm_bem.HMError) as e:
        c.setup(p)
    assert e.value.code == -108
    c.close()


def test_series_variants_agree(finite, oracle):
    """the lane-group kernels of the epsilon-accelerated regimes against the
    one-pair-per-lane kernels and the oracle: same algorithm, partial sums added
    in a different order, so they agree within the reference's own one-ulp band"""
    p, c = finite
    hist = c.pair_histogram()
    assert hist[1].sum() > 0 and hist[2].sum() > 0      # both regimes are exercised
    for w in p.omegas:
        r = oracle.solve_frequency(p, w, want_matrices=True)
        band = _oracle_band(oracle, p, w, r, matrices=True)
        out = {}
        for variant in (0, 1):
            c.set_series_variant(variant)
            c.set_frequency(w)
            c.assemble()
            out[variant] = c.get_system()
            g = c.solve_frequency(w, hm_bem.HM_FP64)
            bad = _check(g, r, band, 1e-10, 1e-8)
            assert not bad, (w, variant, bad)
        for k, name in ((0, "amat"), (1, "smat")):
            assert relerr(out[1][k], out[0][k]) <= max(1e-12, 40.0 * band[name][0]), (w, name)
            assert relerr(out[1][k], r[name]) <= max(1e-12, 20.0 * band[name][0]), (w, name)
    c.set_series_variant(1)
