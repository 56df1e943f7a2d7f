"""The oracle (oracle/hm_oracle.c) pinned before it is trusted.

How each oracle function is pinned:
  special functions   golden vectors of the reference build (bit-level) + the
                      known-answer tables in FinGrnExtSubs.f90's comments
  Limes (Wynn eps)    golden vectors + exact on an alternating series with a
                      closed-form limit (ln 2)
  DISPERSION          golden vectors + residual of both dispersion equations
  FINGREEN3D          golden vectors over all four R/h regimes (<=1e-12) +
                      continuity across the regime boundaries + reciprocity
  INFGREEN3D          golden vectors (<=1e-12)
  SGLINTBD            golden vectors + closed form of a square at its centroid
                      + far-field limit + brute-force Gauss-Legendre + finite
                      differences of the potential for the gradient
  assembly/LU/forces  the reference's own config1 expected outputs (5 printed
                      digits), A/B symmetry, LU residual
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, relerr

G = 9.80665


# ---- special functions ----------------------------------------------------------
def test_special_functions_golden(oracle, golden):
    g = golden("special")
    for x, k in zip(g["x"], g["k01"]):
        assert oracle.ik01a(x) == pytest.approx(tuple(k), rel=1e-14)
    for x, j in zip(g["x"], g["jy"]):
        np.testing.assert_allclose(oracle.jy01b(x), j, rtol=1e-14, atol=1e-300)
    for x, e in zip(g["x"], g["e1"]):
        assert oracle.e1xa(x + 1e-3) == pytest.approx(e, rel=1e-14)
    for x, e in zip(g["x"], g["ei"]):
        assert oracle.eix(x + 1e-3) == pytest.approx(e, rel=1e-14)
    for x, row in zip(g["x"], g["en"]):
        for n in range(6):
            assert oracle.enxa(n, x + 1e-3) == pytest.approx(row[n], rel=1e-13)
    for x, e in zip(g["xe"], g["erf"]):
        assert oracle.error(x) == pytest.approx(e, rel=1e-14, abs=1e-300)
    for s, n, lim in zip(g["lim_s"], g["lim_n"], g["lim"]):
        assert oracle.limes(s[: n + 1]) == pytest.approx(lim, rel=1e-14)


def test_special_function_known_answers(oracle):
    # tables in the source comments, FinGrnExtSubs.f90:52-59, :124-132, :187-195, :243-251
    for x, v in [(1.0, .84270079), (2.0, .99532227), (3.0, .99997791), (4.0, .99999998), (5.0, 1.0)]:
        assert oracle.error(x) == pytest.approx(v, abs=1e-8)
    for n, v in enumerate([.45399930e-05, .41569689e-05, .38302405e-05, .35487626e-05, .33041014e-05,
                           .30897289e-05]):
        assert oracle.enxa(n, 10.0) == pytest.approx(v, rel=2e-5)   # recurrence on a 1e-7-class E1 fit
    for x, v in [(1.0, .2193839), (2.0, .4890051e-01), (3.0, .1304838e-01), (4.0, .3779352e-02),
                 (5.0, .1148296e-02)]:
        assert oracle.e1xa(x) == pytest.approx(v, rel=1e-6)
    for x, v in [(1.0, 1.8951178), (2.0, 4.9542344), (3.0, 9.9338326), (4.0, 19.630874), (5.0, 40.185275)]:
        assert oracle.eix(x) == pytest.approx(v, rel=1e-7)


def test_bessel_against_closed_forms(oracle):
    # K0/K1 are the Zhang-Jin series/asymptotics: full FP64 against scipy
    from scipy import special as sp
    for x in (0.01, 0.3, 1.0, 5.0, 9.0, 12.0, 20.0, 40.0):
        k0, k1 = oracle.ik01a(x)
        assert k0 == pytest.approx(sp.k0(x), rel=1e-13)
        assert k1 == pytest.approx(sp.k1(x), rel=1e-13)
    # J/Y are 1e-8-class polynomial fits (JY01B), so only that accuracy is pinned here
    for x in (0.5, 2.0, 3.9, 4.1, 10.0, 30.0):
        j0, j1, y0, y1 = oracle.jy01b(x)
        assert j0 == pytest.approx(sp.j0(x), abs=1e-8)
        assert y1 == pytest.approx(sp.y1(x), abs=1e-8)


def test_limes_alternating_series(oracle):
    # partial sums of sum (-1)^k/(k+1) -> ln 2; Wynn epsilon reaches it to 1e-12
    s = np.cumsum((-1.0) ** np.arange(30) / (1.0 + np.arange(30)))
    assert oracle.limes(s) == pytest.approx(np.log(2.0), abs=1e-12)


# ---- dispersion ----------------------------------------------------------------------
def test_dispersion_golden_and_residual(oracle, golden):
    g = golden("dispersion")
    for a, w in enumerate(g["w"]):
        for b, h in enumerate(g["h"]):
            wvn = oracle.dispersion(w, h)
            np.testing.assert_array_equal(wvn, g["wvn"][a, b])
            v = w * w / 9.807
            k = wvn[0]
            assert k * np.tanh(k * h) == pytest.approx(v, rel=1e-9)
            km = wvn[1:]
            np.testing.assert_allclose(km * np.tan(km * h), -v, rtol=1e-7, atol=1e-9)


# ---- Green functions -------------------------------------------------------------------
def test_fingreen3d_golden(oracle, golden):
    g = golden("green_finite")
    worst = 0.0
    for k in range(len(g["w"])):
        w, h = g["w"][k], g["h"][k]
        wvn = oracle.dispersion(w, h)
        got = oracle.fingreen3d(g["src"][k], g["fld"][k], w * w / G, wvn, h, int(g["flag"][k]))
        worst = max(worst, relerr(got, g["grn"][k]))
    assert worst < 1e-12


def test_infgreen3d_golden(oracle, golden):
    g = golden("green_infinite")
    worst = 0.0
    for k in range(len(g["w"])):
        got = oracle.infgreen3d(g["src"][k], g["fld"][k], g["w"][k] ** 2 / G, int(g["flag"][k]))
        worst = max(worst, relerr(got, g["grn"][k]))
    assert worst < 1e-12


def test_fingreen3d_regime_continuity(oracle):
    # the eigenfunction, epsilon-accelerated and Pidcock series agree where the
    # regime changes at R/h = 0.05 and 0.5. (The reference is not continuous at
    # its Linton boundary R/h = 5e-4 -- the compiled reference differs there by
    # up to O(1) -- so that branch is pinned by the golden vectors instead.)
    for h, w in ((50.0, 1.0), (50.0, 0.3), (70.0, 2.0)):
        wvn = oracle.dispersion(w, h)
        for rb in (0.05, 0.5):
            lo = oracle.fingreen3d((rb * h * (1 - 1e-9), 0, -10.0), (0, 0, -12.0), w * w / G, wvn, h, 1)
            hi = oracle.fingreen3d((rb * h * (1 + 1e-9), 0, -10.0), (0, 0, -12.0), w * w / G, wvn, h, 1)
            assert relerr(lo, hi) < 1e-3


def test_fingreen3d_reciprocity(oracle):
    # G(P,Q) = G(Q,P) for the wave Green function (value), any regime
    h, w = 40.0, 0.8
    wvn = oracle.dispersion(w, h)
    rng = np.random.default_rng(3)
    for rr in (0.0, 0.1, 1.0, 5.0, 30.0):
        a = np.array([0.0, 0.0, -rng.uniform(1, 30)])
        b = np.array([rr, 0.0, -rng.uniform(1, 30)])
        g1 = oracle.fingreen3d(a, b, w * w / G, wvn, h, 0)
        g2 = oracle.fingreen3d(b, a, w * w / G, wvn, h, 0)
        assert abs(g1[0] - g2[0]) / abs(g1[0]) < 1e-4   # the series stop at 1e-6 absolute


# ---- Rankine panel integrals ---------------------------------------------------------
def test_sglintbd_golden(oracle, golden):
    g = golden("rankine")
    for k in range(len(g["ncn"])):
        s, d = oracle.sglintbd(g["verts"][k][: g["ncn"][k]], g["xf"][k])
        assert s == pytest.approx(g["sij"][k], rel=1e-13, abs=1e-13)
        np.testing.assert_allclose(d, g["dij"][k], rtol=1e-12, atol=1e-12)


def _square(a):
    return np.array([[0, 0, 0], [a, 0, 0], [a, a, 0], [0, a, 0]], dtype=float)


def test_sglintbd_closed_form_and_limits(oracle):
    a = 1.7
    q = _square(a)
    s, _ = oracle.sglintbd(q, q.mean(0))
    assert abs(s) == pytest.approx(4 * a * np.arcsinh(1.0), rel=1e-12)
    far = q.mean(0) + np.array([300.0, -200.0, 150.0])
    s, _ = oracle.sglintbd(q, far)
    assert abs(s) == pytest.approx(a * a / np.linalg.norm(far - q.mean(0)), rel=1e-5)


def test_sglintbd_brute_force_and_gradient(oracle):
    a = 1.3
    q = _square(a)
    x, w = np.polynomial.legendre.leggauss(64)
    u = 0.5 * a * (x + 1)
    wu = 0.5 * a * w
    for xf in (np.array([0.4, 0.9, 0.35]), np.array([2.0, -0.5, -0.8]), np.array([0.65, 0.65, 3.0])):
        r = np.sqrt((u[:, None] - xf[0]) ** 2 + (u[None, :] - xf[1]) ** 2 + xf[2] ** 2)
        ref = np.sum(wu[:, None] * wu[None, :] / r)
        s, d = oracle.sglintbd(q, xf)
        assert abs(s) == pytest.approx(ref, rel=1e-9)
        # gradient of the integral w.r.t. the field point, by central differences
        eps = 1e-5
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            sp_, _ = oracle.sglintbd(q, xf + e)
            sm_, _ = oracle.sglintbd(q, xf - e)
            fd = (sp_ - sm_) / (2 * eps)
            assert abs(abs(d[k]) - abs(fd)) < 1e-6 * max(1.0, abs(fd))


# ---- panel geometry -----------------------------------------------------------------------
def test_panel_geometry_matches_numpy_restatement(oracle):
    import meshes
    b = meshes.cylinder(4.0, 3.0, 16, 3, 2)
    mine = meshes.panel_geometry(b.xyz, b.ncn, b.ncon, b.xr)
    ref = oracle.panel_geometry(b.xyz, b.ncn, b.ncon, b.xr)
    for x, y in zip(mine, ref):
        np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-14)
    # wetted area of the cylinder converges to 2*pi*r*T + pi*r^2 from below
    assert mine[1].sum() == pytest.approx(2 * np.pi * 4 * 3 + np.pi * 16, rel=0.05)


# ---- LU -------------------------------------------------------------------------------------
def test_textbook_lu_residual(oracle):
    rng = np.random.default_rng(9)
    n = 60
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
    x, info = oracle.zgetrf_solve(a, b)
    assert info == 0
    assert np.max(np.abs(a @ x - b)) / (np.max(np.abs(a)) * np.max(np.abs(x))) < 1e-13


# ---- end to end on the reference's own test case -----------------------------------------------
def _load_expected(kind, k):
    return np.loadtxt(os.path.join(GOLDEN, "config1_expected", f"{kind}{k}.txt"))


@pytest.mark.slow
def test_config1_expected_outputs(oracle):
    """test/data/config1 (3 bodies x 234 panels, h = 70 m): added mass, damping and
    excitation of the oracle against the reference's expected_output files, which
    print 5 significant digits (E14.5)."""
    import meshes
    p = meshes.reference_config1_from(os.path.join(GOLDEN, "config1_input"))
    r = oracle.solve_frequency(p, 0.05)
    assert r["info"] == 0
    amas, bdmp, exfc = r["amas"], r["bdmp"], r["exfc"]
    for k in range(6):
        ea = _load_expected("OAMASS", k + 1)[:9]
        eb = _load_expected("ODAMPING", k + 1)[:9]
        q = 0
        for m in range(3):
            for n in range(3):
                got_a = amas[6 * m + k, 6 * n:6 * n + 6]
                got_b = bdmp[6 * m + k, 6 * n:6 * n + 6]
                assert np.max(np.abs(got_a - ea[q, 2:])) <= 5e-5 * np.max(np.abs(amas)) + 1e-6
                assert np.max(np.abs(got_b - eb[q, 2:])) <= 5e-5 * np.max(np.abs(bdmp)) + 1e-6
                q += 1
        ef = _load_expected("OEXFOR", k + 1)[:9]      # rows: heading-major, body-minor
        got = np.array([exfc[6 * m + k, hh] for hh in range(3) for m in range(3)])
        np.testing.assert_allclose(got.real, ef[:, 2], rtol=2e-5, atol=2e-5 * np.max(np.abs(exfc)))
        np.testing.assert_allclose(got.imag, ef[:, 3], rtol=2e-5, atol=2e-5 * np.max(np.abs(exfc)))
    # reciprocity of the radiation coefficients (to the discretisation error)
    assert relerr(amas, amas.T) < 1e-2
    assert relerr(bdmp, bdmp.T) < 1e-2


@pytest.mark.parametrize("h", [31.7, -1.0])
def test_multibody_chain_against_the_reference(oracle, h):
    """The oracle's assembled system and force reductions against the
    reference's own multi-body code (oracle/_ref, ref_shim_multi.f90): its
    CALGREEN_MULTI table fill (CalGreenFuncMulti.f90:112-183, with
    SGLINTBD_MULTI and FINGREEN3D/INFGREEN3D), its element integration
    BODINT_MULTI_LEFT / RBC_MULTI_RIGHT / DBC_MULTI_RIGHT with DINP
    (BodyIntgrMulti.f90:48-173, PatclVelct.f90:83) in the loops of
    ASSB_LEFT/RBC/DBC_MULTI, and its RFORCE_MULTI / EFORCE_MULTI
    (PotentWavForceMulti.f90:45-226) fed with the oracle's potentials. Two
    cylinders 12 m apart: PIDCOCK, EIGENE and EIGEN pairs at h = 31.7 m, and
    infinite depth. The reference stores its tables at GRN_KIND = 4 (FP32,
    WavDynMods.f90:47), so A and the right-hand sides agree to that storage
    (measured ~1e-8 of the largest entry; bar 2e-7); the force reductions read
    the potentials in FP64 and agree to rounding (bar 1e-13)."""
    import meshes
    from oracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    R = Reference()
    bs = [meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0 * i, 0.0)) for i in range(2)]
    p = meshes.reference_parity_constants(meshes.Problem(bs, h=h, omegas=[0.9], beta=np.array([0.0, np.pi / 3])))
    w = 0.9
    r = oracle.solve_frequency(p, w, want_matrices=True)
    assert r["info"] == 0
    R.multi_set(p)
    v = w * w / p.g
    wvn = R.dispersion(w, h) if h > 0 else np.zeros(200)
    R.multi_freq(v, wvn[0] if h > 0 else v, w, wvn)
    rk, cg = R.multi_calgreen()
    assert np.all(np.isfinite(cg)) and np.max(np.abs(cg)) > 0
    n, m = p.n, p.m
    a = R.multi_left()
    assert relerr(r["amat"][:n], a) <= 2e-7
    br = R.multi_rbc()
    assert relerr(r["rhs"][:n, :m], br) <= 2e-7
    for ib, beta in enumerate(p.beta):
        bd = R.multi_dbc(beta)
        assert relerr(r["rhs"][:n, m + ib], bd) <= 2e-7
        mx = np.column_stack([r["mxpot"][:, :m], r["mxpot"][:, m + ib]])
        amas, bdmp, exfc = R.multi_forces(mx, beta)
        assert relerr(r["amas"], amas) <= 1e-13
        assert relerr(r["bdmp"], bdmp) <= 1e-13
        assert relerr(r["exfc"][:, ib], exfc) <= 1e-13


def _example5_expected(kind, k):
    return np.loadtxt(os.path.join(GOLDEN, "example5_expected", f"{kind}{k}.txt"))


@pytest.mark.slow
@pytest.mark.parametrize("f", [0, 9])
def test_example5_irr_expected_outputs(oracle, f):
    """examples/example5_single_body of the reference (234 hull + 383 interior
    waterplane panels, h = 70 m, irregular-frequency removal on, 36 headings),
    read from its input directory: the oracle's whole IRR chain (waterplane
    rows of A, the normal equations C = A^T A and D = A^T B, the solve and the
    force reductions) against the case's Hams_format files, which print 5
    significant digits for A and B (E14.5) and 6 for F (E14.6); the bar is
    that of the config1 case, 5e-5 of the largest entry."""
    import formats
    p, cf = formats.problem_from_control(os.path.join(GOLDEN, "example5_input"))
    assert (p.n, p.ni, cf.irsp, p.nbeta) == (234, 383, 1, 36)
    w = float(p.omegas[f])
    r = oracle.solve_frequency(p, w)
    assert r["info"] == 0
    amas, bdmp, exfc = r["amas"], r["bdmp"], r["exfc"]
    for k in range(6):
        ea = _example5_expected("OAMASS", k + 1)[f]
        eb = _example5_expected("ODAMPING", k + 1)[f]
        assert abs(ea[1] - w) < 5e-4
        assert np.max(np.abs(amas[k] - ea[2:])) <= 5e-5 * np.max(np.abs(amas)) + 1e-6, (k, amas[k], ea[2:])
        assert np.max(np.abs(bdmp[k] - eb[2:])) <= 5e-5 * np.max(np.abs(bdmp)) + 1e-6, (k, bdmp[k], eb[2:])
        ef = _example5_expected("OEXFOR", k + 1)[36 * f:36 * (f + 1)]
        np.testing.assert_allclose(exfc[k].real, ef[:, 2], rtol=2e-5, atol=2e-5 * np.max(np.abs(exfc)))
        np.testing.assert_allclose(exfc[k].imag, ef[:, 3], rtol=2e-5, atol=2e-5 * np.max(np.abs(exfc)))


def test_field_points_reproduce_the_hull_potentials(oracle):
    """hmo_field_points (CalPotentialMulti) pinned by what the boundary-integral
    representation fixes: at a hull panel's centroid it is the collocation
    equation the solve satisfied, so it returns that panel's solved potential
    for the radiation columns (the 2 pi free term; 4 pi would give half) and the
    solved scattered potential plus the incident one for the diffraction
    columns; pressure and elevation are i w rho phi and i w phi / g."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__)))
    import meshes
    bs = [meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0 * i, 0.0)) for i in range(2)]
    p = meshes.Problem(bs, h=40.0, omegas=[0.9], beta=np.array([0.0, np.pi / 3]))
    j = 7
    xfp = np.array([p.xyz_p[j], p.xyz_p[p.n - 3]])
    r = oracle.field_points(p, 0.9, xfp)
    s = oracle.solve_frequency(p, 0.9, want_matrices=True)
    assert r["info"] == 0 and s["info"] == 0
    mx, m = s["mxpot"], 6 * p.nbody
    for k, jj in enumerate((j, p.n - 3)):
        np.testing.assert_allclose(r["pot"][k, :m], mx[jj, :m], rtol=1e-10)
        # diffraction: scattered + incident; the incident part alone is the difference
        inc = r["pot"][k, m:] - mx[jj, m:]
        assert np.all(np.abs(inc) > 1e-3 * np.max(np.abs(mx[:, m:])))
    np.testing.assert_allclose(r["prs"], 1j * 0.9 * p.rho * r["pot"], rtol=1e-14)
    np.testing.assert_allclose(r["elv"], 1j * 0.9 / p.g * r["pot"], rtol=1e-14)


def _motion_system(nb, rng):
    """a plausible multi-body system: symmetric added mass and damping with
    cross-body coupling, body masses, hydrostatics and a weak mooring"""
    m = 6 * nb
    q = rng.normal(size=(m, m))
    amas = 2e4 * (q @ q.T / m + np.eye(m))
    q = rng.normal(size=(m, m))
    bdmp = 3e3 * (q @ q.T / m + 0.5 * np.eye(m))
    matx = np.zeros((nb, 6, 6))
    crs = np.zeros((nb, 6, 6))
    kstf = np.zeros((nb, 6, 6))
    for b in range(nb):
        matx[b] = np.diag([1.16e5, 1.16e5, 1.16e5, 4e5, 4e5, 5e5])
        crs[b] = np.diag([0.0, 0.0, 2.84e5, 6e5, 6e5, 0.0])
        kstf[b] = np.diag([2e3, 2e3, 0.0, 0.0, 0.0, 3e3])
    exfc = 1e4 * (rng.normal(size=m) + 1j * rng.normal(size=m))
    return amas, bdmp, matx, crs, kstf, exfc


def _global(loc):
    nb = loc.shape[0]
    g = np.zeros((6 * nb, 6 * nb))
    for b in range(nb):
        g[6 * b:6 * b + 6, 6 * b:6 * b + 6] = loc[b]
    return g


def test_solve_motion_closed_forms(oracle):
    """hmo_solve_motion (SolveMotionMulti) pinned by: a single degree of freedom
    at resonance, |x| = |F| / (w B); the coupled multi-body system against
    numpy's solve of the block-diagonal assembly (the body matrices' Fortran
    layout, the sign convention RIGHT = -i conj(F)); and, with quadratic
    damping, the converged x satisfying its own fixed-point equation to the
    stopping tolerance"""
    rng = np.random.default_rng(21)
    # one body, heave only: resonance of K - w^2 (M + A)
    nb, w = 1, 0.8
    amas, bdmp = np.zeros((6, 6)), np.zeros((6, 6))
    matx, crs = np.zeros((1, 6, 6)), np.zeros((1, 6, 6))
    for k in range(6):
        matx[0, k, k] = 1.0
        crs[0, k, k] = 1.0 if k != 2 else 0.0
        bdmp[k, k] = 0.1
    amas[2, 2] = 5.0e4
    matx[0, 2, 2] = 1.0e5
    crs[0, 2, 2] = w * w * (1.5e5)
    bdmp[2, 2] = 2.0e4
    exfc = np.zeros(6, dtype=np.complex128)
    exfc[2] = 3.0e4 + 4.0e4j
    x, it, info = oracle.solve_motion(w, amas, bdmp, exfc, matx, crs)
    assert info == 0 and it == 0
    assert abs(abs(x[2]) - abs(exfc[2]) / (w * 2.0e4)) < 1e-12 * abs(x[2])
    # coupled bodies against numpy
    nb, w = 3, 0.7
    amas, bdmp, matx, crs, kstf, exfc = _motion_system(nb, rng)
    x, it, info = oracle.solve_motion(w, amas, bdmp, exfc, matx, crs, kstf)
    lhs = -w * w * (_global(matx) + amas) + 1j * w * bdmp + _global(crs) + _global(kstf)
    ref = np.linalg.solve(lhs, -1j * np.conj(exfc))
    assert info == 0 and it == 0
    assert np.max(np.abs(x - ref)) < 1e-12 * np.max(np.abs(ref))
    # quadratic damping: the fixed point
    bqdr = np.zeros((nb, 6, 6))
    for b in range(nb):
        bqdr[b] = np.diag([4e3, 4e3, 8e3, 1e4, 1e4, 2e3])
    x, it, info = oracle.solve_motion(w, amas, bdmp, exfc, matx, crs, kstf, bqdr=bqdr)
    assert info == 0 and it >= 2
    vd = _global(bqdr) * w * np.abs(x)[None, :]
    res = (lhs + 1j * w * vd) @ x + 1j * np.conj(exfc)
    assert np.linalg.norm(res) < 1e-5 * np.linalg.norm(exfc)
