"""Motion response on the GPU (hm_solve_motion: SolveMotionMulti,
SolveMotionMulti.f90:46-101) against the oracle's restatement
(hmo_solve_motion), on the added mass, damping and excitation of an actual
two-body solve with body mass, hydrostatic restoring and mooring matrices of
the same bodies.

Tolerance: the linear solve 1e-12 of the largest motion (both sides solve the
same 12 x 12 system by partial-pivoting LU; their rounding differs only in the
order of the updates). With quadratic damping both sides run the same fixed
point to the same stopping test (relative change 1e-6), so the iteration
counts agree and the motions within 1e-9."""
import numpy as np
import pytest

import hm_bem
import meshes

pytestmark = pytest.mark.gpu


def _body_matrices(nb, r=3.0, draft=4.0, rho=1025.0, g=9.81):
    vol = np.pi * r * r * draft
    mass = rho * vol
    matx = np.zeros((nb, 6, 6))
    crs = np.zeros((nb, 6, 6))
    kstf = np.zeros((nb, 6, 6))
    bqdr = np.zeros((nb, 6, 6))
    for b in range(nb):
        matx[b] = np.diag([mass, mass, mass, mass * 4.0, mass * 4.0, mass * 4.5])
        matx[b, 0, 4] = matx[b, 4, 0] = -mass * 1.0      # centre of gravity 1 m below the origin
        matx[b, 1, 3] = matx[b, 3, 1] = mass * 1.0
        crs[b, 2, 2] = rho * g * np.pi * r * r
        crs[b, 3, 3] = crs[b, 4, 4] = rho * g * vol * 0.8
        kstf[b] = np.diag([3e4, 3e4, 0.0, 0.0, 0.0, 5e4])
        bqdr[b] = np.diag([2e4, 2e4, 4e4, 1e5, 1e5, 1e4])
    return matx, crs, kstf, bqdr


@pytest.mark.parametrize("h", [40.0, -1.0], ids=["finite", "deep"])
def test_motion_matches_oracle(oracle, h):
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0 * i, 0.0)) for i in range(2)]
    p = meshes.Problem(bs, h=h, omegas=[0.9], beta=np.array([0.0, np.pi / 3]))
    w = 0.9
    matx, crs, kstf, bqdr = _body_matrices(p.nbody)
    blnr = 0.02 * bqdr
    ctx = hm_bem.Context()
    try:
        ctx.setup(p)
        r = ctx.solve_frequency(w, hm_bem.HM_FP64)
        out = []
        for ib in range(p.nbeta):
            f = r["exfc"][:, ib]
            lin = ctx.solve_motion(w, r["amas"], r["bdmp"], f, matx, crs, kstf, blnr=blnr)
            qd = ctx.solve_motion(w, r["amas"], r["bdmp"], f, matx, crs, kstf, blnr=blnr, bqdr=bqdr)
            out.append((f, lin, qd))
    finally:
        ctx.close()
    for ib, (f, lin, qd) in enumerate(out):
        xl, il, infl = oracle.solve_motion(w, r["amas"], r["bdmp"], f, matx, crs, kstf, blnr=blnr)
        xq, iq, infq = oracle.solve_motion(w, r["amas"], r["bdmp"], f, matx, crs, kstf, blnr=blnr, bqdr=bqdr)
        assert infl == 0 and infq == 0
        el = np.max(np.abs(lin[0] - xl)) / np.max(np.abs(xl))
        eq = np.max(np.abs(qd[0] - xq)) / np.max(np.abs(xq))
        print("h", h, "heading", ib, "linear", el, "quadratic", eq, "iterations", qd[1], iq, flush=True)
        assert lin[1] == il == 0
        assert el < 1e-12
        assert qd[1] == iq and iq >= 2
        assert eq < 1e-9


def test_motion_argument_errors():
    """rejected on the host, before any device work"""
    ctx = hm_bem.Context()
    try:
        z = np.zeros((1, 6, 6)) + np.eye(6)
        with pytest.raises(hm_bem.HMError, match="w1 must be"):
            ctx.solve_motion(-1.0, np.eye(6), np.eye(6), np.ones(6, complex), z, z)
        with pytest.raises(hm_bem.HMError, match="nbody"):
            ctx.solve_motion(0.5, np.eye(6), np.eye(6), np.ones(6, complex), np.zeros((0, 6, 6)),
                             np.zeros((0, 6, 6)))
    finally:
        ctx.close()
