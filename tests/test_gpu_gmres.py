"""Restarted GMRES on the GPU (FP32 MFMA products, FP64 residual restarts,
optional block-diagonal preconditioner) against the FP64 LU solve of the same
assembled system and against the oracle.

Bar: GMRES stops when ||B - A X|| <= tol ||B|| (tol = 1e-6) in every column, so
the potentials satisfy that bound exactly (checked on the host with the
assembled matrices) and added mass, damping and excitation agree with the FP64
LU path to ~cond * tol; 2e-5 is used, below the 1e-4 target of the accelerated
path.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _array(nx, ny, h, spacing=12.0, n_theta=16, n_z=4, n_r=2):
    bs = [meshes.cylinder(3.0, 4.0, n_theta, n_z, n_r, center=(spacing * i, spacing * j))
          for j in range(ny) for i in range(nx)]
    return meshes.Problem(bs, h=h, omegas=[0.5, 1.6], beta=np.array([0.0, np.pi / 2]))


@pytest.fixture(scope="module")
def quad():
    p = _array(2, 2, 40.0)
    c = hm_bem.Context()
    c.setup(p)
    yield p, c
    c.close()


def _residuals(c, w, precision):
    c.set_frequency(w)
    c.assemble()
    a, _, rhs = c.get_system()
    c.factor(precision)
    c.solve()
    x = c.get_potentials()
    r = rhs - a @ x
    return np.linalg.norm(r, axis=0) / np.linalg.norm(rhs, axis=0)


@pytest.mark.parametrize("precond", [-1, 0, 1, 2])
def test_gmres_residual_and_forces(quad, precond):
    p, c = quad
    c.set_gmres(restart=50, tol=1e-6, maxiter=500, precond=precond)
    for w in p.omegas:
        ref = c.solve_frequency(w, hm_bem.HM_FP64)
        got = c.solve_frequency(w, hm_bem.HM_GMRES)
        st = c.gmres_stats()
        assert st["converged"] and st["rel_residual"] <= 1e-6, st
        assert 1 <= st["iterations"] <= 100, st
        for k in KEYS:
            assert relerr(got[k], ref[k]) < 2e-5, (w, precond, k, relerr(got[k], ref[k]))
        res = _residuals(c, w, hm_bem.HM_GMRES)
        assert np.max(res) <= 1.05e-6, (w, precond, np.max(res))


def test_block_preconditioner_saves_iterations(quad):
    p, c = quad
    its = {}
    for pc in (0, 1):
        c.set_gmres(restart=50, tol=1e-6, precond=pc)
        c.solve_frequency(1.6, hm_bem.HM_GMRES)
        its[pc] = c.gmres_stats()["iterations"]
    assert its[1] <= its[0], its


def test_explicit_block_inverse_matches_triangular_solves(quad):
    """The preconditioner applied as the blocks' explicit FP32 inverses (one
    batched GEMM a Krylov step, precond 2 and the default) against the blocks'
    LU triangular solves (precond 1), dense and H-matrix: both are FP32
    approximations of the same M^-1, so GMRES takes the same number of steps
    (within one) and converges to the same solution. (The H-matrix form's FP16
    application of the inverses, its default, is held to its FP32 one in
    tests/test_gpu_fp16.py.)"""
    p, c = quad
    c.set_precond_fp16(False)
    try:
        for hmat in (False, True):
            c.set_hmatrix(hmat, eps=1e-7, eps_s=1e-8, kmax=50)
            got, its = {}, {}
            for pc in (1, 2):
                c.set_gmres(restart=50, tol=1e-9, maxiter=500, precond=pc)
                got[pc] = c.solve_frequency(1.1, hm_bem.HM_GMRES)
                st = c.gmres_stats()
                assert st["converged"], (hmat, pc, st)
                its[pc] = st["iterations"]
            assert abs(its[1] - its[2]) <= 1, (hmat, its)
            for k in KEYS:
                assert relerr(got[2][k], got[1][k]) < 1e-7, (hmat, k, relerr(got[2][k], got[1][k]))
    finally:
        c.set_precond_fp16(True)
        c.set_hmatrix(False, eps=1e-5, eps_s=1e-6, kmax=50)
        c.set_gmres()


def test_short_restart_cycles(quad):
    # restart 4 forces several cycles, each starting from the FP64 residual
    p, c = quad
    c.set_gmres(restart=4, tol=1e-6, maxiter=800, precond=0)
    ref = c.solve_frequency(1.6, hm_bem.HM_FP64)
    got = c.solve_frequency(1.6, hm_bem.HM_GMRES)
    st = c.gmres_stats()
    assert st["converged"] and st["cycles"] >= 2, st
    for k in KEYS:
        assert relerr(got[k], ref[k]) < 2e-5
    c.set_gmres()


def test_not_converged_is_reported(quad):
    p, c = quad
    c.set_gmres(restart=2, tol=1e-12, maxiter=3, precond=0)
    c.solve_frequency(1.6, hm_bem.HM_GMRES)
    st = c.gmres_stats()
    assert not st["converged"] and st["iterations"] == 3, st
    c.set_gmres()


def test_gmres_deep_irr_single_body(oracle):
    # one body with irregular-frequency rows: GMRES on the normal equations
    b = meshes.cylinder(5.0, 5.0, 24, 4, 2, waterplane_nr=2)
    p = meshes.Problem([b], h=-1.0, omegas=[0.4, 1.5], beta=np.array([0.0, 0.7]))
    c = hm_bem.Context()
    c.setup(p)
    c.set_gmres(restart=50, tol=1e-6)
    for w in p.omegas:
        r = oracle.solve_frequency(p, w)
        g = c.solve_frequency(w, hm_bem.HM_GMRES)
        assert c.gmres_stats()["converged"]
        for k in KEYS:
            assert relerr(g[k], r[k]) < 2e-5, (w, k, relerr(g[k], r[k]))
    c.close()


def test_gmres_argument_errors():
    c = hm_bem.Context()
    for kw, code in ((dict(restart=0), -102), (dict(restart=129), -102), (dict(tol=0.0), -103),
                     (dict(tol=1.5), -103), (dict(maxiter=0), -104), (dict(precond=3), -105)):
        with pytest.raises(hm_bem.HMError) as e:
            c.set_gmres(**kw)
        assert e.value.code == code, kw
    c.close()


# ---- ZITER_SOLVE_MULTI / HMATRIX_BUILD on an assembled host matrix ----------
def _assembled(c, w):
    c.set_frequency(w)
    c.assemble()
    a, _, rhs = c.get_system()
    return a, rhs


def test_ziter_dense_with_blocks(quad):
    p, c = quad
    c.set_gmres(restart=50, tol=1e-8, maxiter=500, precond=-1)
    a, rhs = _assembled(c, 0.9)
    m = p.n // p.nbody
    x, info = c.ziter_solve(1, a, rhs, block_sizes=[m] * p.nbody)
    assert info == 0
    r = np.linalg.norm(rhs - a @ x, axis=0) / np.linalg.norm(rhs, axis=0)
    assert np.all(r <= 1e-8), r
    x0, info0 = c.ziter_solve(1, a, rhs)                 # no preconditioner
    assert info0 == 0
    assert relerr(x0, x) < 1e-6


def test_ziter_hmatrix(quad):
    p, c = quad
    m = p.n // p.nbody
    c.set_gmres(restart=50, tol=1e-9, maxiter=500, precond=-1)
    c.set_hmatrix(False, eps=1e-9, eps_s=1e-9, kmax=m)   # ACA parameters of the build below
    a, rhs = _assembled(c, 1.3)
    c.hmatrix_build(a, [m] * p.nbody)
    st = c.hmatrix_stats()
    assert st["unconverged"] == 0 and st["mean_rank"] < m, st
    rng = np.random.default_rng(3)
    v = rng.standard_normal(p.n) + 1j * rng.standard_normal(p.n)
    hv = c.hmatrix_apply(0, v)[:, 0]                      # HMATVEC
    assert np.linalg.norm(hv - a @ v) <= 1e-7 * np.linalg.norm(a @ v)
    x, info = c.ziter_solve(2, None, rhs)
    assert info == 0
    xd = np.linalg.solve(a, rhs)
    assert relerr(x, xd) < 1e-6
    c.hmatrix_release()
    with pytest.raises(hm_bem.HMError, match="before hm_hmatrix_build"):
        c.ziter_solve(2, None, rhs)


def test_ziter_hmatrix_after_solve_needs_a_new_assembly(quad):
    """hm_ziter_solve overwrites the context's right-hand sides and solution, so
    after hm_solve -> hm_ziter_solve(2) the forces are refused (HM_ERR_STATE)
    instead of being reduced from the GMRES solution of the host system; a new
    assembly restores the chain and its forces"""
    p, c = quad
    m = p.n // p.nbody
    c.set_gmres(restart=50, tol=1e-9, maxiter=500, precond=-1)
    c.set_hmatrix(False, eps=1e-9, eps_s=1e-9, kmax=m)
    w = 1.1
    ref = c.solve_frequency(w, hm_bem.HM_FP64)
    a, rhs = _assembled(c, w)
    c.factor(hm_bem.HM_FP64)
    c.solve()
    c.hmatrix_build(a, [m] * p.nbody)
    x, info = c.ziter_solve(2, None, 2.0 * rhs)
    assert info == 0
    with pytest.raises(hm_bem.HMError) as e:
        c.forces()
    assert e.value.code == -200 and "before hm_solve" in str(e.value)
    again = c.solve_frequency(w, hm_bem.HM_FP64)
    for k in KEYS:
        assert np.array_equal(again[k], ref[k]), k
    c.hmatrix_release()


def test_ziter_argument_errors(quad):
    p, c = quad
    a, rhs = _assembled(c, 0.9)
    with pytest.raises(hm_bem.HMError, match="isolv"):
        c.ziter_solve(3, a, rhs)
    with pytest.raises(hm_bem.HMError, match="sum to"):
        c.ziter_solve(1, a, rhs, block_sizes=[3, 4])
    with pytest.raises(hm_bem.HMError, match="equal"):
        c.hmatrix_build(a, [p.n - 10, 10])
