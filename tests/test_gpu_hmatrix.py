"""H-matrix form of A and S (dense diagonal block per body, off-diagonal blocks
by adaptive cross approximation evaluated straight from the Green function)
against the dense assembly of the same problem.

Bars:
- products (diagonal blocks assembled pair by pair, bit for bit the dense
  blocks): the ACA stopping test |u_k||v_k| < eps |U V^T|_F bounds each
  block's relative error by about eps (it is an estimate, not a bound, so the
  bar is 20 eps) and the diagonal blocks are the dense ones, so
  ||(H - A) x|| <= 20 eps ||A x|| column by column for random x, for A and S;
- forces at eps = 1e-10 and full rank: within 1e-7 of the dense FP64 LU
  (GMRES tol 1e-10, FP64 residuals);
- forces at eps = 1e-5 for A and 1e-6 for S (rank <= 50, GMRES tol 1e-6):
  within 1e-4 of the dense FP64 LU, the accuracy the accelerated path is held
  to (measured: <= 4e-5 over 0.3..2.5 rad/s on 2x2 and 4x4 arrays);
- forces at the reference's ACA_TOL = 1e-4 for both: within 1e-2.
- diagonal blocks with PIDCOCK and EIGENE pairs assembled together with
  their mirrors (the default): the series terms are formed so that a pair's
  terms with the depths swapped are the mirror's terms bit for bit, and the
  epsilon tables are the whole tables the dense kernels build, so the blocks
  are the dense blocks bit for bit either way (bar 1e-13 on block products,
  the rounding of the products themselves). The reference's epsilon
  algorithm (exact-repeat exit) turns rounding changes of a pair's inputs
  into jumps of up to ~4e-6 of the block's largest entry; near the array's
  trapped-mode frequencies such jumps move A/B/F of the 16-cylinder array by
  ~1e-3, which is why the mirrored blocks must not change a single bit.
- The damping is the small imaginary part of the response at low frequency,
  and a relative Frobenius tolerance on the complex blocks does not resolve
  it (measured up to 4.3e-3 on the deep-water 2x2 array at 0.5 rad/s, 1.3e-3
  on the 4x4 array); added mass and excitation stay within 1.7e-4.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _array(nx, ny, h, spacing=12.0, n_theta=16, n_z=4, n_r=2, omegas=(0.5, 1.6)):
    bs = [meshes.cylinder(3.0, 4.0, n_theta, n_z, n_r, center=(spacing * i, spacing * j))
          for j in range(ny) for i in range(nx)]
    return meshes.Problem(bs, h=h, omegas=list(omegas), beta=np.array([0.0, np.pi / 2]))


def _ctx(p):
    c = hm_bem.Context()
    c.setup(p)
    return c


@pytest.fixture(scope="module", params=[60.0, -1.0], ids=["finite", "deep"])
def pair(request):
    p = _array(2, 2, request.param)
    dense, hm = _ctx(p), _ctx(p)
    yield p, dense, hm
    dense.close()
    hm.close()


@pytest.mark.parametrize("eps", [1e-4, 1e-7])
def test_products_match_dense(pair, eps):
    p, dense, hm = pair
    m = p.n // p.nbody
    hm.set_hmatrix(True, eps=eps, kmax=m)
    hm.set_pair_symmetry(False)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((p.n, 3)) + 1j * rng.standard_normal((p.n, 3))
    for w in p.omegas:
        dense.set_frequency(w)
        dense.assemble()
        a, s, _ = dense.get_system()
        hm.set_frequency(w)
        hm.assemble()
        st = hm.hmatrix_stats()
        print(f"h={p.h} w={w} eps={eps} stats={st}", flush=True)
        assert st["unconverged"] == 0, st
        assert 0 < st["mean_rank"] < m, st
        assert st["stored_fraction"] < 1.0, st
        for mat, ref in ((0, a), (1, s)):
            ya = ref @ x
            yh = hm.hmatrix_apply(mat, x)
            err = np.linalg.norm(yh - ya, axis=0) / np.linalg.norm(ya, axis=0)
            print(f"  mat {mat}: rel err per column {err}", flush=True)
            assert np.all(err < 20 * eps), (mat, err)


def test_forces_full_rank_tight(pair):
    p, dense, hm = pair
    m = p.n // p.nbody
    hm.set_hmatrix(True, eps=1e-10, kmax=m)
    hm.set_pair_symmetry(False)
    hm.set_gmres(restart=50, tol=1e-10, maxiter=500, precond=1)
    for w in p.omegas:
        ref = dense.solve_frequency(w, hm_bem.HM_FP64)
        got = hm.solve_frequency(w, hm_bem.HM_GMRES)
        gs = hm.gmres_stats()
        assert gs["converged"], gs
        for k in KEYS:
            assert relerr(got[k], ref[k]) < 1e-7, (w, k, relerr(got[k], ref[k]))


@pytest.mark.parametrize("eps,eps_s,bar", [(1e-5, 1e-6, 1e-4), (1e-4, 1e-4, 1e-2)])
def test_forces_at_tolerance(pair, eps, eps_s, bar):
    p, dense, hm = pair
    hm.set_hmatrix(True, eps=eps, eps_s=eps_s, kmax=50)
    hm.set_pair_symmetry(True)
    hm.set_gmres(restart=50, tol=1e-6, maxiter=500, precond=-1)
    for w in p.omegas:
        ref = dense.solve_frequency(w, hm_bem.HM_FP64)
        got = hm.solve_frequency(w, hm_bem.HM_GMRES)
        gs, st = hm.gmres_stats(), hm.hmatrix_stats()
        errs = {k: relerr(got[k], ref[k]) for k in KEYS}
        print(f"h={p.h} w={w} gmres={gs} hmatrix={st} errs={errs}", flush=True)
        assert gs["converged"], gs
        assert st["unconverged"] == 0, st
        for k in KEYS:
            assert errs[k] < bar, (w, k, errs)


def test_toggle_back_to_dense(pair):
    p, dense, hm = pair
    w = p.omegas[0]
    ref = dense.solve_frequency(w, hm_bem.HM_FP64)
    hm.set_hmatrix(True)
    hm.solve_frequency(w, hm_bem.HM_GMRES)
    hm.set_hmatrix(False)
    got = hm.solve_frequency(w, hm_bem.HM_FP64)
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


def test_hmatrix_argument_errors():
    # geometries the body clusters do not fit take the cluster tree
    # (tests/test_gpu_htree.py checks their solutions)
    single = _ctx(_array(1, 1, 60.0))
    try:
        single.set_hmatrix(True)
    finally:
        single.close()
    p = _array(2, 1, 60.0)
    c = _ctx(p)
    try:
        with pytest.raises(hm_bem.HMError, match="eps"):
            c.set_hmatrix(True, eps=0.0)
        with pytest.raises(hm_bem.HMError, match="kmax"):
            c.set_hmatrix(True, kmax=0)
        with pytest.raises(hm_bem.HMError, match="leaf"):
            c.set_cluster_tree(0, leaf=65)
        with pytest.raises(hm_bem.HMError, match="no cluster-tree"):
            c.cluster_tree_stats()
        c.set_hmatrix(True)
        c.set_frequency(0.8)
        c.assemble()
        with pytest.raises(hm_bem.HMError, match="no dense system"):
            c.get_system()
        with pytest.raises(hm_bem.HMError, match="HM_GMRES only"):
            c.factor(hm_bem.HM_FP64)
        with pytest.raises(hm_bem.HMError, match="mat must be"):
            c.hmatrix_apply(2, np.zeros(p.n))
    finally:
        c.close()


def test_linton_entries_between_stacked_bodies():
    """Two copies of one cylinder, the second 10 m below the first: their side
    panels share horizontal positions (R = 0, the LINTON regime) across the
    off-diagonal blocks. The cross approximation evaluates those entries one
    lane at a time after its lane-group pass; products against the dense
    assembly at eps = 1e-7, full rank allowed."""
    import dataclasses
    top = meshes.cylinder(3.0, 4.0, 16, 4, 2)
    low = dataclasses.replace(top, xyz=top.xyz + np.array([0.0, 0.0, -10.0]), local=top.local + np.array([0.0, 0.0, -10.0]),
                              xr=(0.0, 0.0, -10.0))
    p = meshes.Problem([top, low], h=60.0, omegas=[0.5, 1.6], beta=np.array([0.0]))
    m = p.n // 2
    xy = p.xyz_p[:, :2]
    rh = np.linalg.norm(xy[:m, None, :] - xy[None, m:, :], axis=-1) / p.h
    assert np.count_nonzero(rh < 5e-4) > 0          # LINTON pairs across the two bodies
    dense, hm = _ctx(p), _ctx(p)
    try:
        eps = 1e-7
        hm.set_hmatrix(True, eps=eps, kmax=m)
        hm.set_pair_symmetry(False)                 # diagonal blocks bit for bit the dense ones
        rng = np.random.default_rng(11)
        x = rng.standard_normal((p.n, 2)) + 1j * rng.standard_normal((p.n, 2))
        for w in p.omegas:
            dense.set_frequency(w)
            dense.assemble()
            a, s, _ = dense.get_system()
            hm.set_frequency(w)
            hm.assemble()
            st = hm.hmatrix_stats()
            assert st["unconverged"] == 0, st
            for mat, ref in ((0, a), (1, s)):
                ya = ref @ x
                err = np.linalg.norm(hm.hmatrix_apply(mat, x) - ya, axis=0) / np.linalg.norm(ya, axis=0)
                print(f"stacked w={w} mat {mat}: rel err {err} stats {st}", flush=True)
                assert np.all(err < 20 * eps), (w, mat, err)
    finally:
        dense.close()
        hm.close()


def test_diagonal_blocks_match_dense(pair):
    """The diagonal blocks of the H-matrix form against the dense assembly: x
    supported on one body picks that body's diagonal block out of H x. With the
    PIDCOCK and EIGENE pairs assembled together with their mirrors, and pair by
    pair, the blocks are the dense blocks bit for bit (a wrong mirror, e.g. the
    pair's own depth derivative, is off at O(1); a mirror equal only to
    rounding moves entries by up to ~4e-6 through the epsilon algorithm)."""
    p, dense, hm = pair
    if p.h < 0:
        pytest.skip("infinite depth: no series regimes, no mirrored pairs")
    m = p.n // p.nbody
    hm.set_hmatrix(True, eps=1e-5, kmax=m)
    rng = np.random.default_rng(13)
    for sym, bar in ((True, 1e-13), (False, 1e-13)):
        hm.set_pair_symmetry(sym)
        hist = hm.pair_histogram()
        assert hist[1].sum() > 0 and hist[2].sum() > 0   # PIDCOCK and EIGENE pairs in the diagonal blocks
        worst = 0.0
        for w in p.omegas:
            dense.set_frequency(w)
            dense.assemble()
            a, s, _ = dense.get_system()
            hm.set_frequency(w)
            hm.assemble()
            for b in range(p.nbody):
                x = np.zeros((p.n, 2), complex)
                x[b * m:(b + 1) * m] = rng.standard_normal((m, 2)) + 1j * rng.standard_normal((m, 2))
                for mat, ref in ((0, a), (1, s)):
                    ya = (ref @ x)[b * m:(b + 1) * m]
                    yh = hm.hmatrix_apply(mat, x)[b * m:(b + 1) * m]
                    err = np.linalg.norm(yh - ya) / np.linalg.norm(ya)
                    worst = max(worst, err)
                    assert err < bar, (sym, w, b, mat, err)
        print(f"h={p.h} mirrored={sym}: largest block product error {worst:.2e}", flush=True)
    hm.set_pair_symmetry(True)


def test_overlap_switch_gives_the_same_solution(pair):
    """the cross approximation on the second stream beside the diagonal blocks
    (the default) or after them on the context's stream: the same kernels on
    the same inputs, so the same A/B/F bit for bit; the stage timer counts one
    entry-kernel record per launch either way"""
    p, dense, hm = pair
    hm.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
    hm.set_pair_symmetry(True)
    hm.set_gmres(restart=50, tol=1e-8, maxiter=500, precond=-1)
    w = p.omegas[1]
    got = {}
    for ov in (True, False):
        hm.set_hmatrix_overlap(ov)
        hm.set_timing(True)
        got[ov] = hm.solve_frequency(w, hm_bem.HM_GMRES)
        st = hm.timing()
        hm.set_timing(False)
        steps = hm.hmatrix_stats()["steps"]
        assert st["aca_entries"][1] == 2 * steps, (ov, st["aca_entries"], steps)
        assert st["asm_cross"][1] == 1
    hm.set_hmatrix_overlap(True)
    for k in KEYS:
        assert np.array_equal(got[True][k], got[False][k]), k
