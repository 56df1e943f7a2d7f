"""H-matrix over the reference's geometric cluster tree (HMATRIX_BUILD,
HMatrix.f90:135-497: leaf 32, eta = 2) for the geometries the body clusters do
not fit: one body with interior waterplane panels (the C1 cylinder, 1024 +
256 panels, whose system is the normal equations C = A^T A), two bodies of
unequal panel counts, and equal bodies on request.

Bars:
- products of the tree H-matrix built from an assembled matrix
  (hm_hmatrix_build_tree): every admissible block stops at |u_k||v_k| <
  eps |U V^T|_F, an estimate of its relative error, and the dense blocks are
  copies, so ||(H - M) X|| <= 20 eps ||M X|| (the bar of the body-cluster
  products); the oracle's restatement of the reference's own route
  (oracle/htree_ref.py: its pivot rules and norm estimate) is held to the same
  bar on the same matrix, and the partition's block counts equal the host
  partition's (hm_cluster_tree_partition);
- GMRES (tol 1e-10) on the tree H-matrix against numpy's dense solve: within
  the compression error times the conditioning, measured below 1e-5 at
  eps = 1e-7; bar 1e-5;
- the assembly-aware form (hm_set_hmatrix choosing the tree): A, B, F within
  1e-4 of the dense FP64 LU at eps = 1e-6 (GMRES tol 1e-8), the bar of the
  accelerated path.
"""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

pytestmark = pytest.mark.gpu
KEYS = ("amas", "bdmp", "exfc")


def _uneven(h=40.0):
    bs = [meshes.cylinder(3.0, 4.0, 32, 8, 4), meshes.cylinder(2.0, 3.0, 24, 6, 3, center=(14.0, 2.0))]
    return meshes.Problem(bs, h=h, omegas=[0.5, 1.4], beta=np.array([0.0, np.pi / 3]))


def _equal():
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(16.0 * i, 0.0)) for i in range(2)]
    return meshes.Problem(bs, h=60.0, omegas=[0.9], beta=np.array([0.0]))


def _ctx(p):
    c = hm_bem.Context()
    c.setup(p)
    return c


def _system(c, p, w):
    """the n x n system matrix and right-hand sides the solve works on"""
    c.set_frequency(w)
    c.assemble()
    if p.ni:
        return c.get_normal_equations()
    a, _, rhs = c.get_system()
    return np.asarray(a), np.asarray(rhs)


@pytest.mark.parametrize("case", ["c1 single body irr", "uneven bodies"])
def test_build_tree_from_an_assembled_matrix(case):
    import htree_ref
    p = meshes.c1_single_cylinder() if case.startswith("c1") else _uneven()
    w = float(p.omegas[len(p.omegas) // 2])
    c = _ctx(p)
    try:
        m, b = _system(c, p, w)
        blocks = [] if p.nbody == 1 else list(np.diff(p.body_start))
        rng = np.random.default_rng(11)
        x = rng.normal(size=(p.n, 4)) + 1j * rng.normal(size=(p.n, 4))
        ref = m @ x
        _, part = hm_bem.cluster_tree_partition(p.xyz_p)
        for eps in (1e-4, 1e-7):
            c.set_hmatrix(False, eps=eps, eps_s=eps)
            print(f"{case}: n={p.n} ni={p.ni} eps={eps}", flush=True)
            c.hmatrix_build_tree(m, p.xyz_p, blocks)
            st, hs = c.cluster_tree_stats(), c.hmatrix_stats()
            print(f"  tree {st}, ranks {hs['mean_rank']:.1f}/{hs['max_rank']}, stored {hs['stored_fraction']:.3f}",
                  flush=True)
            assert st["dense_blocks"] + st["lowrank_blocks"] == len(part)
            assert st["lowrank_blocks"] + st["dense_fallbacks"] == int(part[:, 4].sum())
            assert st["lowrank_blocks"] > 0
            e = relerr(c.hmatrix_apply(0, x), ref)
            perm, hb = htree_ref.build(m, p.xyz_p, tol=eps)
            e_ref = relerr(htree_ref.matvec(perm, hb, x), ref)
            print(f"  products: GPU {e:.2e}, the oracle's restatement of the reference {e_ref:.2e}", flush=True)
            assert e <= 20 * eps
            assert e_ref <= 20 * eps
        c.set_gmres(tol=1e-10)
        got, info = c.ziter_solve(2, None, b)
        assert info == 0
        e = relerr(got, np.linalg.solve(m, b))
        print(f"  GMRES on the tree H-matrix (eps 1e-7) against the dense solve: {e:.2e}, {c.gmres_stats()}",
              flush=True)
        assert e <= 1e-5
    finally:
        c.close()


@pytest.mark.parametrize("case", ["c1 single body irr", "uneven bodies", "equal bodies on request"])
def test_assembly_takes_the_tree(case):
    p = {"c1 single body irr": meshes.c1_single_cylinder, "uneven bodies": _uneven,
         "equal bodies on request": _equal}[case]()
    c = _ctx(p)
    try:
        ws = [float(p.omegas[0]), float(p.omegas[-1])]
        ref = {w: c.solve_frequency(w, hm_bem.HM_FP64) for w in ws}
        if case.startswith("equal"):
            c.set_cluster_tree(1)
        c.set_gmres(tol=1e-8)
        c.set_hmatrix(True, eps=1e-6, eps_s=1e-6)
        for w in ws:
            print(f"{case}: n={p.n} ni={p.ni} nbody={p.nbody} w={w}", flush=True)
            got = c.solve_frequency(w, hm_bem.HM_GMRES)
            st, gs = c.cluster_tree_stats(), c.gmres_stats()
            errs = {k: relerr(got[k], ref[w][k]) for k in KEYS}
            print(f"  tree {st}, gmres {gs}, errors {errs}", flush=True)
            assert gs["converged"]
            for k in KEYS:
                assert errs[k] <= 1e-4, (w, k, errs)
        # HMATRIX_DESTROY: the solve needs the tree again
        c.hmatrix_release()
        with pytest.raises(hm_bem.HMError, match="hm_solve before hm_factor"):
            c.solve()
        got = c.solve_frequency(ws[0], hm_bem.HM_GMRES)
        assert relerr(got["amas"], ref[ws[0]]["amas"]) <= 1e-4
    finally:
        c.close()
