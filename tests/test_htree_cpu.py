"""The cluster tree and block partition of the H-matrix (HMatrix.f90:203-497), on
the host: the library's partition (hm_cluster_tree_partition, no device work)
against the oracle's restatement (oracle/htree_ref.py), and the oracle pinned
by what the construction fixes: the blocks tile the matrix once, every
admissible block meets DIST > eta DIAM_MAX for the points it holds, every
stored-dense inadmissible block pairs two leaves, every split is at the
midpoint of the longest bounding-box side; and the oracle's cross
approximation and product against the dense product of a smooth kernel."""
import numpy as np
import pytest

import hm_bem
import meshes
from conftest import relerr

htree_ref = pytest.importorskip("htree_ref")


_SETS = {}


def _point_sets():
    if _SETS:
        return _SETS
    c1 = meshes.c1_single_cylinder()
    uneven = meshes.Problem([meshes.cylinder(3.0, 4.0, 16, 4, 2), meshes.cylinder(1.5, 2.0, 8, 2, 1, center=(11.0, 3.0))],
                            h=40.0, omegas=[1.0], beta=np.array([0.0]))
    rng = np.random.default_rng(20261204)
    blob = np.concatenate([rng.normal(size=(300, 3)) * [8.0, 2.0, 1.0], rng.normal(size=(200, 3)) + [30.0, 0.0, -2.0]])
    _SETS.update({"c1 cylinder": c1.xyz_p, "uneven bodies": uneven.xyz_p, "two clouds": blob,
                  "tiny": rng.normal(size=(7, 3))})
    return _SETS


@pytest.mark.parametrize("leaf,eta", [(32, 2.0), (16, 1.0), (64, 3.0)])
@pytest.mark.parametrize("name", ["c1 cylinder", "uneven bodies", "two clouds", "tiny"])
def test_partition_matches_the_oracle(name, leaf, eta):
    xyz = _point_sets()[name]
    perm, nodes = htree_ref.cluster_tree(xyz, leaf)
    ref = htree_ref.block_partition(nodes, eta)
    gp, gb = hm_bem.cluster_tree_partition(xyz, leaf, eta)
    assert np.array_equal(gp, perm)
    assert sorted(map(tuple, gb.tolist())) == sorted(ref)


@pytest.mark.parametrize("name", ["c1 cylinder", "uneven bodies", "two clouds", "tiny"])
def test_partition_invariants(name):
    xyz = np.asarray(_point_sets()[name])
    n, leaf, eta = len(xyz), 32, 2.0
    perm, nodes = htree_ref.cluster_tree(xyz, leaf)
    assert sorted(perm) == list(range(n))
    pts = xyz[perm]
    # the tree: children split their parent at the midpoint of its longest side
    for nd in nodes:
        p = pts[nd["s"]:nd["e"]]
        ln = p.max(axis=0) - p.min(axis=0)
        assert nd["diam"] == pytest.approx(np.linalg.norm(ln), rel=1e-15)
        np.testing.assert_allclose(nd["c"], p.mean(axis=0), rtol=1e-12, atol=1e-12)
        if nd["left"] < 0:
            assert len(p) <= leaf
            continue
        assert len(p) > leaf
        l, r = nodes[nd["left"]], nodes[nd["right"]]
        assert (l["s"], l["e"], r["s"], r["e"]) == (nd["s"], l["e"], l["e"], nd["e"])
        d = int(np.argmax(ln))
        mid = 0.5 * (p[:, d].min() + p[:, d].max())
        left_pts = pts[l["s"]:l["e"], d]
        right_pts = pts[r["s"]:r["e"], d]
        # a clamped split (all points on one side) still leaves both children non-empty
        if (left_pts <= mid).all() and (right_pts > mid).all():
            continue
        assert len(left_pts) == 1 or len(right_pts) == 1
    leaves = {(nd["s"], nd["e"]) for nd in nodes if nd["left"] < 0}
    blocks = htree_ref.block_partition(nodes, eta)
    cover = np.zeros((n, n), dtype=np.int32)
    for rs, nr, cs, nc, lr in blocks:
        cover[rs:rs + nr, cs:cs + nc] += 1
        a, b = pts[rs:rs + nr], pts[cs:cs + nc]
        dist = np.linalg.norm(a.mean(axis=0) - b.mean(axis=0))
        dmax = max(np.linalg.norm(a.max(axis=0) - a.min(axis=0)), np.linalg.norm(b.max(axis=0) - b.min(axis=0)))
        admissible = dist > eta * dmax and nr > 1 and nc > 1
        if lr:
            assert admissible and nr > 4 and nc > 4
        elif not admissible:
            assert (rs, rs + nr) in leaves and (cs, cs + nc) in leaves
    assert cover.min() == 1 and cover.max() == 1


def _helmholtz(xyz, k=0.3):
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    np.fill_diagonal(d, 1.0)
    a = np.exp(1j * k * d) / d
    np.fill_diagonal(a, 4.0 + 1.0j)
    return a


def test_oracle_compression_of_a_smooth_kernel():
    """The oracle's HMATRIX_BUILD on e^{ikr}/r over the C1 cylinder's centroids:
    admissible blocks compress to low rank and the product stays within a few
    times the tolerance of the dense product"""
    xyz = meshes.c1_single_cylinder().xyz_p
    a = _helmholtz(xyz)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(len(xyz), 3)) + 1j * rng.normal(size=(len(xyz), 3))
    for tol in (1e-4, 1e-6):
        perm, blocks = htree_ref.build(a, xyz, tol=tol)
        lr = [f for *_, f in blocks if isinstance(f, tuple)]
        assert len(lr) > 0
        ranks = [f[0].shape[1] for f in lr]
        e = relerr(htree_ref.matvec(perm, blocks, x), a @ x)
        print(f"tol {tol}: {len(lr)} low-rank blocks, mean rank {np.mean(ranks):.1f}, product error {e:.2e}")
        assert e <= 10 * tol
        assert max(ranks) < 32
