"""Body frames and shared diagonal blocks on the GPU.

With body frames (hm_set_body_origins) the coordinates are local to each body
and a pair's geometry is the local difference plus the difference of the two
bodies' placements, which is exactly zero within one body. Bodies whose local
hull panels are bitwise equal (translated copies, as in an array of identical
devices) then have bitwise equal diagonal blocks, on the dense path too. The
H-matrix form assembles, factors and inverts such a block once per frequency
and copies it to the bodies that share it; the off-diagonal blocks of pairs of
such bodies at an equal offset share one cross approximation.

Bars:
- the dense FP64 path's diagonal blocks of equal bodies: equal bit for bit;
- the H-matrix solution with and without sharing at equal offsets: equal bit
  for bit (the shared blocks are the blocks each body would get, and the rest
  of the computation is the same launches on the same data);
- body symmetries: the maps found are exactly those of the body's shape (a
  cylinder with 16 panels a ring: all eight; a 2:1 box: identity, 180 degrees
  and the two axis reflections; the cylinder moved off its frame's origin
  along x: identity and the x-axis reflection); sharing across offsets that
  they relate (a permuted copy of the other block's cross approximation, the
  same approximation error) keeps A/B/F within the H-matrix form's 1e-4 of the
  dense FP64 solve, and its block products are held to the dense matrices in
  tests/test_gpu_hmatrix.py (20 eps, the default sharing);
- frames against the global coordinates (the oracle's): infinite depth within
  1e-12 (rounding of the pair differences only); finite depth within 20x the
  oracle's own band for +-1 ulp moves of the nodes (the reference's epsilon
  algorithm turns rounding-level input changes into jumps of that size), at
  least 1e-10 and at most 1e-4, as in the smoke check.
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


def _ctx(p, frames=True):
    c = hm_bem.Context()
    c.setup(p, frames=frames)
    return c


def test_body_classes():
    p = _array(2, 2, 60.0)
    c = _ctx(p)
    try:
        assert c.body_classes().tolist() == [0, 0, 0, 0]
    finally:
        c.close()
    c = _ctx(p, frames=False)
    try:
        assert c.body_classes().tolist() == [0, 1, 2, 3]
    finally:
        c.close()
    # a different third body: its own class; the fourth shares body 0's
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(0.0, 0.0)), meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0, 0.0)),
          meshes.cylinder(3.0, 4.5, 16, 4, 2, center=(0.0, 12.0)), meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0, 12.0))]
    q = meshes.Problem(bs, h=60.0, omegas=[0.5], beta=np.array([0.0]))
    c = _ctx(q)
    try:
        assert c.body_classes().tolist() == [0, 0, 2, 0]
    finally:
        c.close()


@pytest.mark.parametrize("h", [60.0, -1.0], ids=["finite", "deep"])
def test_dense_diagonal_blocks_equal_bodies(h):
    p = _array(2, 2, h)
    m = p.n // p.nbody
    c = _ctx(p)
    try:
        for w in p.omegas:
            c.set_frequency(w)
            c.assemble()
            a, s, _ = c.get_system()
            for b in range(1, p.nbody):
                sl = slice(b * m, (b + 1) * m)
                assert np.array_equal(a[sl, sl], a[:m, :m]), (h, w, b)
                assert np.array_equal(s[sl, sl], s[:m, :m]), (h, w, b)
    finally:
        c.close()


@pytest.mark.parametrize("h", [60.0, -1.0], ids=["finite", "deep"])
def test_hmatrix_sharing_is_bitwise(h):
    p = _array(2, 2, h)
    on, off = _ctx(p), _ctx(p)
    try:
        for c, share in ((on, True), (off, False)):
            c.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
            c.set_gmres(restart=50, tol=1e-6)
            c.set_block_sharing(share, symmetries=False)
        # 2 x 2 equal bodies: one diagonal block; 12 ordered pairs at 8 distinct
        # offsets, per matrix (A and S)
        assert on.block_sharing() == (1, 16)
        assert off.block_sharing() == (4, 24)
        for w in p.omegas:
            a = on.solve_frequency(w, hm_bem.HM_GMRES)
            b = off.solve_frequency(w, hm_bem.HM_GMRES)
            print(f"h={h} w={w}: gmres {on.gmres_stats()} / {off.gmres_stats()}", flush=True)
            for k in KEYS:
                assert np.array_equal(a[k], b[k]), (h, w, k, relerr(a[k], b[k]))
    finally:
        on.close()
        off.close()


def test_frames_against_global_coordinates(oracle):
    # infinite depth: the pair differences round differently, nothing else
    p = _array(2, 1, -1.0, omegas=(0.9,))
    fr, gl = _ctx(p), _ctx(p, frames=False)
    try:
        w = float(p.omegas[0])
        a = fr.solve_frequency(w, hm_bem.HM_FP64)
        b = gl.solve_frequency(w, hm_bem.HM_FP64)
        errs = {k: relerr(a[k], b[k]) for k in KEYS}
        print(f"deep: frames vs global {errs}", flush=True)
        assert max(errs.values()) <= 1e-12, errs
    finally:
        fr.close()
        gl.close()
    # finite depth: against the oracle (global coordinates) within its band
    p = _array(2, 1, 31.7, n_theta=12, n_z=3, n_r=2, omegas=(0.9,))
    c = _ctx(p)
    try:
        w = float(p.omegas[0])
        ref = oracle.solve_frequency(p, w)
        band = 0.0
        for d in (np.inf, -np.inf):
            bodies = [meshes.Body(xyz=np.nextafter(b.xyz, d), ncn=b.ncn, ncon=b.ncon, xr=b.xr) for b in p.bodies]
            pp = meshes.Problem(bodies, h=p.h, omegas=p.omegas, beta=p.beta)
            rp = oracle.solve_frequency(pp, w)
            band = max(band, max(relerr(rp[k], ref[k]) for k in KEYS))
        tol = min(max(1e-10, 20.0 * band), 1e-4)
        got = c.solve_frequency(w, hm_bem.HM_FP64)
        errs = {k: relerr(got[k], ref[k]) for k in KEYS}
        print(f"finite: frames vs oracle {errs}, band {band:.2e}, bar {tol:.2e}", flush=True)
        assert max(errs.values()) <= tol, (errs, tol)
    finally:
        c.close()


def _equal_bodies(body, centers):
    """copies of one body (its nodes taken as its local frame) placed at centers"""
    local = body.local if body.local is not None else body.xyz
    return [meshes.Body(xyz=meshes._placed(local, c), ncn=body.ncn, ncon=body.ncon, xr=(c[0], c[1], 0.0),
                        local=local.copy(), origin=(float(c[0]), float(c[1]))) for c in centers]


def test_sharing_and_fp16_arguments():
    p = _array(2, 1, 40.0)
    c = _ctx(p)
    try:
        with pytest.raises(hm_bem.HMError, match="0, 1 or 2"):
            c._chk(c.L.hm_set_block_sharing(c.ptr, hm_bem._i(3)), "hm_set_block_sharing")
        with pytest.raises(hm_bem.HMError, match="sizes"):
            c.cgemm_f16(np.zeros((0, 2), np.complex64), np.zeros((0, 2), np.complex64))
    finally:
        c.close()


def test_body_symmetries():
    centers = [(0.0, 0.0), (15.0, 0.0)]
    cyl = meshes.cylinder(3.0, 4.0, 16, 4, 2)
    moved = meshes.Body(xyz=cyl.xyz, ncn=cyl.ncn, ncon=cyl.ncon, local=cyl.local + np.array([0.5, 0.0, 0.0]))
    cases = [(cyl, [1, 1, 1, 1, 1, 1, 1, 1]),
             (meshes.box(8.0, 4.0, 3.0, 8, 4, 3), [1, 0, 1, 0, 1, 1, 0, 0]),
             (moved, [1, 0, 0, 0, 1, 0, 0, 0])]
    for body, want in cases:
        p = meshes.Problem(_equal_bodies(body, centers), h=40.0, omegas=[0.5], beta=np.array([0.0]))
        c = _ctx(p)
        try:
            assert c.body_symmetries().astype(int).tolist() == want
        finally:
            c.close()
    # unequal bodies or no frames: none
    p = _array(2, 1, 40.0)
    c = _ctx(p, frames=False)
    try:
        assert not c.body_symmetries().any()
    finally:
        c.close()


@pytest.mark.parametrize("h", [60.0, -1.0], ids=["finite", "deep"])
def test_symmetric_sharing_against_dense(h):
    p = _array(2, 2, h)
    sym, dense = _ctx(p), _ctx(p)
    try:
        sym.set_hmatrix(True, eps=1e-6, eps_s=1e-7, kmax=50)
        sym.set_gmres(restart=50, tol=1e-6)
        # 2 x 2 equal cylinders: the axis offsets are one class and the diagonal
        # ones another, per matrix (A and S)
        assert sym.block_sharing() == (1, 4)
        for w in p.omegas:
            got = sym.solve_frequency(w, hm_bem.HM_GMRES)
            ref = dense.solve_frequency(w, hm_bem.HM_FP64)
            errs = {k: relerr(got[k], ref[k]) for k in KEYS}
            print(f"h={h} w={w}: symmetric sharing against dense FP64 {errs}", flush=True)
            assert max(errs.values()) < 1e-4, errs
    finally:
        sym.close()
        dense.close()
