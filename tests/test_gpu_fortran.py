"""The Fortran drop-in layer executed: tests/fortran/hm_fortran_check, a Fortran
program that calls FINGREEN3D / INFGREEN3D, SGLINTBD_QUAD / SGLINTBD_TRI, the
driver's frequency chain (CALGREEN_MULTI, ASSB_LEFT_MULTI, ASSB_RBC_MULTI per
body, RADIATION_SOLVER_MULTI, then per heading ASSB_DBC_MULTI and
DIFFRACTION_SOLVER_MULTI; for one body with irregular-frequency removal
CALGREEN_IRR, ASSB_LEFT_IRR, ASSB_RBC_IRR, RADIATION_SOLVER_IRR, ASSB_DBC_IRR,
DIFFRACTION_SOLVER_IRR), the H-matrix route of ISOLV = 2
(SETUP_BLOCK_PRECONDITIONER, HMATRIX_BUILD, HMATVEC, ZITER_SOLVE_MULTI,
HMATRIX_DESTROY) on a smooth test kernel, and LAPACK's ZGETRF / ZGETRS with the reference's
argument lists (hams-mrel_amd/fortran/hm_bem.f90), against the same quantities
from the C-ABI through ctypes: the element kernels bit for bit, the potentials
within 1e-12 (the Fortran chain factors the host copy of AMAT or CMAT through
ZGETRF, the ctypes chain the device-resident system), the 4 x 4 solve within
1e-13 of numpy, the H-matrix product within 20 x its 1e-8 tolerance of the
dense product and its GMRES solution (tol 1e-10) within 1e-7 of numpy's."""
import os
import subprocess

import numpy as np
import pytest

import hm_bem
import meshes
from conftest import ROOT, relerr

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "fortran", "hm_fortran_check")


def _write(path, p, w1, pts, flags, sgl, lua):
    i32 = lambda a: np.asarray(a, dtype=np.int32)  # noqa: E731
    f64 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))  # noqa: E731
    iel, jel, irr = (np.array([s[k] for s in sgl]) for k in range(3))
    ix = p.ixyz_p if p.ni else np.zeros((1, 3))
    with open(path, "wb") as fh:
        i32([p.n, p.ni, p.nbody, len(p.xyz), p.nbeta, len(pts), len(sgl)]).tofile(fh)
        for a in (p.body_start, p.ncn):
            i32(a).tofile(fh)
        i32(p.ncon).ravel(order="F").tofile(fh)
        for a in (iel, jel, irr):
            i32(a).tofile(fh)
        for a in (p.xyz, p.xyz_p, p.ds, p.pnsz, p.dxyz_p, ix):
            f64(a).ravel(order="F").tofile(fh)
        np.array([p.h], dtype=np.float64).tofile(fh)
        f64(p.beta).tofile(fh)
        np.array([p.rho, p.g, p.g_disp, p.amp, p.xw[0], p.xw[1], w1], dtype=np.float64).tofile(fh)
        f64(pts).ravel(order="F").tofile(fh)
        f64(flags).tofile(fh)
        np.asfortranarray(lua, dtype=np.complex128).ravel(order="F").tofile(fh)


def _read(path, p, npts, nsgl):
    ncol = p.m + p.nbeta
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a
    grn = take(np.complex128, 4 * npts).reshape(npts, 4)
    sij = take(np.float64, nsgl)
    dij = take(np.float64, 3 * nsgl).reshape(nsgl, 3)
    out = take(np.complex128, p.n * ncol).reshape(ncol, p.n).T
    x = take(np.complex128, 4)
    info = int(take(np.int32, 1)[0])
    hy = take(np.complex128, p.n)
    hz = take(np.complex128, 2 * p.n).reshape(2, p.n).T
    hinfo = int(take(np.int32, 1)[0])
    assert off == len(raw)
    return grn, sij, dij, out, x, info, (hy, hz, hinfo)


def _case(tmp_path, p, w1):
    assert os.path.exists(EXE), "tests/fortran/hm_fortran_check is not built (__graft_entry__.build())"
    rng = np.random.default_rng(17)
    k = 12
    src, fld = p.xyz_p[rng.integers(0, p.n, k)], p.xyz_p[rng.integers(0, p.n, k)]
    fld[0] = src[0] + np.array([1e-3, 0.0, 0.0])
    pts = np.column_stack([src[:, 0], fld[:, 0], src[:, 1], fld[:, 1], src[:, 2], fld[:, 2]])
    flags = rng.integers(0, 2, k)
    sgl = [(int(rng.integers(1, p.n + 1)), int(j), 1) for j in rng.integers(1, p.n + 1, 6)]
    sgl.append((3, 3, 1))                                  # the self term
    if p.ni:
        sgl += [(int(i), int(j), 3) for i, j in zip(rng.integers(1, p.ni + 1, 3), rng.integers(1, p.n + 1, 3))]
    lua = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)) + 4 * np.eye(4)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write(fin, p, w1, pts, flags, sgl, lua)
    print(f"fortran caller: n={p.n} ni={p.ni} nbody={p.nbody} h={p.h} w={w1}", flush=True)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    grn, sij, dij, out, x, info, (hy, hz, hinfo) = _read(fout, p, len(pts), len(sgl))
    c = hm_bem.Context()
    try:
        c.setup(p, frames=False)             # the Fortran layer passes global coordinates
        v = w1 * w1 / p.g
        if p.h > 0:
            wvn = hm_bem.dispersion(w1, p.h, p.g_disp)
            g_c = c.fingreen3d(src, fld, v, wvn, p.h, flags)
        else:
            g_c = c.infgreen3d(src, fld, v, flags)
        assert np.array_equal(grn, g_c)
        for t, (i, j, irr) in enumerate(sgl):
            xf = (p.xyz_p if irr == 1 else p.ixyz_p)[i - 1]
            verts = np.zeros((1, 4, 3))
            nv = int(p.ncn[j - 1])
            verts[0, :nv] = p.xyz[p.ncon[j - 1, :nv] - 1]
            s_c, d_c = c.sglintbd(np.array([nv]), verts, xf[None, :])
            assert sij[t] == s_c[0] and np.array_equal(dij[t], d_c[0]), (t, i, j, irr)
        c.solve_frequency(w1, hm_bem.HM_FP64)
        ref = c.get_potentials()
    finally:
        c.close()
    e = relerr(out, ref)
    print(f"  potentials: Fortran chain against the C-ABI chain {e:.2e}", flush=True)
    assert e <= 1e-12
    assert info == 0
    assert relerr(x, np.linalg.solve(lua, np.arange(1, 5))) <= 1e-13
    # the H-matrix route (HMATRIX_BUILD over the cluster tree at eps 1e-8,
    # HMATVEC, ZITER_SOLVE_MULTI(2), HMATRIX_DESTROY) on the smooth test kernel
    xp = np.asarray(p.xyz_p)
    r = np.sqrt(((xp[:, None, :] - xp[None, :, :]) ** 2).sum(axis=2))
    np.fill_diagonal(r, 1.0)
    k = np.exp(0.3j * r) / r
    np.fill_diagonal(k, 4.0 + 1.0j)
    e = 1.0 + 1j * np.arange(1, p.n + 1) / p.n
    ey, ez = relerr(hy, k @ e), relerr(hz, np.linalg.solve(k, np.column_stack([e, np.conj(e)])))
    print(f"  H-matrix route: HMATVEC {ey:.2e}, ZITER_SOLVE_MULTI(2) {ez:.2e}, info {hinfo}", flush=True)
    assert hinfo == 0
    assert ey <= 2e-7
    assert ez <= 1e-7


def test_fortran_multibody_chain(tmp_path):
    bs = [meshes.cylinder(3.0, 4.0, 12, 3, 2, center=(12.0 * i, 0.0)) for i in range(2)]
    p = meshes.reference_parity_constants(meshes.Problem(bs, h=31.7, omegas=[0.9],
                                                         beta=np.array([0.0, np.pi / 3])))
    _case(tmp_path, p, 0.9)


def test_fortran_single_body_irr_chain(tmp_path):
    b = meshes.cylinder(5.0, 5.0, 16, 3, 2, waterplane_nr=2)
    p = meshes.reference_parity_constants(meshes.Problem([b], h=-1.0, omegas=[0.8],
                                                         beta=np.array([0.0, 0.5, 1.0])))
    _case(tmp_path, p, 0.8)
