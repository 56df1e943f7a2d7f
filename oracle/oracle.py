"""ctypes bindings of the CPU checkers (TEST INFRASTRUCTURE ONLY).

`Oracle` wraps liboracle.so, the project's C restatement of the HAMS-MREL hot
path (oracle/hm_oracle.c). `Reference` wraps oracle/_ref/libhamsref.so, the
reference's own Fortran kernels compiled from its sources with a bind(C) shim.
Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
this module; the product library never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "-f", os.path.join(HERE, "Makefile"), "all"], check=True)


class _Problem(ctypes.Structure):
    _fields_ = [("n", _i), ("ni", _i), ("nbody", _i), ("body_start", _ip),
                ("xyz", _dp), ("ncn", _ip), ("ncon", _ip),
                ("xyz_p", _dp), ("ds", _dp), ("pnsz", _dp), ("dxyz_p", _dp),
                ("ixyz_p", _dp), ("h", _d), ("rho", _d), ("g", _d), ("g_disp", _d),
                ("xw", _d * 2), ("amp", _d), ("nbeta", _i), ("beta", _dp)]


class _Result(ctypes.Structure):
    _fields_ = [("amas", _dp), ("bdmp", _dp), ("exfc", _dp), ("amat", _dp),
                ("smat", _dp), ("rhs", _dp), ("mxpot", _dp)]


class Oracle:
    def __init__(self, path=os.path.join(HERE, "liboracle.so")):
        self.lib = L = ctypes.CDLL(path)
        L.hmo_error.restype = _d
        L.hmo_e1xa.restype = _d
        L.hmo_enxa.restype = _d
        L.hmo_eix.restype = _d
        L.hmo_limes.restype = _d
        L.hmo_flops_get.restype = _d
        for f in ("hmo_error", "hmo_e1xa", "hmo_eix"):
            getattr(L, f).argtypes = [_d]
        L.hmo_enxa.argtypes = [_i, _d]

    # ---- scalar special functions -----------------------------------------
    def error(self, x):
        return self.lib.hmo_error(_d(x))

    def e1xa(self, x):
        return self.lib.hmo_e1xa(_d(x))

    def enxa(self, n, x):
        return self.lib.hmo_enxa(_i(n), _d(x))

    def eix(self, x):
        return self.lib.hmo_eix(_d(x))

    def ik01a(self, x):
        a, b = _d(), _d()
        self.lib.hmo_ik01a(_d(x), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def jy01b(self, x):
        o = [_d() for _ in range(4)]
        self.lib.hmo_jy01b(_d(x), *[ctypes.byref(v) for v in o])
        return tuple(v.value for v in o)

    def limes(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64).copy()
        return self.lib.hmo_limes(_i(len(s) - 1), _ptr(s))

    def dispersion(self, w, h, nrt=200, g_disp=9.807):
        wvn = np.zeros(nrt)
        self.lib.hmo_dispersion(_ptr(wvn), _i(nrt), _d(w), _d(h), _d(g_disp))
        return wvn

    # ---- Green functions -----------------------------------------------------
    def fingreen3d(self, src, fld, v, wvn, h, flag):
        g = np.zeros(4, dtype=np.complex128)
        wvn = np.ascontiguousarray(wvn, dtype=np.float64)
        self.lib.hmo_fingreen3d(_d(src[0]), _d(fld[0]), _d(src[1]), _d(fld[1]), _d(src[2]),
                                _d(fld[2]), _d(v), _ptr(wvn), _i(len(wvn)), _d(h), _ptr(g), _i(flag))
        return g

    def infgreen3d(self, src, fld, v, flag):
        g = np.zeros(4, dtype=np.complex128)
        self.lib.hmo_infgreen3d(_d(src[0]), _d(fld[0]), _d(src[1]), _d(fld[1]), _d(src[2]),
                                _d(fld[2]), _d(v), _ptr(g), _i(flag))
        return g

    def sglintbd(self, verts, xf):
        """verts: (ncn, 3) panel vertices; xf: field point."""
        verts = np.asarray(verts, dtype=np.float64)
        v = np.zeros(12)
        v[:verts.size] = verts.ravel()
        xf = np.ascontiguousarray(xf, dtype=np.float64)
        s = _d()
        d = np.zeros(3)
        self.lib.hmo_sglintbd(_i(verts.shape[0]), _ptr(v), _ptr(xf), ctypes.byref(s), _ptr(d))
        return s.value, d

    def flops_reset(self):
        self.lib.hmo_flops_reset()

    def flops(self):
        return self.lib.hmo_flops_get()

    # ---- geometry and one panel-frequency system -----------------------------
    def panel_geometry(self, xyz, ncn, ncon, rg=(0.0, 0.0, 0.0)):
        ne = len(ncn)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        ncn = np.ascontiguousarray(ncn, dtype=np.int32)
        ncon = np.ascontiguousarray(ncon, dtype=np.int32)
        rg = np.ascontiguousarray(rg, dtype=np.float64)
        xp = np.zeros((ne, 3))
        ds = np.zeros(ne)
        pn = np.zeros(ne)
        dx = np.zeros((ne, 6))
        self.lib.hmo_panel_geometry(_i(ne), _ptr(ncn, _ip), _ptr(ncon, _ip), _ptr(xyz), _ptr(rg),
                                    _ptr(xp), _ptr(ds), _ptr(pn), _ptr(dx))
        return xp, ds, pn, dx

    def solve_frequency(self, prob, w1, want_matrices=False):
        """prob: mesh.Problem-like object (see tests/meshes.py). Returns dict."""
        return self._solve(prob, lambda P, R: self.lib.hmo_solve_frequency(P, _d(w1), R), want_matrices)

    def solve_limit(self, prob, which, want_matrices=False):
        """zero (which=0) or infinite (which=1) wave frequency (SYBO=1)"""
        return self._solve(prob, lambda P, R: self.lib.hmo_solve_limit(P, _i(which), R), want_matrices)

    def field_points(self, prob, w1, xfp):
        """potential, pressure and elevation of every radiation mode and heading
        at the points xfp (nfp x 3) after the frequency solve (CalPotentialMulti,
        CalPressureMulti, CalElevationMulti)"""
        P, keep = self._problem(prob)
        xfp = np.ascontiguousarray(xfp, dtype=np.float64).reshape(-1, 3)
        nfp, ncol = len(xfp), 6 * prob.nbody + len(prob.beta)
        out = {k: np.zeros((nfp, ncol), dtype=np.complex128, order="F") for k in ("pot", "prs", "elv")}
        info = self.lib.hmo_field_points(ctypes.byref(P), _d(w1), _i(nfp), _ptr(xfp), _ptr(out["pot"]),
                                         _ptr(out["prs"]), _ptr(out["elv"]))
        out["info"] = info
        return out

    def solve_motion(self, w1, amas, bdmp, exfc, matx, crs, kstf=None, blnr=None, bqdr=None):
        """motion response of one heading (SolveMotionMulti); body matrices nbody x 6 x 6"""
        nb = np.asarray(matx).shape[0]
        f = lambda a: np.asfortranarray(np.zeros((nb, 6, 6)) if a is None else a, dtype=np.float64)  # noqa: E731
        args = [np.asfortranarray(amas, dtype=np.float64), np.asfortranarray(bdmp, dtype=np.float64),
                f(blnr), f(bqdr), f(matx), f(crs), f(kstf)]
        ex = np.ascontiguousarray(exfc, dtype=np.complex128)
        x = np.zeros(6 * nb, dtype=np.complex128)
        it = ctypes.c_int(0)
        info = self.lib.hmo_solve_motion(_i(nb), _d(w1), *[_ptr(a) for a in args], _ptr(ex), _ptr(x),
                                         ctypes.byref(it))
        return x, it.value, info

    def _problem(self, prob):
        n, ni, nb = prob.n, prob.ni, prob.nbody
        nbeta = len(prob.beta)
        keep = [np.ascontiguousarray(prob.body_start, dtype=np.int32),
                np.ascontiguousarray(prob.xyz, dtype=np.float64),
                np.ascontiguousarray(prob.ncn, dtype=np.int32),
                np.ascontiguousarray(prob.ncon, dtype=np.int32),
                np.ascontiguousarray(prob.xyz_p, dtype=np.float64),
                np.ascontiguousarray(prob.ds, dtype=np.float64),
                np.ascontiguousarray(prob.pnsz, dtype=np.float64),
                np.ascontiguousarray(prob.dxyz_p, dtype=np.float64),
                np.ascontiguousarray(prob.ixyz_p if ni else np.zeros((1, 3)), dtype=np.float64),
                np.ascontiguousarray(prob.beta, dtype=np.float64)]
        P = _Problem(n, ni, nb, _ptr(keep[0], _ip), _ptr(keep[1]), _ptr(keep[2], _ip),
                     _ptr(keep[3], _ip), _ptr(keep[4]), _ptr(keep[5]), _ptr(keep[6]),
                     _ptr(keep[7]), _ptr(keep[8]), _d(prob.h), _d(prob.rho), _d(prob.g),
                     _d(prob.g_disp), (_d * 2)(*prob.xw), _d(prob.amp), nbeta, _ptr(keep[9]))
        return P, keep

    def _solve(self, prob, call, want_matrices):
        n, ni, nb = prob.n, prob.ni, prob.nbody
        m, nbeta = 6 * nb, len(prob.beta)
        keep = [np.ascontiguousarray(prob.body_start, dtype=np.int32),
                np.ascontiguousarray(prob.xyz, dtype=np.float64),
                np.ascontiguousarray(prob.ncn, dtype=np.int32),
                np.ascontiguousarray(prob.ncon, dtype=np.int32),
                np.ascontiguousarray(prob.xyz_p, dtype=np.float64),
                np.ascontiguousarray(prob.ds, dtype=np.float64),
                np.ascontiguousarray(prob.pnsz, dtype=np.float64),
                np.ascontiguousarray(prob.dxyz_p, dtype=np.float64),
                np.ascontiguousarray(prob.ixyz_p if ni else np.zeros((1, 3)), dtype=np.float64),
                np.ascontiguousarray(prob.beta, dtype=np.float64)]
        P = _Problem(n, ni, nb, _ptr(keep[0], _ip), _ptr(keep[1]), _ptr(keep[2], _ip),
                     _ptr(keep[3], _ip), _ptr(keep[4]), _ptr(keep[5]), _ptr(keep[6]),
                     _ptr(keep[7]), _ptr(keep[8]), _d(prob.h), _d(prob.rho), _d(prob.g),
                     _d(prob.g_disp), (_d * 2)(*prob.xw), _d(prob.amp), nbeta, _ptr(keep[9]))
        amas = np.zeros((m, m), order="F")
        bdmp = np.zeros((m, m), order="F")
        exfc = np.zeros((m, nbeta), dtype=np.complex128, order="F")
        R = _Result(_ptr(amas), _ptr(bdmp), _ptr(exfc), None, None, None, None)
        out = {}
        if want_matrices:
            nt = n + ni
            out["amat"] = np.zeros((nt, n), dtype=np.complex128, order="F")
            out["smat"] = np.zeros((nt, n), dtype=np.complex128, order="F")
            out["rhs"] = np.zeros((nt, m + nbeta), dtype=np.complex128, order="F")
            out["mxpot"] = np.zeros((n, m + nbeta), dtype=np.complex128, order="F")
            R.amat, R.smat = _ptr(out["amat"]), _ptr(out["smat"])
            R.rhs, R.mxpot = _ptr(out["rhs"]), _ptr(out["mxpot"])
        info = call(ctypes.byref(P), ctypes.byref(R))
        out.update(amas=amas, bdmp=bdmp, exfc=exfc, info=info)
        return out

    def zgetrf_solve(self, a, b):
        a = np.array(a, dtype=np.complex128, order="F", copy=True)
        b = np.array(b, dtype=np.complex128, order="F", copy=True)
        n = a.shape[0]
        ipiv = np.zeros(n, dtype=np.int32)
        info = self.lib.hmo_zgetrf(_i(n), _ptr(a), _i(n), _ptr(ipiv, _ip))
        self.lib.hmo_zgetrs(_i(n), _i(b.shape[1]), _ptr(a), _i(n), _ptr(ipiv, _ip), _ptr(b), _i(n))
        return b, info


class Reference:
    """The reference's Fortran kernels (oracle/_ref/libhamsref.so)."""

    PATH = os.path.join(HERE, "_ref", "libhamsref.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self):
        self.lib = ctypes.CDLL(self.PATH)

    def fingreen3d(self, src, fld, v, wvn, h, flag):
        g = np.zeros(4, dtype=np.complex128)
        wvn = np.ascontiguousarray(wvn, dtype=np.float64)
        self.lib.ref_fingreen3d(_d(src[0]), _d(fld[0]), _d(src[1]), _d(fld[1]), _d(src[2]),
                                _d(fld[2]), _d(v), _ptr(wvn), _i(len(wvn)), _d(h), _ptr(g), _i(flag))
        return g

    def infgreen3d(self, src, fld, v, flag):
        g = np.zeros(4, dtype=np.complex128)
        self.lib.ref_infgreen3d(_d(src[0]), _d(fld[0]), _d(src[1]), _d(fld[1]), _d(src[2]),
                                _d(fld[2]), _d(v), _ptr(g), _i(flag))
        return g

    def dispersion(self, w, h, nrt=200):
        wvn = np.zeros(nrt)
        self.lib.ref_dispersion(_ptr(wvn), _i(nrt), _d(w), _d(h))
        return wvn

    def sglintbd(self, verts, xf):
        verts = np.asarray(verts, dtype=np.float64)
        v = np.zeros((4, 3))
        v[:verts.shape[0]] = verts
        v = np.ascontiguousarray(v)
        xf = np.ascontiguousarray(xf, dtype=np.float64)
        s = _d()
        d = np.zeros(3)
        self.lib.ref_sglintbd(_i(verts.shape[0]), _ptr(v), _ptr(xf), ctypes.byref(s), _ptr(d))
        return s.value, d

    def ik01a(self, x):
        a, b = _d(), _d()
        self.lib.ref_ik01a(_d(x), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def jy01b(self, x):
        o = [_d() for _ in range(4)]
        self.lib.ref_jy01b(_d(x), *[ctypes.byref(v) for v in o])
        return tuple(v.value for v in o)

    def error(self, x):
        e = _d()
        self.lib.ref_error(_d(x), ctypes.byref(e))
        return e.value

    def e1xa(self, x):
        e = _d()
        self.lib.ref_e1xa(_d(x), ctypes.byref(e))
        return e.value

    def enxa(self, n, x):
        e = _d()
        self.lib.ref_enxa(_i(n), _d(x), ctypes.byref(e))
        return e.value

    def eix(self, x):
        e = _d()
        self.lib.ref_eix(_d(x), ctypes.byref(e))
        return e.value

    def limes(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64).copy()
        r = _d()
        self.lib.ref_limes(_i(len(s) - 1), _ptr(s), ctypes.byref(r))
        return r.value

    # ---- the multi-body chain (ref_shim_multi.f90) ---------------------------
    def multi_set(self, prob, nthreads=1):
        """the reference's *_MULTI_COMB module arrays from a meshes.Problem (hull panels)"""
        keep = [np.ascontiguousarray(prob.body_start, dtype=np.int32),
                np.ascontiguousarray(prob.xyz, dtype=np.float64),
                np.ascontiguousarray(prob.ncn, dtype=np.int32),
                np.ascontiguousarray(prob.ncon, dtype=np.int32),
                np.ascontiguousarray(prob.xyz_p, dtype=np.float64),
                np.ascontiguousarray(prob.ds, dtype=np.float64),
                np.ascontiguousarray(prob.pnsz, dtype=np.float64),
                np.ascontiguousarray(prob.dxyz_p, dtype=np.float64)]
        self.lib.ref_multi_set(_i(prob.nbody), _ptr(keep[0], _ip), _i(prob.n), _i(len(prob.xyz)), _ptr(keep[1]),
                               _ptr(keep[2], _ip), _ptr(keep[3], _ip), _ptr(keep[4]), _ptr(keep[5]),
                               _ptr(keep[6]), _ptr(keep[7]), _d(prob.h), _i(nthreads))
        self._multi = (prob.n, prob.nbody)

    def multi_freq(self, v, wk, w1, wvn):
        wvn = np.ascontiguousarray(wvn, dtype=np.float64)
        assert len(wvn) == 200
        self.lib.ref_multi_freq(_d(v), _d(wk), _d(w1), _ptr(wvn))

    def multi_calgreen(self):
        """CALGREEN_MULTI, then its tables RKBN (n, n, 4) and CGRN (n, n, 4)"""
        n, _ = self._multi
        self.lib.ref_multi_calgreen()
        rk = np.zeros((n, n, 4), order="F")
        cg = np.zeros((n, n, 4), dtype=np.complex128, order="F")
        self.lib.ref_multi_tables(_i(n), _ptr(rk), _ptr(cg))
        return rk, cg

    def multi_left(self):
        n, _ = self._multi
        a = np.zeros((n, n), dtype=np.complex128, order="F")
        self.lib.ref_multi_left(_i(n), _ptr(a))
        return a

    def multi_rbc(self):
        n, nb = self._multi
        b = np.zeros((n, 6 * nb), dtype=np.complex128, order="F")
        self.lib.ref_multi_rbc(_i(n), _i(nb), _ptr(b))
        return b

    def multi_dbc(self, beta):
        n, _ = self._multi
        b = np.zeros(n, dtype=np.complex128)
        self.lib.ref_multi_dbc(_i(n), _d(beta), _ptr(b))
        return b

    def multi_forces(self, mxpot, beta):
        """RFORCE_MULTI / EFORCE_MULTI for potentials mxpot (n, 6 nbody + 1: the
        radiation columns and the diffraction column of heading beta)"""
        n, nb = self._multi
        mx = np.asfortranarray(mxpot, dtype=np.complex128)
        assert mx.shape == (n, 6 * nb + 1)
        amas = np.zeros((6 * nb, 6 * nb), order="F")
        bdmp = np.zeros((6 * nb, 6 * nb), order="F")
        exfc = np.zeros(6 * nb, dtype=np.complex128)
        self.lib.ref_multi_forces(_i(n), _i(nb), _ptr(mx), _d(beta), _ptr(amas), _ptr(bdmp), _ptr(exfc))
        return amas, bdmp, exfc
