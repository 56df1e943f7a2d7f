"""ctypes binding of the MI355X hot-path library (lib/libhm_bem.so, include/hm_bem.h).

This is the Python view of the same C-ABI the Fortran layer (fortran/hm_bem.f90)
binds. Loading fails loudly when the HIP library has not been built: there is no
CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhm_bem.so")

HM_FP64 = 0
HM_FP32_IR = 1
HM_GMRES = 2
SUR_PARAMS = 12995          # network parameters (see include/hm_bem.h)
STAGES = ("classify", "assembly", "rhs_gemm", "irr_normal_eq", "lu", "solve", "forces",
          "asm_linton", "asm_pidcock", "asm_eigene", "asm_eigen", "asm_surrogate", "asm_cross",
          "aca_entries")

_d = ctypes.c_double
_i = ctypes.c_int
_vp = ctypes.c_void_p

# every symbol include/hm_bem.h declares
EXPORTS = ("hm_create", "hm_destroy", "hm_error_string", "hm_set_constants", "hm_set_geometry",
           "hm_set_depth", "hm_set_pair_symmetry", "hm_set_headings", "hm_set_frequency", "hm_assemble", "hm_factor",
           "hm_solve", "hm_forces", "hm_solve_frequency", "hm_run_frequencies", "hm_set_timing",
           "hm_get_timing", "hm_get_system", "hm_get_potentials", "hm_get_normal_equations", "hm_fingreen3d_batch",
           "hm_infgreen3d_batch", "hm_fingreen3d_batch_noalloc", "hm_sglintbd_batch", "hm_ik01a_batch", "hm_field_points", "hm_solve_motion", "hm_dispersion", "hm_zgetrf", "hm_zgetrs",
           "hm_get_pair_histogram", "hm_set_gmres", "hm_get_gmres_stats",
           "hm_set_series_variant", "hm_surrogate_blob_size", "hm_surrogate_train", "hm_surrogate_get",
           "hm_surrogate_set", "hm_set_surrogate", "hm_surrogate_counts", "hm_surrogate_eval",
           "hm_surrogate_grad", "hm_set_hmatrix", "hm_set_hmatrix_overlap", "hm_get_hmatrix_stats",
           "hm_hmatrix_apply", "hm_set_frequency_limit", "hm_batch_create", "hm_batch_destroy",
           "hm_batch_error_string", "hm_batch_set_constants", "hm_batch_add_problem", "hm_batch_prepare",
           "hm_batch_solve", "hm_batch_info", "hm_hmatrix_build", "hm_hmatrix_release", "hm_ziter_solve",
           "hm_surrogate_predict", "hm_set_cluster_tree", "hm_get_cluster_tree_stats", "hm_hmatrix_build_tree",
           "hm_cluster_tree_partition", "hm_set_lu_panel", "hm_batch_set_waterplane", "hm_batch_tuning", "hm_batch_set_gmres", "hm_batch_gmres_stats", "hm_surrogate_train_deep",
           "hm_set_surrogate_deep", "hm_surrogate_deep_count", "hm_surrogate_eval_deep", "hm_set_body_origins",
           "hm_get_body_classes", "hm_set_block_sharing", "hm_get_block_sharing",
           "hm_set_workers", "hm_set_precond_fp16", "hm_cgemm_f16_check", "hm_cgemm_3m_check",
           "hm_get_body_symmetries")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C hams-mrel_amd` "
                               "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        L.hm_error_string.restype = ctypes.c_char_p
        L.hm_error_string.argtypes = [_vp]
        L.hm_batch_error_string.restype = ctypes.c_char_p
        L.hm_batch_error_string.argtypes = [_vp]
        for name in EXPORTS:
            if name not in ("hm_error_string", "hm_batch_error_string"):
                getattr(L, name).restype = _i
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class HMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hm_bem error {code}: {msg}")
        self.code = code


def dispersion(w, h, g_disp, nrt=200):
    wvn = np.zeros(nrt)
    rc = load().hm_dispersion(_p(wvn), _i(nrt), _d(w), _d(h), _d(g_disp))
    if rc:
        raise HMError(rc, "hm_dispersion")
    return wvn


def cluster_tree_partition(xyz, leaf=32, eta=2.0):
    """The cluster tree's permutation and block partition over the points xyz
    (n x 3), on the host: (perm, blocks) with blocks rows (rs, nr, cs, nc,
    admissible) in cluster positions"""
    x = np.asfortranarray(xyz, dtype=np.float64)
    n = x.shape[0]
    perm = np.zeros(n, dtype=np.int32)
    nb = _i(0)
    L = load()
    rc = L.hm_cluster_tree_partition(_i(n), _p(x), _i(n), _i(leaf), _d(eta), _p(perm), _i(0), ctypes.byref(nb), None)
    if rc:
        raise HMError(rc, "hm_cluster_tree_partition")
    blocks = np.zeros((nb.value, 5), dtype=np.int32)
    rc = L.hm_cluster_tree_partition(_i(n), _p(x), _i(n), _i(leaf), _d(eta), _p(perm), _i(nb.value),
                                     ctypes.byref(nb), _p(blocks))
    if rc:
        raise HMError(rc, "hm_cluster_tree_partition")
    return perm, blocks


class Context:
    """One HIP stream + the HBM-resident state of one problem."""

    def __init__(self):
        self.L = load()
        self.ptr = _vp()
        rc = self.L.hm_create(ctypes.byref(self.ptr))
        if rc:
            raise HMError(rc, "hm_create (no GPU?)")
        self.problem = None

    def close(self):
        if self.ptr:
            self.L.hm_destroy(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise HMError(rc, f"{what}: {self.L.hm_error_string(self.ptr).decode()}")

    # ---- problem set-up -------------------------------------------------------
    def setup(self, prob, frames=True):
        """prob: meshes.Problem. frames: where the problem's bodies carry body
        frames (local meshes and placements), upload the local geometry and
        the placements (hm_set_body_origins) instead of the global geometry"""
        self.problem = prob
        fr = frames and getattr(prob, "frames", False)
        src = prob.local if fr else dict(xyz=prob.xyz, xyz_p=prob.xyz_p, ds=prob.ds, pnsz=prob.pnsz,
                                         dxyz_p=prob.dxyz_p, ixyz_p=prob.ixyz_p)
        keep = dict(bs=np.ascontiguousarray(prob.body_start, dtype=np.int32),
                    xyz=np.ascontiguousarray(src["xyz"], dtype=np.float64),
                    ncn=np.ascontiguousarray(prob.ncn, dtype=np.int32),
                    ncon=np.ascontiguousarray(prob.ncon, dtype=np.int32),
                    xp=np.ascontiguousarray(src["xyz_p"], dtype=np.float64),
                    ds=np.ascontiguousarray(src["ds"], dtype=np.float64),
                    pn=np.ascontiguousarray(src["pnsz"], dtype=np.float64),
                    dx=np.ascontiguousarray(src["dxyz_p"], dtype=np.float64),
                    ix=np.ascontiguousarray(src["ixyz_p"] if prob.ni else np.zeros((1, 3)), dtype=np.float64),
                    beta=np.ascontiguousarray(prob.beta, dtype=np.float64))
        self._chk(self.L.hm_set_constants(self.ptr, _d(prob.rho), _d(prob.g), _d(prob.g_disp), _d(prob.amp),
                                          _d(prob.xw[0]), _d(prob.xw[1])), "hm_set_constants")
        self._chk(self.L.hm_set_geometry(self.ptr, _i(prob.nbody), _p(keep["bs"]), _i(len(prob.xyz)),
                                         _p(keep["xyz"]), _p(keep["ncn"]), _p(keep["ncon"]), _p(keep["xp"]),
                                         _p(keep["ds"]), _p(keep["pn"]), _p(keep["dx"]), _i(prob.ni),
                                         _p(keep["ix"])), "hm_set_geometry")
        self.frames = fr
        if fr:
            org = np.ascontiguousarray(prob.local["origin"], dtype=np.float64)
            wps = np.ascontiguousarray(prob.local["wp_start"], dtype=np.int32)
            self._chk(self.L.hm_set_body_origins(self.ptr, _i(prob.nbody), _p(org), _p(wps) if prob.ni else None),
                      "hm_set_body_origins")
        self._chk(self.L.hm_set_depth(self.ptr, _d(prob.h)), "hm_set_depth")
        self._chk(self.L.hm_set_headings(self.ptr, _i(prob.nbeta), _p(keep["beta"])), "hm_set_headings")

    def body_classes(self):
        """rep[b]: the first body whose local hull panels equal body b's"""
        rep = np.zeros(self.problem.nbody, dtype=np.int32)
        self._chk(self.L.hm_get_body_classes(self.ptr, _i(self.problem.nbody), _p(rep)), "hm_get_body_classes")
        return rep

    def block_sharing(self, with_rank=False):
        """(diagonal blocks assembled, cross-approximation processes run) per
        frequency of the H-matrix form; with_rank: and the mean rank of the
        processes that ran at the last assembly"""
        d, a, r = _i(0), _i(0), _d(0.0)
        self._chk(self.L.hm_get_block_sharing(self.ptr, ctypes.byref(d), ctypes.byref(a), ctypes.byref(r)),
                  "hm_get_block_sharing")
        return (d.value, a.value, r.value) if with_rank else (d.value, a.value)

    def body_symmetries(self):
        """flags of the maps identity, rotations by 90/180/270 degrees,
        reflections in x axis, y axis, y = x, y = -x that take the (equal)
        bodies' hull panels onto themselves"""
        ok = np.zeros(8, dtype=np.int32)
        self._chk(self.L.hm_get_body_symmetries(self.ptr, _p(ok)), "hm_get_body_symmetries")
        return ok.astype(bool)

    def set_block_sharing(self, enable=True, symmetries=True):
        """enable: equal bodies share diagonal blocks and, at equal offsets,
        cross approximations; symmetries: also at offsets that a symmetry of the
        body (rotations by multiples of 90 degrees, reflections in the axes and
        diagonals) maps onto each other"""
        mode = (2 if symmetries else 1) if enable else 0
        self._chk(self.L.hm_set_block_sharing(self.ptr, _i(mode)), "hm_set_block_sharing")

    def set_frequency(self, w1):
        self._chk(self.L.hm_set_frequency(self.ptr, _d(w1)), "hm_set_frequency")

    def assemble(self):
        self._chk(self.L.hm_assemble(self.ptr), "hm_assemble")

    def factor(self, precision=HM_FP64):
        self._chk(self.L.hm_factor(self.ptr, _i(precision)), "hm_factor")

    def solve(self):
        it = _i(0)
        self._chk(self.L.hm_solve(self.ptr, ctypes.byref(it)), "hm_solve")
        return it.value

    def forces(self):
        p = self.problem
        m, nb = p.m, p.nbeta
        amas = np.zeros((m, m), order="F")
        bdmp = np.zeros((m, m), order="F")
        exfc = np.zeros((m, max(nb, 1)), dtype=np.complex128, order="F")
        self._chk(self.L.hm_forces(self.ptr, _p(amas), _p(bdmp), _p(exfc)), "hm_forces")
        return amas, bdmp, exfc[:, :nb]

    def set_frequency_limit(self, which):
        """which: 0 zero frequency, 1 infinite frequency"""
        self._chk(self.L.hm_set_frequency_limit(self.ptr, _i(which)), "hm_set_frequency_limit")

    def solve_limit(self, which, precision=HM_FP64):
        self.set_frequency_limit(which)
        return self._finish(precision)

    def solve_frequency(self, w1, precision=HM_FP64):
        self.set_frequency(w1)
        return self._finish(precision)

    def _finish(self, precision):
        self.assemble()
        self.factor(precision)
        iters = self.solve()
        a, b, f = self.forces()
        return dict(amas=a, bdmp=b, exfc=f, iters=iters)

    def set_precond_fp16(self, enable=True):
        self._chk(self.L.hm_set_precond_fp16(self.ptr, _i(1 if enable else 0)), "hm_set_precond_fp16")

    def cgemm_f16(self, a, b):
        """a @ b with FP16 operands and FP32 accumulation on the device (a, b complex)"""
        a = np.asfortranarray(a, dtype=np.complex64)
        b = np.asfortranarray(b, dtype=np.complex64)
        m, k = a.shape
        n = b.shape[1]
        out = np.zeros((m, n), dtype=np.complex64, order="F")
        self._chk(self.L.hm_cgemm_f16_check(self.ptr, _i(m), _i(n), _i(k), _p(a), _p(b), _p(out)), "hm_cgemm_f16_check")
        return out

    def cgemm_3m(self, a, b, trans_a=False):
        """a @ b (a.T @ b with trans_a, not conjugated) in FP64 on the device,
        each complex product from three real ones"""
        a = np.asfortranarray(a, dtype=np.complex128)
        b = np.asfortranarray(b, dtype=np.complex128)
        k, m = a.shape if trans_a else a.shape[::-1]
        n = b.shape[1]
        out = np.zeros((m, n), dtype=np.complex128, order="F")
        self._chk(self.L.hm_cgemm_3m_check(self.ptr, _i(int(trans_a)), _i(m), _i(n), _i(k), _p(a), _p(b), _p(out)),
                  "hm_cgemm_3m_check")
        return out

    def set_workers(self, n=1):
        """frequencies in flight in run_frequencies (clones of this context)"""
        self._chk(self.L.hm_set_workers(self.ptr, _i(n)), "hm_set_workers")

    def run_frequencies(self, omegas, precision=HM_FP64):
        p = self.problem
        m, nb = p.m, p.nbeta
        w = np.ascontiguousarray(omegas, dtype=np.float64)
        amas = np.zeros((len(w), m, m))
        bdmp = np.zeros((len(w), m, m))
        exfc = np.zeros((len(w), nb, m), dtype=np.complex128)
        self._chk(self.L.hm_run_frequencies(self.ptr, _i(len(w)), _p(w), _i(precision), _p(amas), _p(bdmp),
                                            _p(exfc)), "hm_run_frequencies")
        # each frequency block is column-major (m x m), (m x nb)
        return amas.transpose(0, 2, 1), bdmp.transpose(0, 2, 1), exfc.transpose(0, 2, 1)

    def get_system(self):
        p = self.problem
        nt, n, ncol = p.n + p.ni, p.n, p.m + p.nbeta
        amat = np.zeros((nt, n), dtype=np.complex128, order="F")
        smat = np.zeros((nt, n), dtype=np.complex128, order="F")
        rhs = np.zeros((nt, ncol), dtype=np.complex128, order="F")
        self._chk(self.L.hm_get_system(self.ptr, _p(amat), _p(smat), _p(rhs)), "hm_get_system")
        return amat, smat, rhs

    def get_normal_equations(self):
        """(C, D): A^T A (n x n) and A^T of the right-hand sides, with waterplane panels"""
        p = self.problem
        cm = np.zeros((p.n, p.n), dtype=np.complex128, order="F")
        d = np.zeros((p.n, p.m + p.nbeta), dtype=np.complex128, order="F")
        self._chk(self.L.hm_get_normal_equations(self.ptr, _p(cm), _p(d)), "hm_get_normal_equations")
        return cm, d

    def get_potentials(self):
        p = self.problem
        x = np.zeros((p.n, p.m + p.nbeta), dtype=np.complex128, order="F")
        self._chk(self.L.hm_get_potentials(self.ptr, _p(x)), "hm_get_potentials")
        return x

    def pair_histogram(self):
        """(4, 128) pair counts by regime (Linton, Pidcock, eps-accelerated,
        eigenfunction) and series length"""
        h = np.zeros(512, dtype=np.int64)
        self._chk(self.L.hm_get_pair_histogram(self.ptr, _p(h), _i(512)), "hm_get_pair_histogram")
        return h.reshape(4, 128)

    def set_gmres(self, restart=50, tol=1e-6, maxiter=500, precond=-1):
        self._chk(self.L.hm_set_gmres(self.ptr, _i(restart), _d(tol), _i(maxiter), _i(precond)), "hm_set_gmres")

    # ---- learned surrogate ------------------------------------------------------
    def surrogate_train(self, nsamples=16_000_000, epochs=60, batch=4096, lr=2e-3, seed=20261204, quant=True):
        st = np.zeros(8)
        self._chk(self.L.hm_surrogate_train(self.ptr, ctypes.c_longlong(nsamples), _i(epochs), _i(batch), _d(lr),
                                            ctypes.c_ulonglong(seed), _i(1 if quant else 0), _p(st)),
                  "hm_surrogate_train")
        keys = ("train_loss", "val_rmse", "val_max_err", "seconds", "cells_below_1e-4", "cells_below_1e-3",
                "steps", "samples")
        return dict(zip(keys, st.tolist()))

    def surrogate_get(self):
        n = self.L.hm_surrogate_blob_size()
        b = np.zeros(n, dtype=np.float32)
        self._chk(self.L.hm_surrogate_get(self.ptr, _p(b), _i(n)), "hm_surrogate_get")
        return b

    def surrogate_set(self, blob):
        b = np.ascontiguousarray(blob, dtype=np.float32)
        self._chk(self.L.hm_surrogate_set(self.ptr, _p(b), _i(len(b))), "hm_surrogate_set")

    def set_surrogate(self, enable=True, tol=1e-4):
        self._chk(self.L.hm_set_surrogate(self.ptr, _i(1 if enable else 0), _d(tol)), "hm_set_surrogate")

    def surrogate_counts(self):
        k = np.zeros(3, dtype=np.int64)
        self._chk(self.L.hm_surrogate_counts(self.ptr, _p(k)), "hm_surrogate_counts")
        return dict(network=int(k[0]), fallback_eigene=int(k[1]), fallback_eigen=int(k[2]))

    def surrogate_eval(self, rh, zs, zd, kh):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (rh, zs, zd, kh)]
        n = len(a[0])
        tn = np.zeros((n, 3))
        tp = np.zeros((n, 3))
        bd = np.zeros(n)
        self._chk(self.L.hm_surrogate_eval(self.ptr, _i(n), *[_p(v) for v in a], _p(tn), _p(tp), _p(bd)),
                  "hm_surrogate_eval")
        return tn, tp, bd

    def surrogate_train_deep(self, nsamples=4_000_000, epochs=40, batch=4096, lr=2e-3, seed=20261205, quant=True):
        """the infinite-depth branch: Re GF and Re GFh of the Havelock function of (KR, K(z+zeta))"""
        st = np.zeros(8)
        self._chk(self.L.hm_surrogate_train_deep(self.ptr, ctypes.c_longlong(nsamples), _i(epochs), _i(batch),
                                                 _d(lr), ctypes.c_ulonglong(seed), _i(1 if quant else 0), _p(st)),
                  "hm_surrogate_train_deep")
        keys = ("train_loss", "val_rmse", "val_max_err", "seconds", "cells_below_1e-4", "cells_below_1e-3",
                "steps", "samples")
        return dict(zip(keys, st.tolist()))

    def set_surrogate_deep(self, enable=True, tol=1e-4):
        self._chk(self.L.hm_set_surrogate_deep(self.ptr, _i(1 if enable else 0), _d(tol)), "hm_set_surrogate_deep")

    def surrogate_deep_count(self):
        k = ctypes.c_longlong(0)
        self._chk(self.L.hm_surrogate_deep_count(self.ptr, ctypes.byref(k)), "hm_surrogate_deep_count")
        return k.value

    def surrogate_eval_deep(self, kr, kv):
        kr = np.ascontiguousarray(kr, dtype=np.float64)
        kv = np.ascontiguousarray(kv, dtype=np.float64)
        n = len(kr)
        net, plain, bd = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros(n)
        self._chk(self.L.hm_surrogate_eval_deep(self.ptr, _i(n), _p(kr), _p(kv), _p(net), _p(plain), _p(bd)),
                  "hm_surrogate_eval_deep")
        return net, plain, bd

    def surrogate_grad(self, x, u, quant=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        u = np.ascontiguousarray(u, dtype=np.float32)
        g = np.zeros(self.L.hm_surrogate_blob_size(), dtype=np.float32)
        loss = _d()
        self._chk(self.L.hm_surrogate_grad(self.ptr, _i(len(x)), _p(x), _p(u), _i(1 if quant else 0), _p(g),
                                           ctypes.byref(loss)), "hm_surrogate_grad")
        return g[:SUR_PARAMS], loss.value

    def set_hmatrix(self, enable=True, eps=1e-4, eps_s=None, kmax=50):
        """eps: ACA tolerance of A; eps_s: of S (default eps)"""
        self._chk(self.L.hm_set_hmatrix(self.ptr, _i(1 if enable else 0), _d(eps), _d(eps if eps_s is None else eps_s),
                                        _i(kmax)), "hm_set_hmatrix")

    def set_hmatrix_overlap(self, enable=True):
        """cross approximation beside the diagonal blocks (True) or after them"""
        self._chk(self.L.hm_set_hmatrix_overlap(self.ptr, _i(1 if enable else 0)), "hm_set_hmatrix_overlap")

    def set_pair_symmetry(self, enable=True):
        """H-matrix diagonal blocks: PIDCOCK pairs assembled with their mirrors"""
        self._chk(self.L.hm_set_pair_symmetry(self.ptr, _i(1 if enable else 0)), "hm_set_pair_symmetry")

    def hmatrix_stats(self):
        un, st, mx = _i(), _i(), _i()
        mr, fr = _d(), _d()
        self._chk(self.L.hm_get_hmatrix_stats(self.ptr, ctypes.byref(un), ctypes.byref(st), ctypes.byref(mr),
                                              ctypes.byref(mx), ctypes.byref(fr)), "hm_get_hmatrix_stats")
        return dict(unconverged=un.value, steps=st.value, mean_rank=mr.value, max_rank=mx.value,
                    stored_fraction=fr.value)

    def hmatrix_apply(self, mat, x):
        """H-matrix product of A (mat 0) or S (mat 1) with x (n x ncol complex)"""
        x = np.asarray(x, dtype=np.complex128)
        if x.ndim == 1:
            x = x[:, None]
        xf = np.asfortranarray(x)
        y = np.zeros_like(xf, order="F")
        self._chk(self.L.hm_hmatrix_apply(self.ptr, _i(mat), _i(x.shape[1]), _p(xf), _p(y)), "hm_hmatrix_apply")
        return y

    def hmatrix_build(self, a, block_sizes):
        """H-matrix of an assembled matrix a (n x n) with the given equal cluster blocks"""
        a = np.asfortranarray(a, dtype=np.complex128)
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        self._chk(self.L.hm_hmatrix_build(self.ptr, _i(a.shape[0]), _p(a), _i(a.shape[0]), _i(len(bs)), _p(bs)),
                  "hm_hmatrix_build")

    def hmatrix_build_tree(self, a, xyz, block_sizes=()):
        """H-matrix of an assembled matrix a (n x n) over the cluster tree of the
        points xyz (n x 3), with the block-diagonal preconditioner of block_sizes
        (none if empty)"""
        a = np.asfortranarray(a, dtype=np.complex128)
        x = np.asfortranarray(xyz, dtype=np.float64)
        n = a.shape[0]
        assert x.shape == (n, 3)
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        self._chk(self.L.hm_hmatrix_build_tree(self.ptr, _i(n), _p(a), _i(n), _p(x), _i(n), _i(len(bs)),
                                               _p(bs) if len(bs) else None), "hm_hmatrix_build_tree")

    def set_lu_panel(self, cooperative=True):
        """dense LU panels: one cooperative launch each (True) or per-column launches"""
        self._chk(self.L.hm_set_lu_panel(self.ptr, _i(1 if cooperative else 0)), "hm_set_lu_panel")

    def set_cluster_tree(self, mode=0, leaf=32, eta=2.0):
        self._chk(self.L.hm_set_cluster_tree(self.ptr, _i(mode), _i(leaf), _d(eta)), "hm_set_cluster_tree")

    def cluster_tree_stats(self):
        v = [_i() for _ in range(5)]
        self._chk(self.L.hm_get_cluster_tree_stats(self.ptr, *[ctypes.byref(t) for t in v]),
                  "hm_get_cluster_tree_stats")
        return dict(zip(("clusters", "leaves", "dense_blocks", "lowrank_blocks", "dense_fallbacks"),
                        (t.value for t in v)))

    def hmatrix_release(self):
        self._chk(self.L.hm_hmatrix_release(self.ptr), "hm_hmatrix_release")

    def ziter_solve(self, isolv, a, b, block_sizes=()):
        """GMRES solve of a x = b (isolv 1: dense a; 2: the built H-matrix); returns (x, info)"""
        b = np.asfortranarray(b, dtype=np.complex128)
        if b.ndim == 1:
            b = b[:, None]
        n, nrhs = b.shape
        x = np.zeros_like(b, order="F")
        a = np.asfortranarray(a, dtype=np.complex128) if a is not None else None
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        info = _i(0)
        self._chk(self.L.hm_ziter_solve(self.ptr, _i(isolv), _i(n), _p(a), _i(n), _p(b), _i(n), _p(x), _i(n),
                                        _i(nrhs), _i(len(bs)), _p(bs) if len(bs) else None, ctypes.byref(info)),
                  "hm_ziter_solve")
        return x, info.value

    def surrogate_predict(self, x):
        """scaled network outputs u (n x 3) for normalised inputs x (n x 4)"""
        x = np.ascontiguousarray(x, dtype=np.float32)
        u = np.zeros((x.shape[0], 3), dtype=np.float32)
        self._chk(self.L.hm_surrogate_predict(self.ptr, _i(x.shape[0]), _p(x), _p(u)), "hm_surrogate_predict")
        return u

    def set_series_variant(self, variant):
        self._chk(self.L.hm_set_series_variant(self.ptr, _i(variant)), "hm_set_series_variant")

    def gmres_stats(self):
        it, cy, cv = _i(), _i(), _i()
        rr = _d()
        self._chk(self.L.hm_get_gmres_stats(self.ptr, ctypes.byref(it), ctypes.byref(cy), ctypes.byref(cv),
                                            ctypes.byref(rr)), "hm_get_gmres_stats")
        return dict(iterations=it.value, cycles=cy.value, converged=bool(cv.value), rel_residual=rr.value)

    def set_timing(self, on=True):
        self._chk(self.L.hm_set_timing(self.ptr, _i(1 if on else 0)), "hm_set_timing")

    def timing(self):
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), dtype=np.int32)
        self._chk(self.L.hm_get_timing(self.ptr, _p(ms), _p(cnt), _i(len(STAGES))), "hm_get_timing")
        return {s: (float(ms[k]), int(cnt[k])) for k, s in enumerate(STAGES)}

    # ---- element-wise kernels --------------------------------------------------
    def fingreen3d(self, src, fld, v, wvn, h, flag):
        """src, fld: (n,3) source/field points; flag: (n,) -> (n,4) complex GRN"""
        src = np.asarray(src, dtype=np.float64)
        fld = np.asarray(fld, dtype=np.float64)
        n = len(src)
        cols = [np.ascontiguousarray(a) for a in (src[:, 0], fld[:, 0], src[:, 1], fld[:, 1], src[:, 2], fld[:, 2])]
        fl = np.ascontiguousarray(flag, dtype=np.int32)
        wvn = np.ascontiguousarray(wvn, dtype=np.float64)
        out = np.zeros((n, 4), dtype=np.complex128)
        self._chk(self.L.hm_fingreen3d_batch(self.ptr, _i(n), *[_p(c) for c in cols], _d(v), _p(wvn), _i(len(wvn)),
                                             _d(h), _p(fl), _p(out)), "hm_fingreen3d_batch")
        return out

    def infgreen3d(self, src, fld, v, flag):
        src = np.asarray(src, dtype=np.float64)
        fld = np.asarray(fld, dtype=np.float64)
        n = len(src)
        cols = [np.ascontiguousarray(a) for a in (src[:, 0], fld[:, 0], src[:, 1], fld[:, 1], src[:, 2], fld[:, 2])]
        fl = np.ascontiguousarray(flag, dtype=np.int32)
        out = np.zeros((n, 4), dtype=np.complex128)
        self._chk(self.L.hm_infgreen3d_batch(self.ptr, _i(n), *[_p(c) for c in cols], _d(v), _p(fl), _p(out)),
                  "hm_infgreen3d_batch")
        return out

    def sglintbd(self, ncn, verts, xf):
        ncn = np.ascontiguousarray(ncn, dtype=np.int32)
        verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(len(ncn), 12)
        xf = np.ascontiguousarray(xf, dtype=np.float64).reshape(len(ncn), 3)
        s = np.zeros(len(ncn))
        d = np.zeros((len(ncn), 3))
        self._chk(self.L.hm_sglintbd_batch(self.ptr, _i(len(ncn)), _p(ncn), _p(verts), _p(xf), _p(s), _p(d)),
                  "hm_sglintbd_batch")
        return s, d

    def field_points(self, xfp):
        """potential, pressure and elevation of every radiation mode and heading at
        the points xfp (nfp x 3) after solve (CalPotentialMulti and its pressure and
        elevation forms); each nfp x (6 nbody + nbeta) complex"""
        xfp = np.ascontiguousarray(xfp, dtype=np.float64).reshape(-1, 3)
        nfp = len(xfp)
        ncol = self.problem.m + self.problem.nbeta
        out = {k: np.zeros((nfp, ncol), dtype=np.complex128, order="F") for k in ("pot", "prs", "elv")}
        self._chk(self.L.hm_field_points(self.ptr, _i(nfp), _p(xfp), _p(out["pot"]), _p(out["prs"]), _p(out["elv"])),
                  "hm_field_points")
        return out

    def solve_motion(self, w1, amas, bdmp, exfc, matx, crs, kstf=None, blnr=None, bqdr=None):
        """motion response of one heading (SolveMotionMulti): amas, bdmp (6nb x 6nb),
        exfc (6nb complex), body matrices matx, crs, kstf, blnr, bqdr (nbody x 6 x 6)"""
        nb = np.asarray(matx).shape[0]
        f = lambda a: np.asfortranarray(np.zeros((nb, 6, 6)) if a is None else a, dtype=np.float64)  # noqa: E731
        args = [np.asfortranarray(amas, dtype=np.float64), np.asfortranarray(bdmp, dtype=np.float64),
                f(blnr), f(bqdr), f(matx), f(crs), f(kstf)]
        ex = np.ascontiguousarray(exfc, dtype=np.complex128)
        x = np.zeros(6 * nb, dtype=np.complex128)
        it = ctypes.c_int(0)
        self._chk(self.L.hm_solve_motion(self.ptr, _i(nb), _d(w1), *[_p(a) for a in args], _p(ex), _p(x),
                                         ctypes.byref(it)), "hm_solve_motion")
        return x, it.value

    def ik01a(self, x, fast=False):
        """K0(x), K1(x) on the GPU: the reference's IK01A evaluation, or (fast) the
        mixed-precision kernels' one"""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        k0, k1 = np.zeros(len(x)), np.zeros(len(x))
        self._chk(self.L.hm_ik01a_batch(self.ptr, _i(len(x)), _p(x), _p(k0), _p(k1), _i(1 if fast else 0)),
                  "hm_ik01a_batch")
        return k0, k1

    def zgetrf(self, a):
        a = np.array(a, dtype=np.complex128, order="F", copy=True)
        n = a.shape[0]
        ipiv = np.zeros(n, dtype=np.int32)
        info = _i(0)
        self._chk(self.L.hm_zgetrf(self.ptr, _i(n), _i(a.shape[1]), _p(a), _i(n), _p(ipiv), ctypes.byref(info)),
                  "hm_zgetrf")
        return a, ipiv, info.value

    def zgetrs(self, lu, ipiv, b):
        b = np.array(b, dtype=np.complex128, order="F", copy=True)
        if b.ndim == 1:
            b = b[:, None]
        n = lu.shape[0]
        info = _i(0)
        self._chk(self.L.hm_zgetrs(self.ptr, ctypes.c_char(b"N"), _i(n), _i(b.shape[1]), _p(lu), _i(n), _p(ipiv),
                                   _p(b), _i(n), ctypes.byref(info)), "hm_zgetrs")
        return b


class Batch:
    """Many independent problems resident at once, solved together per frequency
    (hm_batch_* in include/hm_bem.h)."""

    def __init__(self, problems, rho=None, g=None, g_disp=None, amp=None, xw=None):
        self.L = load()
        self.ptr = _vp()
        rc = self.L.hm_batch_create(ctypes.byref(self.ptr))
        if rc:
            raise HMError(rc, "hm_batch_create (no GPU?)")
        self.problems = list(problems)
        p0 = self.problems[0]
        self._chk(self.L.hm_batch_set_constants(
            self.ptr, _d(p0.rho if rho is None else rho), _d(p0.g if g is None else g),
            _d(p0.g_disp if g_disp is None else g_disp), _d(p0.amp if amp is None else amp),
            _d((p0.xw if xw is None else xw)[0]), _d((p0.xw if xw is None else xw)[1])), "hm_batch_set_constants")
        for prob in self.problems:
            keep = [np.ascontiguousarray(prob.body_start, dtype=np.int32),
                    np.ascontiguousarray(prob.xyz, dtype=np.float64),
                    np.ascontiguousarray(prob.ncn, dtype=np.int32),
                    np.ascontiguousarray(prob.ncon, dtype=np.int32),
                    np.ascontiguousarray(prob.xyz_p, dtype=np.float64),
                    np.ascontiguousarray(prob.ds, dtype=np.float64),
                    np.ascontiguousarray(prob.pnsz, dtype=np.float64),
                    np.ascontiguousarray(prob.dxyz_p, dtype=np.float64),
                    np.ascontiguousarray(prob.beta if prob.nbeta else np.zeros(1), dtype=np.float64)]
            rc = self.L.hm_batch_add_problem(self.ptr, _i(prob.nbody), _p(keep[0]), _i(len(prob.xyz)), _p(keep[1]),
                                             _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]), _p(keep[6]),
                                             _p(keep[7]), _d(prob.h), _i(prob.nbeta), _p(keep[8]))
            if rc < 0:
                self._chk(rc, "hm_batch_add_problem")
            if prob.ni:
                ix = np.ascontiguousarray(prob.ixyz_p, dtype=np.float64)
                self._chk(self.L.hm_batch_set_waterplane(self.ptr, _i(rc), _i(prob.ni), _p(ix)),
                          "hm_batch_set_waterplane")

    def close(self):
        if self.ptr:
            self.L.hm_batch_destroy(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise HMError(rc, f"{what}: {self.L.hm_batch_error_string(self.ptr).decode()}")

    def prepare(self, precision=HM_FP64, autotune=True, use_graph=True, nrefine=3):
        self._chk(self.L.hm_batch_prepare(self.ptr, _i(precision), _i(1 if autotune else 0),
                                          _i(1 if use_graph else 0), _i(nrefine)), "hm_batch_prepare")

    def solve(self, w1):
        """w1: one frequency for all members, or one per member. Returns a list of
        dict(amas, bdmp, exfc) per member."""
        w = np.broadcast_to(np.asarray(w1, dtype=np.float64), (len(self.problems),)).copy()
        na = sum(p.m * p.m for p in self.problems)
        ne = sum(p.m * p.nbeta for p in self.problems)
        amas = np.zeros(max(na, 1))
        bdmp = np.zeros(max(na, 1))
        exfc = np.zeros(max(ne, 1), dtype=np.complex128)
        self._chk(self.L.hm_batch_solve(self.ptr, _p(w), _p(amas), _p(bdmp), _p(exfc)), "hm_batch_solve")
        out, oa, oe = [], 0, 0
        for p in self.problems:
            m = p.m
            out.append(dict(amas=amas[oa:oa + m * m].reshape((m, m), order="F"),
                            bdmp=bdmp[oa:oa + m * m].reshape((m, m), order="F"),
                            exfc=exfc[oe:oe + m * p.nbeta].reshape((m, p.nbeta), order="F")))
            oa += m * m
            oe += m * p.nbeta
        return out

    def set_gmres(self, restart=50, tol=1e-6, maxiter=500):
        """settings of the HM_GMRES members (before prepare)"""
        self._chk(self.L.hm_batch_set_gmres(self.ptr, _i(restart), _d(tol), _i(maxiter)), "hm_batch_set_gmres")

    def gmres_stats(self):
        it, cy, ok = _i(), _i(), _i()
        rr = _d()
        self._chk(self.L.hm_batch_gmres_stats(self.ptr, ctypes.byref(it), ctypes.byref(cy), ctypes.byref(ok),
                                              ctypes.byref(rr)), "hm_batch_gmres_stats")
        return dict(max_iterations=it.value, max_cycles=cy.value, all_converged=bool(ok.value),
                    worst_rel_residual=rr.value)

    def tuning(self):
        """the autotuner's choices: series variant, K chunk of the grouped GEMM
        tiles, assembly time, grouped LU time at K chunks 8, 16, 32"""
        var, kb, ams = _i(), _i(), _d()
        lu = np.zeros(3)
        self._chk(self.L.hm_batch_tuning(self.ptr, ctypes.byref(var), ctypes.byref(kb), ctypes.byref(ams), _p(lu)),
                  "hm_batch_tuning")
        return dict(series_variant=var.value, gemm_kchunk=kb.value, assembly_ms=ams.value, lu_ms=lu.tolist())

    def info(self):
        mem, nmax, var, grid, es = _i(), _i(), _i(), _i(), _i()
        tms, corr = _d(), _d()
        nbytes = ctypes.c_longlong()
        self._chk(self.L.hm_batch_info(self.ptr, ctypes.byref(mem), ctypes.byref(nmax), ctypes.byref(var),
                                       ctypes.byref(grid), ctypes.byref(tms), ctypes.byref(corr), ctypes.byref(es),
                                       ctypes.byref(nbytes)), "hm_batch_info")
        return dict(members=mem.value, nmax=nmax.value, series_variant=var.value, grid_blocks=grid.value,
                    tuned_assembly_ms=tms.value, last_correction=corr.value, extra_sweeps=es.value,
                    device_bytes=nbytes.value)
