"""Writes tests/golden/*.npz from the reference's own kernels.

Runs in the build container only: it needs oracle/_ref/libhamsref.so, built by
oracle/Makefile from the unmodified reference sources (FinGreen3D.f90,
FinGrnExtSubs.f90, InfGreen_Appr.f90, SingularIntgr.f90, WavDynSubs.f90,
WavDynMods.f90) with ROCm flang. The fixtures are inputs + reference outputs only.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import Reference  # noqa: E402

G = 9.80665          # WavDynMods.f90:146 (V = w^2/G)


def green_finite(R, rng):
    """Stratified (R/h, z/h, zeta/h, omega, h) grid over all four regimes,
    including R = 0 and both FLAG values."""
    rows = []
    for h in (20.0, 50.0, 70.0):
        for w in (0.1, 0.5, 1.0, 2.0):
            wvn = R.dispersion(w, h)
            for rh in (0.0, 2e-5, 3e-4, 4.9e-4, 5e-4, 6e-4, 0.01, 0.049, 0.05, 0.06, 0.3, 0.49, 0.5, 0.6, 2.0, 5.0):
                for flag in (0, 1):
                    ang = rng.uniform(0, 2 * np.pi)
                    rr = rh * h
                    src = (1.0 + rr * np.cos(ang), 2.0 + rr * np.sin(ang), -rng.uniform(0.02, 0.95) * h)
                    fld = (1.0, 2.0, -rng.uniform(0.02, 0.95) * h)
                    g = R.fingreen3d(src, fld, w * w / G, wvn, h, flag)
                    rows.append((src, fld, w, h, flag, g))
    src = np.array([r[0] for r in rows])
    fld = np.array([r[1] for r in rows])
    return dict(src=src, fld=fld, w=np.array([r[2] for r in rows]), h=np.array([r[3] for r in rows]),
                flag=np.array([r[4] for r in rows], dtype=np.int32), grn=np.array([r[5] for r in rows]))


def green_infinite(R, rng):
    rows = []
    for w in (0.1, 0.5, 1.0, 2.0):
        for k in range(60):
            src = rng.uniform(-10, 10, 3)
            src[2] = -rng.uniform(0.01, 10)
            fld = rng.uniform(-10, 10, 3)
            fld[2] = -rng.uniform(0.01, 10)
            if k % 6 == 0:
                fld[:2] = src[:2]                      # R = 0
            flag = k % 2
            rows.append((src, fld, w, flag, R.infgreen3d(src, fld, w * w / G, flag)))
    return dict(src=np.array([r[0] for r in rows]), fld=np.array([r[1] for r in rows]),
                w=np.array([r[2] for r in rows]), flag=np.array([r[3] for r in rows], dtype=np.int32),
                grn=np.array([r[4] for r in rows]))


def rankine(R, rng):
    rows = []
    for k in range(200):
        c = rng.uniform(-3, 3, 3)
        e1, e2 = rng.normal(size=3), rng.normal(size=3)
        sz = rng.uniform(0.2, 2.0)
        q = np.array([c, c + sz * e1, c + sz * (e1 + e2), c + sz * e2])
        q[3] += rng.normal(size=3) * 0.01 * sz              # slightly warped quads
        ncn = 3 if k % 5 == 0 else 4
        if k % 11 == 0:
            q[3] = q[0]                                      # degenerate edge (centre-of-disk quads)
        xf = c + rng.normal(size=3) * sz * (0.1, 1.0, 5.0, 50.0)[k % 4]
        if k % 7 == 0:
            xf = q[:ncn].mean(0)                             # self panel
        s, d = R.sglintbd(q[:ncn], xf)
        v = np.zeros((4, 3))
        v[:ncn] = q[:ncn]
        rows.append((ncn, v, xf, s, d))
    return dict(ncn=np.array([r[0] for r in rows], dtype=np.int32), verts=np.array([r[1] for r in rows]),
                xf=np.array([r[2] for r in rows]), sij=np.array([r[3] for r in rows]),
                dij=np.array([r[4] for r in rows]))


def special(R, rng):
    x = np.concatenate([rng.uniform(0, 60, 200), [1e-6, 0.5, 1, 2, 3, 4, 8.9, 9, 9.1, 17.9, 18, 18.1, 35, 50, 59]])
    k = np.array([R.ik01a(v) for v in x])
    j = np.array([R.jy01b(v) for v in x])
    e1 = np.array([R.e1xa(v + 1e-3) for v in x])
    ei = np.array([R.eix(v + 1e-3) for v in x])
    en = np.array([[R.enxa(n, v + 1e-3) for n in range(6)] for v in x])
    xe = np.concatenate([rng.uniform(-6, 6, 200), [-3.5, 3.5, 0.0, 1.0, 2.0]])
    er = np.array([R.error(v) for v in xe])
    seqs = []
    for n in (3, 10, 30, 54, 99):
        s = np.cumsum(rng.normal(size=n + 1) * (-0.7) ** np.arange(n + 1))
        seqs.append((s, R.limes(s)))
    lim_s = np.zeros((len(seqs), 100))
    lim_n = np.array([len(s) - 1 for s, _ in seqs], dtype=np.int32)
    for i, (s, _) in enumerate(seqs):
        lim_s[i, :len(s)] = s
    return dict(x=x, k01=k, jy=j, e1=e1, ei=ei, en=en, xe=xe, erf=er, lim_s=lim_s, lim_n=lim_n,
                lim=np.array([r for _, r in seqs]))


def dispersion(R):
    ws = np.array([0.05, 0.1, 0.5, 1.0, 2.0, 2.5, 3.0])
    hs = np.array([5.0, 30.0, 50.0, 70.0, 200.0])
    out = np.array([[R.dispersion(w, h) for h in hs] for w in ws])
    return dict(w=ws, h=hs, wvn=out)


def main():
    R = Reference()
    rng = np.random.default_rng(20261203)
    np.savez_compressed(os.path.join(HERE, "green_finite.npz"), **green_finite(R, rng))
    np.savez_compressed(os.path.join(HERE, "green_infinite.npz"), **green_infinite(R, rng))
    np.savez_compressed(os.path.join(HERE, "rankine.npz"), **rankine(R, rng))
    np.savez_compressed(os.path.join(HERE, "special.npz"), **special(R, rng))
    np.savez_compressed(os.path.join(HERE, "dispersion.npz"), **dispersion(R))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
