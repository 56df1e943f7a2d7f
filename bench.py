"""Benchmark: panel-frequency systems solved per second on one MI355X.

One step = one panel-frequency system of the workload (assemble the influence
matrix and right-hand sides, factor, solve the 6*NBODY radiation + NBETA
diffraction problems, reduce to added mass / damping / excitation), cycling
through the workload's wave frequencies. Inputs are resident in HBM when the
timed region starts (geometry uploaded and pairs classified beforehand).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5]

Prints one JSON line: metric, value (systems/s of the mixed-precision path),
unit, the plain FP64 path's rate beside it, per-stage milliseconds, the
roofline of the dominant assembly kernel and the CPU baseline.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hams-mrel_amd"))

METRIC = "panel-freq systems solved/s, mixed+surrogate vs plain FP64 GPU at ≤1e-4 rel err"
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            pass
    raise RuntimeError("HIP runtime not found")


def sync():
    _hip().hipDeviceSynchronize()


REGIME_STAGE = {"asm_linton": 0, "asm_pidcock": 1, "asm_eigene": 2, "asm_eigen": 3}

# HBM bytes per launch of the dominant kernel, measured on one MI355X by two
# counter passes of `bench.py --steps 1 --assemble-only` (rocprofv3 --pmc
# FETCH_SIZE, then --pmc WRITE_SIZE, kernel filter k_assemble|k_aca_eval), scaled
# by the readings of a 1 GiB device-to-device copy (fetch x2, write x1). The
# H-matrix path's PIDCOCK kernel assembles pairs with their mirrors from the
# cached frequency-independent term halves (31 GB read per launch); it runs at 5
# waves per SIMD, and register-spill traffic (scratch that misses the L2) is
# most of the rest.
MEASURED_TRAFFIC = {
    ("c5", "hmatrix", "asm_pidcock"): 57449828.7 * 1024 * 2 + 71234060.3 * 1024,
}


def flop_model(name):
    """algorithmic flops per pair by classification key (tests/golden/flop_model.json)"""
    path = os.path.join(ROOT, "tests", "golden", "flop_model.json")
    with open(path) as f:
        t = json.load(f)[name]
    return {int(k): v for k, v in t["flops_per_pair"].items()}, t["deep"]


def calibrate_copy(nbytes=1 << 30):
    """one device-to-device copy of known size: the reference reading for the
    FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)"""
    hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes))
    hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes))
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration_copy_bytes": nbytes}))


def workload(name):
    import meshes
    if name == "c1":
        return meshes.c1_single_cylinder()
    if name == "c2":
        return meshes.c2_three_cylinders()
    if name == "c4":
        return meshes.c4_eight_cylinders()
    if name == "c5":
        return meshes.c5_sixteen_cylinders()
    raise SystemExit(f"unknown workload {name}")


PATH_DESC = {
    "gmres": ("FP64 assembly (learned BF16 surrogate for the evanescent sums where its validated error bound "
              "allows, plain series elsewhere) + restarted GMRES(50), tol 1e-6: FP32 complex GEMM products on "
              "MFMA, FP32 block-diagonal LU preconditioner (one block per body), FP64 residuals"),
    "fp32ir": "FP64 assembly + FP32 LU (MFMA) + FP64 iterative refinement",
    "hmatrix": ("H-matrix: dense FP64 diagonal block per body (learned BF16 surrogate for the evanescent sums "
                "where its validated error bound allows), off-diagonal blocks by adaptive cross approximation "
                "evaluated straight from the FP64 Green function, right-hand sides S W from the H-matrix of S; "
                "restarted GMRES(50), tol 1e-6, with FP32 MFMA H-matrix products, FP32 block-diagonal LU "
                "preconditioner and FP64 H-matrix residuals"),
}


def run_steps(ctx, omegas, k0, steps, precision):
    out = []
    for s in range(steps):
        w = float(omegas[(k0 + s) % len(omegas)])
        out.append(ctx.solve_frequency(w, precision))
    return out


def time_path(ctx, prob, precision, steps, warmup, stage_times=True):
    import hm_bem
    run_steps(ctx, prob.omegas, 0, warmup, precision)
    sync()
    ctx.set_timing(stage_times)
    t0 = time.perf_counter()
    res = run_steps(ctx, prob.omegas, warmup, steps, precision)
    sync()
    dt = time.perf_counter() - t0
    stages = ctx.timing()
    ctx.set_timing(False)
    return dt, res, stages


def cpu_baseline(prob, seconds=20.0, threads=16):
    """The reference's own CPU kernels timed on the host: the body of its
    CALGREEN table-fill loop (SGLINTBD + FINGREEN3D/INFGREEN3D, compiled from the
    reference sources into oracle/_ref with OpenMP, CalGreenFunc.f90:58-115) over
    a random sample of the workload's hull panel pairs on `threads` cores, plus
    LAPACK ZGETRF/ZGETRS (numpy) at N <= 3000 scaled by N^3 to the workload's N.
    Returns systems/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as om
    if not om.Reference.available():
        return {"value": None, "unit": "systems/s", "cores": threads, "kind": "reference",
                "sample": "oracle/_ref/libhamsref.so not built"}
    R = om.Reference()
    L = R.lib
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
    n, nt = prob.n, prob.n + prob.ni
    xyz = np.ascontiguousarray(prob.xyz)
    ncon = np.ascontiguousarray(prob.ncon, dtype=np.int32)
    ncn = np.ascontiguousarray(prob.ncn, dtype=np.int32)
    xp = np.ascontiguousarray(prob.xyz_p)
    pn = np.ascontiguousarray(prob.pnsz)
    L.ref_set_mesh(ctypes.c_int(n), ctypes.c_int(len(xyz)), dp(xyz), dp(ncn), dp(ncon), dp(xp), dp(pn))
    w = float(prob.omegas[len(prob.omegas) // 2])
    v = w * w / prob.g
    wvn = R.dispersion(w, prob.h) if prob.h > 0 else np.zeros(200)
    rng = np.random.default_rng(20261203)
    csum = np.zeros(2)

    def run(np_):
        iel = np.ascontiguousarray(rng.integers(1, n + 1, np_), dtype=np.int32)
        jel = np.ascontiguousarray(rng.integers(1, n + 1, np_), dtype=np.int32)
        t0 = time.perf_counter()
        L.ref_pairs_omp(ctypes.c_int(np_), dp(iel), dp(jel), ctypes.c_double(prob.h), ctypes.c_double(v), dp(wvn),
                        ctypes.c_int(200), ctypes.c_int(threads), dp(csum))
        return time.perf_counter() - t0

    t = run(4000)
    npairs = int(min(5e7, max(4000, 4000 * 0.7 * seconds / max(t, 1e-6))))
    t = run(npairs)
    pair_rate = npairs / t
    m = min(n, 3000)
    a = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    b = rng.normal(size=(m, prob.m + prob.nbeta)) + 1j * rng.normal(size=(m, prob.m + prob.nbeta))
    t0 = time.perf_counter()
    np.linalg.solve(a, b)
    t_lu = (time.perf_counter() - t0) * (n / m) ** 3
    t_sys = nt * n / pair_rate + t_lu
    return {"value": 1.0 / t_sys, "unit": "systems/s", "cores": threads, "kind": "reference",
            "sample": (f"{npairs} random hull panel pairs of the workload at omega={w:.3f} through the reference's "
                       f"CALGREEN loop body (oracle/_ref, OpenMP, {threads} threads) in {t:.1f} s = {pair_rate:.3g} "
                       f"pairs/s, x {nt * n} pairs per system; + LAPACK zgesv (numpy) at N={m} scaled by "
                       f"(N/{m})^3 = {t_lu:.2f} s; {t_sys:.1f} s per system")}


def bench_batch(args):
    """The batching workload: 64 independent random problems (boxes and
    cylinders, 512-6144 panels, infinite or finite depth), 4 headings, solved
    together per frequency by the batch path (assembly launches across members,
    grouped LU, one HIP graph replay per frequency) against the same problems
    solved one after another by single-problem contexts (plain FP64)."""
    import hm_bem
    import meshes
    probs = meshes.c3_random_problems()
    omegas = probs[0].omegas
    nsys = len(probs)
    print(f"# workload C3: {nsys} problems, n {min(p.n for p in probs)}..{max(p.n for p in probs)}, "
          f"{sum(p.h < 0 for p in probs)} infinite depth, sum n^2 = {sum(p.n * p.n for p in probs):.3e}",
          file=sys.stderr, flush=True)
    prec = hm_bem.HM_FP32_IR if args.path == "fp32ir" else hm_bem.HM_FP64
    b = hm_bem.Batch(probs)
    t0 = time.perf_counter()
    b.prepare(prec, autotune=True, use_graph=True, nrefine=3)
    sync()
    info = b.info()
    print(f"# batch prepared in {time.perf_counter() - t0:.1f} s: {info}", file=sys.stderr, flush=True)
    for s in range(args.warmup):
        b.solve(float(omegas[s % len(omegas)]))
    sync()
    t0 = time.perf_counter()
    res = []
    for s in range(args.steps):
        res.append(b.solve(float(omegas[(args.warmup + s) % len(omegas)])))
    sync()
    dt = time.perf_counter() - t0
    info = b.info()
    b.close()
    value = nsys * args.steps / dt
    print(f"# batch: {args.steps} frequencies x {nsys} problems in {dt:.3f} s", file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        ctxs = []
        for p in probs:
            c = hm_bem.Context()
            c.setup(p)
            ctxs.append(c)
        sync()
        for c in ctxs:
            c.solve_frequency(float(omegas[0]), hm_bem.HM_FP64)
        sync()
        t0 = time.perf_counter()
        ref = []
        for s in range(args.steps):
            w = float(omegas[(args.warmup + s) % len(omegas)])
            ref.append([c.solve_frequency(w, hm_bem.HM_FP64) for c in ctxs])
        sync()
        dtp = time.perf_counter() - t0
        for c in ctxs:
            c.close()
        err = 0.0
        for rs, ps in zip(res, ref):
            for r, p in zip(rs, ps):
                for k in ("amas", "bdmp", "exfc"):
                    mx = np.max(np.abs(p[k]))
                    if mx > 0:
                        err = max(err, float(np.max(np.abs(r[k] - p[k])) / mx))
        plain = {"value": nsys * args.steps / dtp, "ms_per_step": 1e3 * dtp / args.steps,
                 "speedup_batch_vs_sequential": dtp / dt, "max_rel_err_batch_vs_sequential": err}
        print(f"# sequential single-problem FP64: {dtp:.3f} s; {plain}", file=sys.stderr, flush=True)
    line = {"metric": "panel-freq systems solved/s (batch of independent problems)", "value": value,
            "unit": "systems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if prec == hm_bem.HM_FP64 else "f64+f32",
            "data": "synthetic (seeded random boxes and cylinders of the named workload)",
            "config": {"workload": "C3 64 random problems", "problems": nsys,
                       "panels_min": min(p.n for p in probs), "panels_max": max(p.n for p in probs),
                       "headings": probs[0].nbeta, "frequencies": len(omegas),
                       "path": ("batch: one assembly launch per regime over all members' pairs, grouped LU and "
                                "solves, one HIP graph replay per frequency, autotuned series variant; "
                                + ("FP64 LU" if prec == hm_bem.HM_FP64 else "FP32 LU + FP64 refinement")),
                       "parallelism": "1 GPU"},
            "batch": info, "sequential_fp64": plain}
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain FP64 path")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--path", choices=("hmatrix", "gmres", "fp32ir"), default="hmatrix",
                    help="the mixed-precision solve that `value` measures")
    ap.add_argument("--aca-eps", type=float, default=2e-5, help="cross-approximation tolerance of A")
    ap.add_argument("--aca-eps-s", type=float, default=1e-6, help="cross-approximation tolerance of S")
    ap.add_argument("--aca-kmax", type=int, default=50, help="largest rank of a cross-approximated block")
    ap.add_argument("--surrogate-tol", type=float, default=1.5e-3,
                    help="cell error bound up to which the learned surrogate replaces the plain series (0: off)")
    ap.add_argument("--surrogate-samples", type=int, default=16_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=60)
    ap.add_argument("--gmres-tol", type=float, default=1e-6)
    ap.add_argument("--gmres-precond", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="block preconditioner: 1 triangular solves, 2 explicit FP32 inverses, 0 none, "
                         "-1 library default")
    ap.add_argument("--series-variant", type=int, default=1, choices=(0, 1, 2),
                    help="epsilon-accelerated regimes: 1 lane groups of 8/16 (default), 2 of 4/8, 0 one pair per lane")
    ap.add_argument("--with-fp32ir", action="store_true", help="also time the FP32 LU + refinement path")
    ap.add_argument("--field-points", type=int, default=0,
                    help="also time the potential, pressure and elevation at this many field points after the "
                         "last measured system (CalPotentialMulti)")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run the traffic-counter calibration copy")
    ap.add_argument("--assemble-only", action="store_true",
                    help="only assemble --steps systems (for hardware-counter passes over the assembly kernels)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: frequencies shard across GPUs "
                                   "without communication, but only --gpus 1 is supported here"}))
        return 2
    if args.workload == "c3":
        return bench_batch(args)
    import hm_bem
    prob = workload(args.workload)
    print(f"# workload {prob.name}: n={prob.n} ni={prob.ni} modes={prob.m} headings={prob.nbeta} "
          f"freqs={len(prob.omegas)}", file=sys.stderr, flush=True)
    ctx = hm_bem.Context()
    t0 = time.perf_counter()
    ctx.setup(prob)
    sync()
    print(f"# setup (upload + pair classification) {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    ctx.set_series_variant(args.series_variant)
    if args.assemble_only:
        if args.path == "hmatrix":
            ctx.set_hmatrix(True, eps=args.aca_eps, eps_s=args.aca_eps_s, kmax=args.aca_kmax)
        ctx.set_timing(True)
        for s in range(args.steps):
            ctx.set_frequency(float(prob.omegas[s % len(prob.omegas)]))
            ctx.assemble()
        sync()
        st = ctx.timing()
        ctx.set_timing(False)
        print(json.dumps({"assembled_systems": args.steps, "workload": prob.name, "path": args.path,
                          "series_variant": args.series_variant, "stages_ms": {k: v[0] for k, v in st.items()}}))
        ctx.close()
        return 0

    # learned surrogate of the evanescent sums: trained on the GPU before the
    # timed region (finite depth only)
    sur = None
    if args.surrogate_tol > 0 and prob.h > 0:
        t0 = time.perf_counter()
        st = ctx.surrogate_train(nsamples=args.surrogate_samples, epochs=args.surrogate_epochs, batch=4096,
                                 lr=2e-3, seed=20261204, quant=True)
        sur = {"train": st, "train_wall_s": time.perf_counter() - t0, "tol": args.surrogate_tol}
        print(f"# surrogate trained: {st}", file=sys.stderr, flush=True)
        ctx.set_surrogate(True, args.surrogate_tol)
    # mixed precision path: GMRES with FP32 MFMA products and FP64 residuals
    # (default), or FP32 LU on MFMA + FP64 refinement
    prec = {"hmatrix": hm_bem.HM_GMRES, "gmres": hm_bem.HM_GMRES, "fp32ir": hm_bem.HM_FP32_IR}[args.path]
    if prec == hm_bem.HM_GMRES:
        ctx.set_gmres(restart=50, tol=args.gmres_tol, precond=args.gmres_precond)
    hmat = None
    hm_setup_s = None
    if args.path == "hmatrix":
        # once per geometry and depth, outside the timed region: the diagonal
        # blocks' pair classification and the frequency-independent halves of
        # their PIDCOCK terms
        sync()
        t0 = time.perf_counter()
        ctx.set_hmatrix(True, eps=args.aca_eps, eps_s=args.aca_eps_s, kmax=args.aca_kmax)
        sync()
        hm_setup_s = time.perf_counter() - t0
        print(f"# hmatrix set-up (diagonal-block classification, PIDCOCK term cache) {hm_setup_s:.2f} s",
              file=sys.stderr, flush=True)
    dt_mx, res_mx, st_mx = time_path(ctx, prob, prec, args.steps, args.warmup)
    gm = ctx.gmres_stats() if prec == hm_bem.HM_GMRES else None
    field = None
    if args.field_points > 0:
        # a grid of points around the bodies, from the free surface down to the draft
        k = args.field_points
        lo, hi = prob.xyz_p.min(axis=0), prob.xyz_p.max(axis=0)
        rng = np.random.default_rng(20261205)
        xfp = np.column_stack([rng.uniform(lo[0] - 5.0, hi[0] + 5.0, k), rng.uniform(lo[1] - 5.0, hi[1] + 5.0, k),
                               rng.uniform(lo[2], 0.0, k)])
        ctx.field_points(xfp[:8])
        sync()
        t0 = time.perf_counter()
        ctx.field_points(xfp)
        sync()
        dt_f = time.perf_counter() - t0
        field = {"points": k, "ms": 1e3 * dt_f, "point_panel_pairs_per_s": k * prob.n / dt_f,
                 "columns": prob.m + prob.nbeta}
        print(f"# field points: {field}", file=sys.stderr, flush=True)
    hist_mx = ctx.pair_histogram().ravel()     # pairs assembled densely on the measured path
    if args.path == "hmatrix":
        hmat = dict(ctx.hmatrix_stats(), eps=args.aca_eps, eps_s=args.aca_eps_s, kmax=args.aca_kmax,
                    setup_s=hm_setup_s)
        ctx.set_hmatrix(False)
        print(f"# hmatrix {hmat}", file=sys.stderr, flush=True)
    if sur is not None:
        sur["pairs"] = ctx.surrogate_counts()
        ctx.set_surrogate(False)
    print(f"# {args.path}: {args.steps} systems in {dt_mx:.3f} s; stages {st_mx}; gmres {gm}; surrogate {sur}",
          file=sys.stderr, flush=True)
    others = {}
    if args.path != "fp32ir" and args.with_fp32ir:
        dt_ir, _, st_ir = time_path(ctx, prob, hm_bem.HM_FP32_IR, args.steps, args.warmup)
        print(f"# fp32ir: {args.steps} systems in {dt_ir:.3f} s; stages {st_ir}", file=sys.stderr, flush=True)
        others["fp32_lu_ir"] = {"value": args.steps / dt_ir, "ms_per_step": 1e3 * dt_ir / args.steps}
    plain = None
    if not args.no_plain:
        dt_pl, res_pl, st_pl = time_path(ctx, prob, hm_bem.HM_FP64, args.steps, args.warmup)
        print(f"# fp64: {args.steps} systems in {dt_pl:.3f} s; stages {st_pl}", file=sys.stderr, flush=True)
        nw, ew, ew3, inbar, cnt = {}, {}, {}, {}, {}
        for a, b in zip(res_mx, res_pl):
            for k in ("amas", "bdmp", "exfc"):
                mx = np.max(np.abs(b[k]))
                d = np.abs(a[k] - b[k])
                nw[k] = max(nw.get(k, 0.0), float(np.max(d) / mx))
                # entries below 1e-6 of the largest are compared absolutely, against 1e-6 * max
                den = np.maximum(np.abs(b[k]), 1e-6 * mx)
                ew[k] = max(ew.get(k, 0.0), float(np.max(d / den)))
                # entries of at least 1e-3 of the largest (the cross couplings that are
                # not near a zero crossing)
                big = np.abs(b[k]) >= 1e-3 * mx
                ew3[k] = max(ew3.get(k, 0.0), float(np.max(d[big] / np.abs(b[k][big]))))
                inbar[k] = inbar.get(k, 0) + int(np.count_nonzero(d <= 1e-4 * den))
                cnt[k] = cnt.get(k, 0) + d.size
        plain = {"value": args.steps / dt_pl, "ms_per_step": 1e3 * dt_pl / args.steps,
                 "speedup_mixed_vs_plain": (args.steps / dt_mx) / (args.steps / dt_pl),
                 "rel_err_mixed_vs_fp64_normwise": nw, "rel_err_mixed_vs_fp64_entrywise": ew,
                 "rel_err_mixed_vs_fp64_entries_above_1e-3_max": ew3,
                 "fraction_of_entries_within_1e-4": {k: inbar[k] / cnt[k] for k in cnt},
                 "max_rel_err_mixed_vs_fp64": max(nw.values()),
                 "stages_ms": {k: v[0] for k, v in st_pl.items()}}

    # dominant kernel of the measured path: the assembly launch (per regime for
    # finite depth) with most time
    fpp, deep = flop_model(args.workload)
    hist = hist_mx
    st_rf = st_mx
    asm = {k: v for k, v in st_rf.items() if k in REGIME_STAGE and v[1] > 0}
    if asm:
        dom = max(asm, key=lambda k: asm[k][0])
        dom_ms = asm[dom][0] / asm[dom][1]
        r = REGIME_STAGE[dom]
        keys = range(128 * r, 128 * (r + 1))
        flops = sum(float(hist[k]) * fpp.get(k, fpp.get(128 * r, 0.0)) for k in keys)
        units = int(sum(hist[k] for k in keys))
    else:
        dom, dom_ms = "asm_deep", st_rf["assembly"][0] / max(st_rf["assembly"][1], 1)
        units = (prob.n + prob.ni) * prob.n
        flops = units * fpp[0]
    achieved = flops / (dom_ms * 1e-3) / 1e12
    traffic = MEASURED_TRAFFIC.get((args.workload, args.path, dom))
    roofline = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "traffic_gbs": traffic / (dom_ms * 1e-3) / 1e9 if traffic else None,
                "avg_ms": dom_ms, "pairs_per_launch": units, "algorithmic_flops_per_launch": flops,
                "algorithmic_bytes_per_launch": 36 * units,
                "note": ("FP64 VALU-bound Green-function assembly over the classified pairs of one regime; "
                         "peak = MI355X FP64 vector peak (= FP64 matrix peak); algorithmic flops = the reference's "
                         "per-pair work counted by the oracle (tests/golden/flop_model.json) for every pair the "
                         "launch assembles; on the H-matrix path the PIDCOCK kernel assembles each pair with its "
                         "mirror and reads the frequency-independent halves of its terms from a per-geometry "
                         "cache, executing 41 % of the FP64 instructions of a pair-by-pair evaluation; traffic = "
                         "measured HBM bytes per launch, algorithmic bytes = 32 B of A and S + 4 B pair index "
                         "per pair")}

    line = {"metric": METRIC, "value": args.steps / dt_mx, "unit": "systems/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt_mx / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32+bf16",
            "data": "synthetic (deterministic meshes of the named workload)",
            "config": {"workload": prob.name, "panels": prob.n, "waterplane_panels": prob.ni,
                       "modes": prob.m, "headings": prob.nbeta, "frequencies": len(prob.omegas),
                       "path": PATH_DESC[args.path], "parallelism": "1 GPU"},
            "stages_ms": {k: v[0] for k, v in st_mx.items()},
            "plain_fp64": plain, "roofline": roofline}
    if gm is not None:
        line["gmres"] = gm
    if hmat is not None:
        line["hmatrix"] = hmat
    if field is not None:
        line["field_points"] = field
    if sur is not None:
        line["surrogate"] = sur
    if others:
        line["other_paths"] = others
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(prob, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
