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
KERNEL_OF_STAGE = {"asm_linton": "k_assemble_list<0>", "asm_pidcock": "k_assemble_sym<1,16,7,true>",
                   "asm_eigene": "k_assemble_sym<2,8,7,false>", "asm_eigen": "k_assemble_list<3>",
                   "aca_entries": "k_aca_eval_row/col<8,false,true>"}
ACCURACY_BAR = 1e-4

# HBM bytes per launch of the candidate kernels, measured on one MI355X by two
# counter passes of `bench.py --steps 1 --assemble-only` (rocprofv3 --pmc
# FETCH_SIZE, then --pmc WRITE_SIZE, kernel filter
# k_assemble_sym|k_aca_eval|k_assemble_list), scaled by the readings of a 1 GiB
# device-to-device copy (fetch x2, write x1). The H-matrix path's PIDCOCK
# kernel assembles the one distinct diagonal block's pairs with their mirrors
# from the cached frequency-independent term halves (1.9 GB read per launch);
# it runs at 5 waves per SIMD, and register-spill traffic (scratch that misses
# the L2) is most of the rest.
MEASURED_TRAFFIC = {
    ("c5", "hmatrix", "asm_pidcock"): 2685335.9375 * 1024 * 2 + 2737365.875 * 1024,
    ("c5", "hmatrix", "asm_eigene"): 1121673.8125 * 1024 * 2 + 2776009.15625 * 1024,
    # per launch, mean of the row and column launches (54 of each)
    ("c5", "hmatrix", "aca_entries"): (252819.0 + 179202.3125) / 108 * 1024 * 2 + (23670.625 + 31830.25) / 108 * 1024,
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
    "hmatrix": ("H-matrix: dense FP64 diagonal block per body (each pair with its mirror, bit for bit the plain "
                "path's blocks; bodies whose local hull panels are equal share one block, assembled, factored and "
                "inverted once per frequency), off-diagonal blocks by adaptive cross approximation (tolerance 1e-6 for A, 1e-7 "
                "for S) whose entries come straight from the FP64 Green function with directly summed evanescent "
                "modes, right-hand sides S W from the H-matrix of S; restarted GMRES(50), tol 1e-6, with FP32 "
                "MFMA H-matrix products, FP32 block-diagonal preconditioner (explicit inverses) and FP64 H-matrix "
                "residuals; several frequencies in flight (clones of the context, one host thread each)"),
}


def run_steps(ctx, omegas, k0, steps, precision, ranks=None, progress=None):
    out = []
    for s in range(steps):
        w = float(omegas[(k0 + s) % len(omegas)])
        out.append(ctx.solve_frequency(w, precision))
        if progress:
            print(f"# {progress}: system {s + 1}/{steps} (omega {w:.4f})", file=sys.stderr, flush=True)
        if ranks is not None:
            # mean rank over all cross-approximated blocks, and over the processes
            # that ran (the others copy an equal block's factors)
            ranks.append((ctx.hmatrix_stats()["mean_rank"], ctx.block_sharing(with_rank=True)[2]))
    return out


def time_path(ctx, prob, precision, steps, warmup, stage_times=True, ranks=None, progress=None):
    run_steps(ctx, prob.omegas, 0, warmup, precision, progress=progress and progress + " warm-up")
    sync()
    ctx.set_timing(stage_times)
    t0 = time.perf_counter()
    res = run_steps(ctx, prob.omegas, warmup, steps, precision, ranks, progress=progress)
    sync()
    dt = time.perf_counter() - t0
    stages = ctx.timing()
    ctx.set_timing(False)
    return dt, res, stages


def time_path_inflight(ctx, prob, precision, steps, warmup, workers, stage_times=True):
    """the same frequencies through hm_run_frequencies with `workers`
    frequencies in flight (clones of the context made by the warm-up call);
    stage times are those of the frequencies the context itself took"""
    om = prob.omegas
    ctx.set_workers(workers)
    ctx.run_frequencies([float(om[k % len(om)]) for k in range(warmup)], precision)
    sync()
    ctx.set_timing(stage_times)
    w = [float(om[(warmup + k) % len(om)]) for k in range(steps)]
    t0 = time.perf_counter()
    a, b, f = ctx.run_frequencies(w, precision)
    sync()
    dt = time.perf_counter() - t0
    stages = ctx.timing()
    ctx.set_timing(False)
    ctx.set_workers(1)
    res = [dict(amas=a[k], bdmp=b[k], exfc=f[k]) for k in range(steps)]
    return dt, res, stages


def host_cores():
    """physical cores among the CPUs this process may run on (lscpu's core ids,
    one thread per core); the affinity count where lscpu is unavailable"""
    import subprocess
    cpus = sorted(os.sched_getaffinity(0))
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        core = {}
        for ln in out.split("\n"):
            if ln and not ln.startswith("#"):
                c, k, so = ln.split(",")[:3]
                core[int(c)] = (so, k)
        phys = {core[c] for c in cpus if c in core}
        return max(1, len(phys)) if phys else len(cpus)
    except (OSError, ValueError, subprocess.CalledProcessError):
        return len(cpus)


def offdiag_flops_per_entry(prob, fpp, step=7):
    """average counted flops (reference algorithm, tests/golden/flop_model.json)
    of the entries between different bodies, over every step-th pair"""
    bs = prob.body_start
    tot, cnt = 0.0, 0
    h = prob.h
    for b1 in range(prob.nbody):
        xi = prob.xyz_p[bs[b1]:bs[b1 + 1]:step]
        for b2 in range(prob.nbody):
            if b1 == b2:
                continue
            xj = prob.xyz_p[bs[b2]:bs[b2 + 1]:step]
            rr = np.hypot(xi[:, None, 0] - xj[None, :, 0], xi[:, None, 1] - xj[None, :, 1]).ravel()
            rh = rr / h
            reg = np.where(rh < 5e-4, 0, np.where(rh < 0.05, 1, np.where(rh < 0.5, 2, 3)))
            nt = np.where(reg == 2, (np.float32(-88.89) * rh + np.float32(54.45)).astype(int),
                          np.where(reg == 1, (np.float32(-1010.10) * rh + np.float32(100.50)).astype(int), 0))
            keys = reg * 128 + nt
            u, c = np.unique(keys, return_counts=True)
            for k, cc in zip(u, c):
                tot += cc * fpp.get(int(k), fpp.get(128 * int(k // 128), 0.0))
            cnt += len(keys)
    return tot / max(cnt, 1)


def cpu_baseline(prob, seconds=20.0, threads=None):
    """The reference's own CPU kernels timed on this host's physical cores: the
    body of its CALGREEN table-fill loop (SGLINTBD + FINGREEN3D/INFGREEN3D,
    compiled from the reference sources into oracle/_ref with OpenMP,
    CalGreenFunc.f90:58-115, CalGreenFuncMulti.f90:123-180) over a random
    sample of the workload's panel pairs, and the rate of a LAPACK ZGETRF/ZGETRS
    (numpy's, at a stated N <= 6144) applied to the workload's 8N^3/3 LU flops.
    GRN_KIND (WavDynMods.f90:47) only sets the precision the table is stored in;
    the loop body computes in FP64 either way and is timed without the store.
    Returns systems/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as om
    threads = threads or host_cores()
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

    t = run(20000)
    npairs = int(min(1e8, max(20000, 20000 * 0.75 * seconds / max(t, 1e-6))))
    t = run(npairs)
    pair_rate = npairs / t
    m = min(n, 6144)
    a = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    b = rng.normal(size=(m, prob.m + prob.nbeta)) + 1j * rng.normal(size=(m, prob.m + prob.nbeta))
    np.linalg.solve(a[:256, :256], b[:256])
    t0 = time.perf_counter()
    np.linalg.solve(a, b)
    t_small = time.perf_counter() - t0
    lu_rate = (8.0 * m ** 3 / 3.0 + 8.0 * m * m * (prob.m + prob.nbeta)) / t_small
    t_lu = (8.0 * n ** 3 / 3.0 + 8.0 * n * n * (prob.m + prob.nbeta)) / lu_rate
    t_sys = nt * n / pair_rate + t_lu
    return {"value": 1.0 / t_sys, "unit": "systems/s", "cores": threads, "kind": "reference",
            "sample": (f"{npairs} random panel pairs of the workload at omega={w:.3f} through the reference's "
                       f"CALGREEN loop body (oracle/_ref, OpenMP, {threads} threads = the physical cores this "
                       f"process may use) in {t:.1f} s = {pair_rate:.3g} pairs/s, x {nt * n} pairs per system "
                       f"= {nt * n / pair_rate:.1f} s; LU + solves at the rate of LAPACK zgesv (numpy) measured "
                       f"at N={m}: {lu_rate / 1e9:.1f} GFLOP/s, for 8N^3/3 + 8N^2(M+NB) at N={n}: {t_lu:.1f} s; "
                       f"{t_sys:.1f} s per system. GRN_KIND sets only the table's storage precision; the loop "
                       f"body is FP64 either way")}


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
    tuning = b.tuning()
    print(f"# batch prepared in {time.perf_counter() - t0:.1f} s: {info}; tuning {tuning}", file=sys.stderr, flush=True)
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
            c.setup(p, frames=False)         # the batch takes global coordinates
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
                                "solves, one HIP graph replay per frequency, autotuned series variant and GEMM "
                                "tile K chunk; "
                                + ("FP64 LU" if prec == hm_bem.HM_FP64 else "FP32 LU + FP64 refinement")),
                       "parallelism": "1 GPU"},
            "batch": info, "tuning": tuning, "sequential_fp64": plain}
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
    ap.add_argument("--aca-eps", type=float, default=1e-6, help="cross-approximation tolerance of A")
    ap.add_argument("--aca-eps-s", type=float, default=1e-7, help="cross-approximation tolerance of S")
    ap.add_argument("--aca-kmax", type=int, default=50, help="largest rank of a cross-approximated block")
    ap.add_argument("--surrogate-tol", type=float, default=None,
                    help="cell error bound up to which the learned surrogate replaces the plain series (0: off; "
                         "default: off for C5, whose near-trapped-mode frequencies amplify any change of the "
                         "diagonal blocks' entries above 1e-4 in A/B/F; 1e-3 for C4)")
    ap.add_argument("--surrogate-samples", type=int, default=16_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=60)
    ap.add_argument("--gmres-tol", type=float, default=1e-6)
    ap.add_argument("--gmres-precond", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="block preconditioner: 1 triangular solves, 2 explicit FP32 inverses, 0 none, "
                         "-1 library default")
    ap.add_argument("--series-variant", type=int, default=1, choices=(0, 1, 2),
                    help="epsilon-accelerated regimes: 1 lane groups of 8/16 (default), 2 of 4/8, 0 one pair per lane")
    ap.add_argument("--lu-panel", choices=("coop", "column"), default="coop",
                    help="dense LU panels: one cooperative launch each (default) or per-column launches (for "
                         "profilers that do not handle cooperative launches)")
    ap.add_argument("--workers", type=int, default=None,
                    help="frequencies in flight on the measured path (hm_set_workers: clones of the context, one "
                         "host thread each); default 4 on the H-matrix path, 1 otherwise")
    ap.add_argument("--no-block-sharing", action="store_true",
                    help="H-matrix path: assemble, factor and invert every body's diagonal block, also where bodies "
                         "are equal (the blocks are the same bit for bit either way)")
    ap.add_argument("--no-symmetries", action="store_true",
                    help="H-matrix path: share cross approximations between equal bodies at equal offsets only, not "
                         "across offsets that a symmetry of the body relates")
    ap.add_argument("--no-extras", action="store_true",
                    help="H-matrix path: skip the passes after the timed region (the cross approximation alone, "
                         "the path without block sharing), e.g. for profiler runs")
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
    ctx.set_lu_panel(args.lu_panel == "coop")
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
    if args.surrogate_tol is None:
        args.surrogate_tol = 1e-3 if args.workload == "c4" else 0.0
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
        ctx.set_block_sharing(not args.no_block_sharing, symmetries=not args.no_symmetries)
        ctx.set_hmatrix(True, eps=args.aca_eps, eps_s=args.aca_eps_s, kmax=args.aca_kmax)
        sync()
        hm_setup_s = time.perf_counter() - t0
        print(f"# hmatrix set-up (diagonal-block classification, PIDCOCK term cache) {hm_setup_s:.2f} s",
              file=sys.stderr, flush=True)
    ranks_mx = []
    if args.workers is None:
        args.workers = 4 if args.path == "hmatrix" else 1
    if args.workers > 1:
        dt_mx, res_mx, st_mx = time_path_inflight(ctx, prob, prec, args.steps, args.warmup, args.workers)
    else:
        dt_mx, res_mx, st_mx = time_path(ctx, prob, prec, args.steps, args.warmup,
                                         ranks=ranks_mx if args.path == "hmatrix" else None)
    # systems the context itself solved in the timed region (its stage times cover those)
    nsys_main = st_mx["forces"][1] if st_mx.get("forces", (0, 0))[1] > 0 else args.steps
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
    aca_alone = None
    if args.path == "hmatrix" and args.steps > 0 and not args.no_extras:
        # the cross approximation's entry kernels without the diagonal blocks
        # beside them (on the context's stream after them), for the timed
        # region's first two frequencies once more, outside it: their launch
        # durations are their own, not shared with the other stream's kernels
        ctx.set_hmatrix_overlap(False)
        ranks_alone = []
        _, _, st_alone = time_path(ctx, prob, prec, min(2, args.steps), args.warmup, ranks=ranks_alone)
        ctx.set_hmatrix_overlap(True)
        aca_alone = {"systems": min(2, args.steps), "stages_ms": {k: v for k, v in st_alone.items() if v[1]},
                     "ranks": [r[0] for r in ranks_alone], "ranks_of_processes_run": [r[1] for r in ranks_alone],
                     "processes_run": ctx.block_sharing()[1]}
        print(f"# cross approximation alone: {aca_alone}", file=sys.stderr, flush=True)
    no_share = None
    if args.path == "hmatrix" and not args.no_block_sharing and args.steps > 0 and not args.no_extras:
        # the same path with every body's diagonal block assembled, factored and
        # inverted and every cross approximation run (outside the timed region;
        # the outputs are the same bit for bit with sharing at equal offsets only,
        # within the cross approximation's tolerance with the body symmetries)
        ctx.set_block_sharing(False)
        k = min(3, args.steps)
        dt_ns, res_ns, _ = time_path(ctx, prob, prec, k, args.warmup, stage_times=False)
        same = all(np.array_equal(a[q], b[q]) for a, b in zip(res_ns, res_mx[:k]) for q in ("amas", "bdmp", "exfc"))
        dev = max(float(np.max(np.abs(a[q] - b[q])) / max(float(np.max(np.abs(a[q]))), 1e-300))
                  for a, b in zip(res_ns, res_mx[:k]) for q in ("amas", "bdmp", "exfc"))
        no_share = {"systems": k, "value": k / dt_ns, "ms_per_step": 1e3 * dt_ns / k, "outputs_bitwise_equal": same,
                    "largest_normwise_deviation": dev}
        ctx.set_block_sharing(True, symmetries=not args.no_symmetries)
        print(f"# without block sharing: {no_share}", file=sys.stderr, flush=True)
    if args.path == "hmatrix":
        hmat = dict(ctx.hmatrix_stats(), eps=args.aca_eps, eps_s=args.aca_eps_s, kmax=args.aca_kmax,
                    setup_s=hm_setup_s, distinct_bodies=int(len(set(ctx.body_classes().tolist()))),
                    diag_blocks_assembled=ctx.block_sharing()[0], aca_processes_run=ctx.block_sharing()[1],
                    block_sharing=not args.no_block_sharing,
                    body_symmetries_used=(not args.no_block_sharing and not args.no_symmetries
                                          and int(ctx.body_symmetries().sum())),
                    without_block_sharing=no_share)
        ctx.set_hmatrix(False)
        print(f"# hmatrix {hmat}", file=sys.stderr, flush=True)
    if sur is not None:
        sur["pairs"] = ctx.surrogate_counts()
        ctx.set_surrogate(False)
    print(f"# {args.path}: {args.steps} systems in {dt_mx:.3f} s ({args.workers} in flight); stages {st_mx}; gmres {gm}; "
          f"surrogate {sur}",
          file=sys.stderr, flush=True)
    others = {}
    if args.path != "fp32ir" and args.with_fp32ir:
        dt_ir, _, st_ir = time_path(ctx, prob, hm_bem.HM_FP32_IR, args.steps, args.warmup)
        print(f"# fp32ir: {args.steps} systems in {dt_ir:.3f} s; stages {st_ir}", file=sys.stderr, flush=True)
        others["fp32_lu_ir"] = {"value": args.steps / dt_ir, "ms_per_step": 1e3 * dt_ir / args.steps}
    plain = None
    if not args.no_plain:
        # one progress line per system (5 s each at C5)
        dt_pl, res_pl, st_pl = time_path(ctx, prob, hm_bem.HM_FP64, args.steps, args.warmup, progress="fp64")
        print(f"# fp64: {args.steps} systems in {dt_pl:.3f} s; stages {st_pl}", file=sys.stderr, flush=True)
        nw, ew, ew3, inbar, cnt = {}, {}, {}, {}, {}
        per_freq = []
        for s_, (a, b) in enumerate(zip(res_mx, res_pl)):
            w = float(prob.omegas[(args.warmup + s_) % len(prob.omegas)])
            per_freq.append([round(w, 6)] + [float(np.max(np.abs(a[k] - b[k])) / np.max(np.abs(b[k])))
                                              for k in ("amas", "bdmp", "exfc")])
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
                 "rel_err_mixed_vs_fp64_normwise": nw,
                 "rel_err_per_frequency": {"columns": ["omega", "amas", "bdmp", "exfc"], "rows": per_freq},
                 "rel_err_mixed_vs_fp64_entrywise": ew,
                 "rel_err_mixed_vs_fp64_entries_above_1e-3_max": ew3,
                 "fraction_of_entries_within_1e-4": {k: inbar[k] / cnt[k] for k in cnt},
                 "max_rel_err_mixed_vs_fp64": max(nw.values()),
                 "stages_ms": {k: v[0] for k, v in st_pl.items()}}

    # dominant kernel of the measured path: the kernel with the most time over
    # the timed region, among the diagonal-block (or dense) regime launches and
    # the cross approximation's entry launches
    fpp, deep = flop_model(args.workload)
    hist = hist_mx
    st_rf = st_mx
    # per system: the regime launches' own times (one launch each) and the
    # cross approximation's entry launches measured alone
    cand = {k: (v[0] / nsys_main, v[1] / nsys_main) for k, v in st_rf.items()
            if k in REGIME_STAGE and v[1] > 0}
    if aca_alone is not None and "aca_entries" in aca_alone["stages_ms"]:
        ms, cnt = aca_alone["stages_ms"]["aca_entries"]
        cand["aca_entries"] = (ms / aca_alone["systems"], cnt / aca_alone["systems"])
    if cand:
        dom = max(cand, key=lambda k: cand[k][0])
        dom_ms = cand[dom][0] / cand[dom][1]
        if dom == "aca_entries":
            # entries evaluated: one row and one column of m entries per rank step
            # of each of the P processes (A and S of every ordered pair of bodies)
            m = prob.n // prob.nbody
            procs = aca_alone["processes_run"]
            entries = sum(2.0 * m * procs * r for r in aca_alone["ranks_of_processes_run"]) / aca_alone["systems"]
            units = entries / cand[dom][1]
            per_entry = offdiag_flops_per_entry(prob, fpp)
            flops = units * per_entry
            abytes = 16.0 * units
            what = (f"cross-approximation entry kernels (residual rows and columns of all open A and S "
                    f"processes; {procs} processes run, the others copy an equal block's factors), measured alone "
                    f"after the timed region: {cand[dom][1]:.0f} launches and "
                    f"{entries:.4g} entries per system (2 m P x the mean rank), {per_entry:.0f} counted flop per entry (the "
                    f"reference's FINGREEN3D + SGLINTBD work for the off-diagonal pairs of this geometry); the "
                    f"kernels sum the evanescent modes directly to a bounded tail instead of the epsilon tables")
        else:
            r = REGIME_STAGE[dom]
            keys = range(128 * r, 128 * (r + 1))
            flops = sum(float(hist[k]) * fpp.get(k, fpp.get(128 * r, 0.0)) for k in keys)
            units = int(sum(hist[k] for k in keys))
            abytes = 36.0 * units
            what = ("Green-function assembly over the classified pairs of one regime (on the H-matrix path the "
                    "diagonal blocks, each stored pair assembled with its mirror); algorithmic flops = the "
                    "reference's per-pair work counted by the oracle for every pair the launch assembles; "
                    "algorithmic bytes = 32 B of A and S + 4 B pair index per pair")
    else:
        dom, dom_ms = "asm_deep", st_rf["assembly"][0] / max(st_rf["assembly"][1], 1)
        units = (prob.n + prob.ni) * prob.n
        flops = units * fpp[0]
        abytes = 36.0 * units
        what = "infinite-depth pair grid"
    achieved = flops / (dom_ms * 1e-3) / 1e12
    traffic = MEASURED_TRAFFIC.get((args.workload, args.path, dom))
    roofline = {"bound": "valu", "kernel": KERNEL_OF_STAGE.get(dom, dom), "stage": dom, "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": traffic, "traffic_gbs": traffic / (dom_ms * 1e-3) / 1e9 if traffic else None,
                "avg_ms": dom_ms, "launches_per_system": cand[dom][1] if cand else 1, "units_per_launch": units,
                "kernels_ms_per_system": {k: round(v[0], 2) for k, v in cand.items()},
                "algorithmic_flops_per_launch": flops, "algorithmic_bytes_per_launch": abytes,
                "note": ("FP64 vector (VALU) bound; peak = MI355X FP64 vector peak, equal to its FP64 matrix peak; "
                         "counted flops: each addition, multiplication, division and square root 1, each "
                         "transcendental 20 (oracle operation counter, tests/golden/flop_model.json); avg_ms = "
                         "HIP events around each launch on its own stream. " + what)}

    acc_ok = None
    if plain is not None:
        acc_ok = bool(plain["max_rel_err_mixed_vs_fp64"] <= ACCURACY_BAR)
    value = args.steps / dt_mx if acc_ok is not False else None
    line = {"metric": METRIC, "value": value, "unit": "systems/s", "n_gpus": 1,
            "accuracy_ok": acc_ok, "accuracy_bar": ACCURACY_BAR,
            "accuracy_note": ("largest normwise deviation (max |mixed - FP64| / max |FP64| per output) over every "
                              "timed frequency; the line's value is null when it exceeds the bar"),
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt_mx / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": ("f64+f32+f16+bf16" if sur is not None else "f64+f32+f16") if args.path == "hmatrix" else
                     ("f64+f32+bf16" if sur is not None else "f64+f32"),
            "data": "synthetic (deterministic meshes of the named workload)",
            "config": {"workload": prob.name, "panels": prob.n, "waterplane_panels": prob.ni,
                       "modes": prob.m, "headings": prob.nbeta, "frequencies": len(prob.omegas),
                       "path": PATH_DESC[args.path], "frequencies_in_flight": args.workers,
                       "parallelism": "1 GPU"},
            "workers": args.workers,
            "stages_ms_per_system": {k: v[0] / nsys_main for k, v in st_mx.items()},
            "stages_note": (f"per system, over the {nsys_main} of the {args.steps} timed systems that the context itself "
                            f"solved (the other {args.steps - nsys_main} ran on its clones, concurrently)"
                            if nsys_main != args.steps else "per system over the timed systems"),
            "plain_fp64": plain, "roofline": roofline}
    if value is None:
        line["value_without_accuracy"] = args.steps / dt_mx
    if aca_alone is not None:
        line["cross_approximation_alone"] = aca_alone
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
