// hm_batch_api.cpp -- batches of independent problems (include/hm_bem.h, hm_batch_*).
//
// The reference solves one problem at a time: its driver loops over the wave
// frequencies and calls the assembly, LU and force routines serially
// (HAMS_Prog.f90:615-724). A batch holds many independent problems (members)
// of different panel counts and depths resident in HBM at once and runs the
// per-frequency chain for all of them together:
//   * assembly: one launch per finite-depth regime over the pairs of
//     every finite-depth member (classified once into one list sorted by regime
//     and series length across members), one over all pairs of the
//     infinite-depth members;
//   * right-hand sides, LU and triangular solves as grouped launches
//     (blockIdx.z = member, each member at its own order);
//   * FP64 LU, or FP32 LU on MFMA with a fixed number of FP64 refinement sweeps
//     inside the chain and more after it where the correction has not yet
//     fallen below 1e-13 of the solution;
//   * the whole chain captured once in a HIP graph and replayed per frequency
//     (the kernels read the members' wave state from device memory, which is
//     rewritten before each replay).
// An autotuner times the assembly launches of one frequency for the series
// variants (the lane-group shapes G x K of the epsilon-accelerated regimes:
// one pair per lane, 16 x 7 / 8 x 7, 8 x 13 / 4 x 13) and the grouped LU for
// the K chunk of its MFMA GEMM tiles (8, 16, 32), and keeps the fastest of
// each before the graph is captured.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hm_assembly.hpp"
#include "hm_bem.h"
#include "hm_gmres.hpp"
#include "hm_linalg.hpp"

namespace {

struct Member {
  int n, nbody, nbeta, ncol, m;
  int ni = 0;                 // interior waterplane field points (irregular-frequency removal)
  double h;
  std::vector<int> body_start;
  std::vector<double> verts, fp, ds, pnsz, dxyz, beta;   // fp: the n hull centroids, then the ni waterplane points
  std::vector<int> ncn;
  int nt() const { return n + ni; }
};

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

}  // namespace

struct hm_batch {
  hipStream_t stream = nullptr;
  std::string err;
  double rho = 1025.0, g = 9.81, g_disp = 9.81, amp = 1.0, xw = 0.0, yw = 0.0;
  std::vector<Member> mem;
  bool prepared = false;
  int precision = HM_FP64, nrefine = 3, use_graph = 1;
  // tuning
  int variant = 1, grid = 0, cus = 256;   // grid: assembly workgroups per frequency
  double tuned_ms = 0.0;
  int gemm_kb = 16;                        // K chunk of the grouped GEMM tiles (8, 16, 32)
  double tuned_lu_ms[3] = {0.0, 0.0, 0.0}; // the grouped LU at K chunks 8, 16, 32
  // device
  void* pool = nullptr;
  size_t pool_bytes = 0;
  hm::BatchProb* d_prob = nullptr;
  hm::BatchFreq* d_freq = nullptr;
  hm::BatchFreq* h_freq = nullptr;    // pinned
  uint64_t* d_codes = nullptr;
  long long seg_off[4] = {0, 0, 0, 0}, seg_cnt[4] = {0, 0, 0, 0};
  int* d_deep = nullptr;
  long long* d_deep_prefix = nullptr;
  int ndeep = 0;
  long long deep_total = 0;
  hm::GemmDesc<double>* d_rhs = nullptr;   // R = S W
  hm::GemmDesc<double>* d_nrmC = nullptr;  // C = A^T A of the waterplane members
  hm::GemmDesc<double>* d_nrmD = nullptr;  // D = A^T R of the waterplane members
  int nirr = 0, nmax_irr = 0, ncolmax_irr = 0, ntmax = 0;
  hm::GemmDesc<double>* d_res = nullptr;   // Rw = Rw - A X
  hm::LuDesc<double>* d_lu64 = nullptr;
  hm::LuDesc<float>* d_lu32 = nullptr;
  hm::ConvDesc* d_convA = nullptr;         // A -> A32
  hm::ConvDesc* d_convR = nullptr;         // Rw -> X32 (demote)
  hm::ConvDesc* d_convX = nullptr;         // X32 -> X (promote); correction ratio
  hm::CopyDesc* d_copyR = nullptr;         // R -> Rw
  int* d_info = nullptr;                   // info[2*count]: info, counter per member
  double* d_ratio = nullptr;
  std::vector<size_t> off_amas, off_bdmp, off_exfc;
  std::vector<hm::BatchProb> h_prob;       // host copy of the member descriptors
  std::vector<const float*> h_a32;         // the members' FP32 copies of A (HM_GMRES)
  double *d_amas = nullptr, *d_bdmp = nullptr, *d_exfc = nullptr;
  size_t out_amas = 0, out_exfc = 0;
  int nmax = 0, ncolmax = 0, mmax = 0;
  long long max_nn = 0, max_ncol_elems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double last_ratio = 0.0;
  int extra_sweeps = 0;
  // HM_GMRES members: restarted GMRES per member after the captured assembly,
  // FP32 copies of A for the Krylov products, FP64 residuals, no preconditioner
  int gm_restart = 50, gm_maxiter = 500;
  double gm_tol = 1e-6;
  void* d_gws = nullptr;              // one workspace, sized for the largest member
  std::vector<hm::GmresStats> gm_stats;

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return HM_OK;
    (void)hipGetLastError();
    return fail(HM_ERR_HIP - (int)e, "%s: %s", what, hipGetErrorString(e));
  }
  int count() const { return (int)mem.size(); }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    if (pool) (void)hipFree(pool);
    pool = nullptr;
    if (d_gws) (void)hipFree(d_gws);
    d_gws = nullptr;
    if (h_freq) (void)hipHostFree(h_freq);
    h_freq = nullptr;
    prepared = false;
  }
};

#define BCHK(b, call, what)                     \
  do {                                          \
    int _rc = (b)->hip((call), (what));         \
    if (_rc != HM_OK) return _rc;               \
  } while (0)

extern "C" int hm_dispersion(double* wvn, int nrt, double w, double h, double g_disp);

extern "C" int hm_batch_create(hm_batch** out) {
  if (!out) return HM_ERR_ARG - 1;
  *out = nullptr;
  hm_batch* b = new hm_batch();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return HM_ERR_HIP;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    b->cus = prop.multiProcessorCount;
  *out = b;
  return HM_OK;
}

extern "C" int hm_batch_destroy(hm_batch* b) {
  if (!b) return HM_ERR_ARG - 1;
  b->release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HM_OK;
}

extern "C" const char* hm_batch_error_string(const hm_batch* b) { return b ? b->err.c_str() : "null batch"; }

extern "C" int hm_batch_set_constants(hm_batch* b, double rho, double g, double g_disp, double amp, double xw,
                                      double yw) {
  if (!b) return HM_ERR_ARG - 1;
  if (b->prepared) return b->fail(HM_ERR_STATE, "hm_batch_set_constants after hm_batch_prepare");
  if (!(rho > 0.0) || !(g > 0.0) || !(g_disp > 0.0)) return b->fail(HM_ERR_ARG - 2, "rho, g and g_disp must be > 0");
  b->rho = rho; b->g = g; b->g_disp = g_disp; b->amp = amp; b->xw = xw; b->yw = yw;
  return HM_OK;
}

extern "C" int hm_batch_add_problem(hm_batch* b, int nbody, const int* body_start, int ntnd, const double* xyz,
                                    const int* ncn, const int* ncon, const double* xyz_p, const double* ds,
                                    const double* pnsz, const double* dxyz_p, double h, int nbeta,
                                    const double* beta) {
  if (!b) return HM_ERR_ARG - 1;
  if (b->prepared) return b->fail(HM_ERR_STATE, "hm_batch_add_problem after hm_batch_prepare");
  if (nbody < 1) return b->fail(HM_ERR_ARG - 2, "nbody must be >= 1");
  if (!body_start || body_start[0] != 0) return b->fail(HM_ERR_ARG - 3, "body_start must start at 0");
  const int n = body_start[nbody];
  for (int k = 0; k < nbody; ++k)
    if (body_start[k + 1] <= body_start[k]) return b->fail(HM_ERR_ARG - 3, "body %d has no panels", k);
  if (n >= (1 << 20)) return b->fail(HM_ERR_ARG - 3, "n = %d exceeds the 20-bit pair encoding", n);
  if (b->count() >= (1 << 23)) return b->fail(HM_ERR_STATE, "too many members");
  if (ntnd < 3 || !xyz) return b->fail(HM_ERR_ARG - 5, "nodes missing");
  if (!ncn || !ncon || !xyz_p || !ds || !pnsz || !dxyz_p) return b->fail(HM_ERR_ARG - 6, "panel arrays missing");
  if (h == 0.0) return b->fail(HM_ERR_ARG - 12, "depth must be > 0 (finite) or < 0 (infinite)");
  if (nbeta < 0 || (nbeta > 0 && !beta)) return b->fail(HM_ERR_ARG - 13, "bad headings");
  Member m;
  m.n = n; m.nbody = nbody; m.nbeta = nbeta; m.m = 6 * nbody; m.ncol = 6 * nbody + nbeta; m.h = h;
  m.body_start.assign(body_start, body_start + nbody + 1);
  m.verts.assign((size_t)n * 12, 0.0);
  for (int j = 0; j < n; ++j) {
    if (ncn[j] != 3 && ncn[j] != 4) return b->fail(HM_ERR_ARG - 6, "panel %d has %d vertices", j + 1, ncn[j]);
    for (int v = 0; v < ncn[j]; ++v) {
      const int node = ncon[4 * j + v];
      if (node < 1 || node > ntnd) return b->fail(HM_ERR_ARG - 7, "panel %d vertex %d out of range", j + 1, v + 1);
      for (int k = 0; k < 3; ++k) m.verts[12 * (size_t)j + 3 * v + k] = xyz[3 * (size_t)(node - 1) + k];
    }
  }
  m.fp.assign(xyz_p, xyz_p + 3 * (size_t)n);
  m.ds.assign(ds, ds + n);
  m.pnsz.assign(pnsz, pnsz + n);
  m.dxyz.assign(dxyz_p, dxyz_p + 6 * (size_t)n);
  m.ncn.assign(ncn, ncn + n);
  m.beta.assign(beta, beta + nbeta);
  b->mem.push_back(std::move(m));
  return b->count() - 1;
}

extern "C" int hm_batch_set_waterplane(hm_batch* b, int member, int ni, const double* ixyz_p) {
  if (!b) return HM_ERR_ARG - 1;
  if (b->prepared) return b->fail(HM_ERR_STATE, "hm_batch_set_waterplane after hm_batch_prepare");
  if (member < 0 || member >= b->count()) return b->fail(HM_ERR_ARG - 2, "no member %d", member);
  if (ni < 0 || (ni > 0 && !ixyz_p)) return b->fail(HM_ERR_ARG - 3, "bad waterplane points");
  Member& m = b->mem[member];
  if (m.n + ni >= (1 << 20)) return b->fail(HM_ERR_ARG - 3, "n + ni = %d exceeds the 20-bit pair encoding", m.n + ni);
  m.fp.resize(3 * (size_t)m.n);
  if (ni) m.fp.insert(m.fp.end(), ixyz_p, ixyz_p + 3 * (size_t)ni);
  m.ni = ni;
  return HM_OK;
}

namespace {

// the per-frequency chain (assembly .. forces) on b->stream, with the current
// tuning; capturable (no memset, no synchronisation, fixed shapes)
int issue_chain(hm_batch* b) {
  hipStream_t s = b->stream;
  const int cnt = b->count();
  // assembly
  BCHK(b, hm::batch_assemble_all(s, true, b->d_prob, b->d_freq, b->d_deep, b->d_deep_prefix, b->ndeep,
                                 b->deep_total), "batch assembly (deep)");
  for (int r = 0; r < 4; ++r)
    BCHK(b, hm::batch_assemble_regime(s, r, b->d_prob, b->d_freq, b->d_codes + b->seg_off[r], b->seg_cnt[r],
                                      b->variant), "batch assembly");
  // right-hand sides R = S W
  BCHK(b, hm::batch_build_w(s, b->d_prob, b->d_freq, cnt, b->nmax), "batch rhs");
  BCHK(b, hm::cgemm_grouped<double>(s, false, b->d_rhs, cnt, b->ntmax, b->ncolmax, b->gemm_kb), "batch rhs gemm");
  // waterplane members: least squares through the normal equations with plain
  // transpose, C = A^T A and D = A^T R (AssbMatx_irr.f90:53-180)
  if (b->nirr) {
    BCHK(b, hm::cgemm_grouped<double>(s, true, b->d_nrmC, b->nirr, b->nmax_irr, b->nmax_irr, b->gemm_kb),
         "batch A^T A");
    BCHK(b, hm::cgemm_grouped<double>(s, true, b->d_nrmD, b->nirr, b->nmax_irr, b->ncolmax_irr, b->gemm_kb),
         "batch A^T R");
  }
  // factor and solve
  BCHK(b, hm::fill_i32(s, b->d_info, 2LL * cnt, 0), "batch factor");
  if (b->precision == HM_GMRES) {
    // the FP32 copies for the Krylov products; the solves and the forces follow
    // outside the chain (GMRES reads its residual norms on the host)
    BCHK(b, hm::convert_grouped(s, b->d_convA, cnt, b->max_nn, true, false), "batch demote");
    return HM_OK;
  }
  if (b->precision == HM_FP64) {
    BCHK(b, hm::lu_factor_grouped<double>(s, b->d_lu64, cnt, b->nmax, b->gemm_kb), "batch lu");
    BCHK(b, hm::lu_solve_grouped<double>(s, b->d_lu64, cnt, b->nmax, b->ncolmax), "batch solve");
  } else {
    BCHK(b, hm::convert_grouped(s, b->d_convA, cnt, b->max_nn, true, false), "batch demote");
    BCHK(b, hm::lu_factor_grouped<float>(s, b->d_lu32, cnt, b->nmax, b->gemm_kb), "batch lu fp32");
    for (int it = 0; it < b->nrefine; ++it) {
      // residual Rw = R - A X (the first sweep has X = 0: Rw = R)
      BCHK(b, hm::copy_grouped(s, b->d_copyR, cnt, b->max_ncol_elems), "batch residual");
      if (it > 0)
        BCHK(b, hm::cgemm_grouped<double>(s, false, b->d_res, cnt, b->nmax, b->ncolmax, b->gemm_kb), "batch residual");
      BCHK(b, hm::convert_grouped(s, b->d_convR, cnt, b->max_ncol_elems, true, false), "batch demote");
      BCHK(b, hm::lu_solve_grouped<float>(s, b->d_lu32, cnt, b->nmax, b->ncolmax), "batch solve fp32");
      BCHK(b, hm::convert_grouped(s, b->d_convX, cnt, b->max_ncol_elems, false, it == 0), "batch update");
    }
  }
  BCHK(b, hm::batch_forces(s, b->d_prob, b->d_freq, cnt, b->mmax, b->ncolmax), "batch forces");
  return HM_OK;
}

// one more FP32 refinement sweep outside the graph (same launches as a sweep in it)
int extra_sweep(hm_batch* b) {
  hipStream_t s = b->stream;
  const int cnt = b->count();
  BCHK(b, hm::copy_grouped(s, b->d_copyR, cnt, b->max_ncol_elems), "batch residual");
  BCHK(b, hm::cgemm_grouped<double>(s, false, b->d_res, cnt, b->nmax, b->ncolmax, b->gemm_kb), "batch residual");
  BCHK(b, hm::convert_grouped(s, b->d_convR, cnt, b->max_ncol_elems, true, false), "batch demote");
  BCHK(b, hm::lu_solve_grouped<float>(s, b->d_lu32, cnt, b->nmax, b->ncolmax), "batch solve fp32");
  BCHK(b, hm::convert_grouped(s, b->d_convX, cnt, b->max_ncol_elems, false, false), "batch update");
  return HM_OK;
}

int write_freq(hm_batch* b, const double* w1) {
  for (int p = 0; p < b->count(); ++p) {
    const Member& m = b->mem[p];
    hm::BatchFreq& f = b->h_freq[p];
    if (!(w1[p] > 0.0)) return b->fail(HM_ERR_ARG - 2, "w1[%d] must be > 0", p);
    f.w1 = w1[p];
    f.v = w1[p] * w1[p] / b->g;
    f.h = m.h;
    f.rho = b->rho; f.g = b->g; f.amp = b->amp; f.xw = b->xw; f.yw = b->yw;
    if (m.h > 0.0) {
      hm_dispersion(f.wvn, hm::kNk, w1[p], m.h, b->g_disp);
      f.wk = f.wvn[0];
    } else {
      f.wk = f.v;
      for (int k = 0; k < hm::kNk; ++k) f.wvn[k] = 0.0;
    }
  }
  BCHK(b, hipMemcpyAsync(b->d_freq, b->h_freq, sizeof(hm::BatchFreq) * b->count(), hipMemcpyHostToDevice, b->stream),
       "batch frequency");
  return HM_OK;
}

// time the assembly launches of one frequency for the current tuning
int time_assembly(hm_batch* b, double* ms) {
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0), "tune");
  BCHK(b, hipEventCreate(&e1), "tune");
  int rc = HM_OK;
  float best = 1e30f;
  for (int rep = 0; rep < 2 && rc == HM_OK; ++rep) {
    (void)hipEventRecord(e0, b->stream);
    rc = b->hip(hm::batch_assemble_all(b->stream, true, b->d_prob, b->d_freq, b->d_deep, b->d_deep_prefix, b->ndeep,
                                       b->deep_total), "tune");
    for (int r = 0; r < 4 && rc == HM_OK; ++r)
      rc = b->hip(hm::batch_assemble_regime(b->stream, r, b->d_prob, b->d_freq, b->d_codes + b->seg_off[r],
                                            b->seg_cnt[r], b->variant), "tune");
    (void)hipEventRecord(e1, b->stream);
    if (rc == HM_OK) rc = b->hip(hipEventSynchronize(e1), "tune");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    best = t < best ? t : best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return rc;
}

// time the grouped LU of the members' systems of one frequency for each K
// chunk of the GEMM tiles and keep the fastest. The systems are assembled
// first (with the normal equations of the waterplane members); the FP64 LU
// factors them in place, so each FP64 trial assembles again, while the FP32
// trials factor fresh FP32 copies.
int tune_lu(hm_batch* b) {
  hipStream_t s = b->stream;
  const int cnt = b->count();
  const int kbs[3] = {8, 16, 32};
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0), "tune");
  BCHK(b, hipEventCreate(&e1), "tune");
  int rc = HM_OK;
  double best = 1e30;
  int bk = 16;
  auto assemble = [&]() -> int {
    BCHK(b, hm::batch_assemble_all(s, true, b->d_prob, b->d_freq, b->d_deep, b->d_deep_prefix, b->ndeep,
                                   b->deep_total), "tune");
    for (int r = 0; r < 4; ++r)
      BCHK(b, hm::batch_assemble_regime(s, r, b->d_prob, b->d_freq, b->d_codes + b->seg_off[r], b->seg_cnt[r],
                                        b->variant), "tune");
    if (b->nirr) BCHK(b, hm::cgemm_grouped<double>(s, true, b->d_nrmC, b->nirr, b->nmax_irr, b->nmax_irr), "tune");
    return HM_OK;
  };
  if (b->precision == HM_FP32_IR) rc = assemble();
  for (int t = 0; t < 3 && rc == HM_OK; ++t) {
    if (b->precision == HM_FP64) rc = assemble();
    else rc = b->hip(hm::convert_grouped(s, b->d_convA, cnt, b->max_nn, true, false), "tune");
    if (rc == HM_OK) rc = b->hip(hm::fill_i32(s, b->d_info, 2LL * cnt, 0), "tune");
    if (rc != HM_OK) break;
    (void)hipEventRecord(e0, s);
    rc = b->hip(b->precision == HM_FP64 ? hm::lu_factor_grouped<double>(s, b->d_lu64, cnt, b->nmax, kbs[t])
                                        : hm::lu_factor_grouped<float>(s, b->d_lu32, cnt, b->nmax, kbs[t]), "tune");
    (void)hipEventRecord(e1, s);
    if (rc == HM_OK) rc = b->hip(hipEventSynchronize(e1), "tune");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    b->tuned_lu_ms[t] = ms;
    if (ms < best) { best = ms; bk = kbs[t]; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == HM_OK) b->gemm_kb = bk;
  return rc;
}

}  // namespace

extern "C" int hm_batch_tuning(hm_batch* b, int* series_variant, int* gemm_kchunk, double* assembly_ms,
                               double* lu_ms) {
  if (!b) return HM_ERR_ARG - 1;
  if (series_variant) *series_variant = b->variant;
  if (gemm_kchunk) *gemm_kchunk = b->gemm_kb;
  if (assembly_ms) *assembly_ms = b->tuned_ms;
  if (lu_ms)
    for (int t = 0; t < 3; ++t) lu_ms[t] = b->tuned_lu_ms[t];
  return HM_OK;
}

extern "C" int hm_batch_prepare(hm_batch* b, int precision, int autotune, int use_graph, int nrefine) {
  if (!b) return HM_ERR_ARG - 1;
  if (b->count() == 0) return b->fail(HM_ERR_STATE, "hm_batch_prepare with no problems");
  if (precision != HM_FP64 && precision != HM_FP32_IR && precision != HM_GMRES)
    return b->fail(HM_ERR_ARG - 2, "precision must be HM_FP64, HM_FP32_IR or HM_GMRES");
  if (precision == HM_GMRES)
    for (int p = 0; p < b->count(); ++p)
      if (b->mem[p].ni)
        return b->fail(HM_ERR_ARG - 2, "member %d has waterplane points: HM_GMRES members solve the collocation "
                                       "system only (use HM_FP64 or HM_FP32_IR)", p);
  if (nrefine < 1 || nrefine > 20) return b->fail(HM_ERR_ARG - 5, "nrefine must be in 1..20");
  b->release();
  b->precision = precision;
  b->use_graph = use_graph != 0;
  b->nrefine = nrefine;
  const int cnt = b->count();
  const bool ir = precision == HM_FP32_IR;
  const bool gm = precision == HM_GMRES;
  // ---- layout of the pool
  struct Off { size_t body, fp, verts, ncn, ds, pnsz, dxyz, beta, A, S, W, R, Rw, A32, X32, C, D, ipiv, amas, bdmp, exfc; };
  std::vector<Off> off(cnt);
  size_t tot = 0;
  auto take = [&](size_t bytes) { size_t o = tot; tot += align_up(bytes ? bytes : 8); return o; };
  b->nmax = b->ncolmax = b->mmax = b->ntmax = 0;
  b->nirr = b->nmax_irr = b->ncolmax_irr = 0;
  b->max_nn = b->max_ncol_elems = 0;
  long long finite_pairs = 0;
  for (int p = 0; p < cnt; ++p) {
    const Member& m = b->mem[p];
    const size_t n = m.n, nt = m.nt(), nc = m.ncol, mm = m.m;
    Off& o = off[p];
    o.body = take(sizeof(int) * (m.nbody + 1));
    o.fp = take(24 * nt); o.verts = take(96 * n); o.ncn = take(4 * n); o.ds = take(8 * n); o.pnsz = take(8 * n);
    o.dxyz = take(48 * n); o.beta = take(8 * (size_t)m.nbeta);
    // A, S: nt x n; W, R, Rw: nt x ncol (leading dimension nt for all of them)
    o.A = take(16 * nt * n); o.S = take(16 * nt * n); o.W = take(16 * nt * nc); o.R = take(16 * nt * nc);
    o.Rw = ir ? take(16 * nt * nc) : 0;
    o.A32 = ir || gm ? take(8 * n * n) : 0;
    o.X32 = ir ? take(8 * n * nc) : 0;
    // waterplane members: C = A^T A (n x n), D = A^T R (n x ncol, leading dimension nt)
    o.C = m.ni ? take(16 * n * n) : 0;
    o.D = m.ni ? take(16 * nt * nc) : 0;
    o.ipiv = take(4 * n);
    b->ntmax = m.nt() > b->ntmax ? m.nt() : b->ntmax;
    if (m.ni) {
      ++b->nirr;
      b->nmax_irr = m.n > b->nmax_irr ? m.n : b->nmax_irr;
      b->ncolmax_irr = m.ncol > b->ncolmax_irr ? m.ncol : b->ncolmax_irr;
    }
    o.amas = take(8 * mm * mm); o.bdmp = take(8 * mm * mm); o.exfc = take(16 * mm * (size_t)m.nbeta);
    b->nmax = m.n > b->nmax ? m.n : b->nmax;
    b->ncolmax = m.ncol > b->ncolmax ? m.ncol : b->ncolmax;
    b->mmax = m.m > b->mmax ? m.m : b->mmax;
    b->max_nn = (long long)n * n > b->max_nn ? (long long)n * n : b->max_nn;
    b->max_ncol_elems = (long long)n * nc > b->max_ncol_elems ? (long long)n * nc : b->max_ncol_elems;
    if (m.h > 0.0) finite_pairs += (long long)nt * n;
  }
  const size_t o_prob = take(sizeof(hm::BatchProb) * cnt), o_freq = take(sizeof(hm::BatchFreq) * cnt);
  // pair list: every (key, member) run padded to a multiple of 64 entries
  long long nfinite = 0;
  for (int p = 0; p < cnt; ++p) nfinite += b->mem[p].h > 0.0;
  const size_t code_cap = (size_t)finite_pairs + 64 * (size_t)hm::kNumKeys * (size_t)nfinite;
  const size_t o_codes = take(8 * code_cap);
  const size_t o_deep = take(4 * (size_t)cnt), o_prefix = take(8 * (size_t)(cnt + 1));
  const size_t o_rhs = take(sizeof(hm::GemmDesc<double>) * cnt), o_res = take(sizeof(hm::GemmDesc<double>) * cnt);
  const size_t o_nrmC = take(sizeof(hm::GemmDesc<double>) * (size_t)b->nirr);
  const size_t o_nrmD = take(sizeof(hm::GemmDesc<double>) * (size_t)b->nirr);
  const size_t o_lu64 = take(sizeof(hm::LuDesc<double>) * cnt), o_lu32 = take(sizeof(hm::LuDesc<float>) * cnt);
  const size_t o_convA = take(sizeof(hm::ConvDesc) * cnt), o_convR = take(sizeof(hm::ConvDesc) * cnt);
  const size_t o_convX = take(sizeof(hm::ConvDesc) * cnt), o_copyR = take(sizeof(hm::CopyDesc) * cnt);
  const size_t o_info = take(8 * (size_t)cnt), o_ratio = take(8 * (size_t)cnt);
  const size_t o_hist = take(sizeof(unsigned) * hm::kNumKeys), o_cursor = take(8 * hm::kNumKeys);
  BCHK(b, hipMalloc(&b->pool, tot), "batch alloc");
  b->pool_bytes = tot;
  BCHK(b, hipHostMalloc((void**)&b->h_freq, sizeof(hm::BatchFreq) * cnt, hipHostMallocDefault), "batch alloc");
  char* base = static_cast<char*>(b->pool);
  auto at = [&](size_t o) { return static_cast<void*>(base + o); };
  b->d_prob = (hm::BatchProb*)at(o_prob);
  b->d_freq = (hm::BatchFreq*)at(o_freq);
  b->d_codes = (uint64_t*)at(o_codes);
  b->d_deep = (int*)at(o_deep);
  b->d_deep_prefix = (long long*)at(o_prefix);
  b->d_rhs = (hm::GemmDesc<double>*)at(o_rhs);
  b->d_res = (hm::GemmDesc<double>*)at(o_res);
  b->d_nrmC = (hm::GemmDesc<double>*)at(o_nrmC);
  b->d_nrmD = (hm::GemmDesc<double>*)at(o_nrmD);
  b->d_lu64 = (hm::LuDesc<double>*)at(o_lu64);
  b->d_lu32 = (hm::LuDesc<float>*)at(o_lu32);
  b->d_convA = (hm::ConvDesc*)at(o_convA);
  b->d_convR = (hm::ConvDesc*)at(o_convR);
  b->d_convX = (hm::ConvDesc*)at(o_convX);
  b->d_copyR = (hm::CopyDesc*)at(o_copyR);
  b->d_info = (int*)at(o_info);
  b->d_ratio = (double*)at(o_ratio);
  // ---- descriptors and geometry
  std::vector<hm::BatchProb> prob(cnt);
  std::vector<hm::GemmDesc<double>> rhs(cnt), res(cnt), nrmC, nrmD;
  std::vector<hm::LuDesc<double>> lu64(cnt);
  std::vector<hm::LuDesc<float>> lu32(cnt);
  std::vector<hm::ConvDesc> convA(cnt), convR(cnt), convX(cnt);
  std::vector<hm::CopyDesc> copyR(cnt);
  std::vector<int> deep;
  std::vector<long long> prefix(1, 0);
  b->off_amas.assign(cnt, 0);
  b->h_a32.assign(cnt, nullptr);
  b->off_bdmp.assign(cnt, 0);
  b->off_exfc.assign(cnt, 0);
  b->out_amas = b->out_exfc = 0;
  for (int p = 0; p < cnt; ++p) {
    const Member& m = b->mem[p];
    const Off& o = off[p];
    const size_t n = m.n, nt = m.nt();
    BCHK(b, hipMemcpyAsync(at(o.body), m.body_start.data(), sizeof(int) * (m.nbody + 1), hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.fp), m.fp.data(), 24 * nt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.verts), m.verts.data(), 96 * n, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.ncn), m.ncn.data(), 4 * n, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.ds), m.ds.data(), 8 * n, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.pnsz), m.pnsz.data(), 8 * n, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(at(o.dxyz), m.dxyz.data(), 48 * n, hipMemcpyHostToDevice, b->stream), "batch upload");
    if (m.nbeta)
      BCHK(b, hipMemcpyAsync(at(o.beta), m.beta.data(), 8 * (size_t)m.nbeta, hipMemcpyHostToDevice, b->stream), "batch upload");
    hm::BatchProb& bp = prob[p];
    bp.g.n = m.n; bp.g.ni = m.ni; bp.g.nt = m.nt(); bp.g.nbody = m.nbody;
    bp.g.body_start = (const int*)at(o.body);
    bp.g.fp = (const double*)at(o.fp);
    bp.g.xyz_p = (const double*)at(o.fp);
    bp.g.verts = (const double*)at(o.verts);
    bp.g.ncn = (const int*)at(o.ncn);
    bp.g.ds = (const double*)at(o.ds);
    bp.g.pnsz = (const double*)at(o.pnsz);
    bp.g.dxyz = (const double*)at(o.dxyz);
    bp.A = (double*)at(o.A); bp.S = (double*)at(o.S); bp.W = (double*)at(o.W); bp.R = (double*)at(o.R);
    bp.ld = (long long)nt;
    bp.ncol = m.ncol;
    bp.nbeta = m.nbeta;
    bp.beta = (const double*)at(o.beta);
    bp.amas = (double*)at(o.amas); bp.bdmp = (double*)at(o.bdmp); bp.exfc = (double*)at(o.exfc);
    // R = S W
    rhs[p] = {(int)nt, m.ncol, m.n, 1.0, 0.0, 0.0, 0.0, bp.S, (long long)nt, bp.W, (long long)nt, bp.R, (long long)nt};
    // the n x n system and its n x ncol right-hand sides: A and R, or with
    // waterplane points C = A^T A and D = A^T R
    double* sysA = bp.A;
    long long lds = (long long)nt;
    double* sysB = bp.R;
    if (m.ni) {
      sysA = (double*)at(o.C);
      lds = (long long)n;
      sysB = (double*)at(o.D);
      nrmC.push_back({m.n, m.n, (int)nt, 1.0, 0.0, 0.0, 0.0, bp.A, (long long)nt, bp.A, (long long)nt, sysA, lds});
      nrmD.push_back({m.n, m.ncol, (int)nt, 1.0, 0.0, 0.0, 0.0, bp.A, (long long)nt, bp.R, (long long)nt, sysB,
                      (long long)nt});
    }
    bp.X = sysB;   // FP64: the solve overwrites the right-hand sides (IR: set below)
    int* info = b->d_info + 2 * p;
    lu64[p] = {sysA, lds, m.n, (int*)at(o.ipiv), info, (unsigned*)(info + 1), sysB, (long long)nt, m.ncol};
    if (gm) {
      // GMRES: the solution in W's slot (dead once R = S W is formed), the
      // right-hand sides kept in R; A demoted for the Krylov products
      bp.X = bp.W;
      b->h_a32[p] = (const float*)at(o.A32);
      convA[p] = {sysA, lds, (float*)at(o.A32), (long long)n, m.n, m.n};
    }
    if (ir) {
      // IR: sysB keeps the right-hand sides, X accumulates in W's slot (W is
      // dead once R = S W is formed), Rw holds residuals, X32 the FP32 corrections
      double* X = bp.W;
      double* Rw = (double*)at(o.Rw);
      float* A32 = (float*)at(o.A32);
      float* X32 = (float*)at(o.X32);
      bp.X = X;
      lu32[p] = {A32, (long long)n, m.n, (int*)at(o.ipiv), info, (unsigned*)(info + 1), X32, (long long)n, m.ncol};
      convA[p] = {sysA, lds, A32, (long long)n, m.n, m.n};
      // demote Rw -> X32; promote X32 -> X (overwrite on the first sweep, add later)
      convR[p] = {Rw, (long long)nt, X32, (long long)n, m.n, m.ncol};
      convX[p] = {X, (long long)nt, X32, (long long)n, m.n, m.ncol};
      copyR[p] = {sysB, (long long)nt, Rw, (long long)nt, m.n, m.ncol};
      res[p] = {m.n, m.ncol, m.n, -1.0, 0.0, 1.0, 0.0, sysA, lds, X, (long long)nt, Rw, (long long)nt};
    }
    if (m.h < 0.0) {
      deep.push_back(p);
      prefix.push_back(prefix.back() + (((long long)nt * n + 63) & ~63LL));
    }
    b->off_amas[p] = b->out_amas;
    b->off_bdmp[p] = b->out_amas;
    b->off_exfc[p] = b->out_exfc;
    b->out_amas += (size_t)m.m * m.m;
    b->out_exfc += 2 * (size_t)m.m * m.nbeta;
  }
  b->h_prob = prob;
  b->ndeep = (int)deep.size();
  b->deep_total = prefix.back();
  BCHK(b, hipMemcpyAsync(b->d_prob, prob.data(), sizeof(hm::BatchProb) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
  BCHK(b, hipMemcpyAsync(b->d_rhs, rhs.data(), sizeof(rhs[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
  BCHK(b, hipMemcpyAsync(b->d_lu64, lu64.data(), sizeof(lu64[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
  if (b->nirr) {
    BCHK(b, hipMemcpyAsync(b->d_nrmC, nrmC.data(), sizeof(nrmC[0]) * nrmC.size(), hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_nrmD, nrmD.data(), sizeof(nrmD[0]) * nrmD.size(), hipMemcpyHostToDevice, b->stream), "batch upload");
  }
  if (gm) {
    BCHK(b, hipMemcpyAsync(b->d_convA, convA.data(), sizeof(convA[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    size_t ws = 0;
    for (int p = 0; p < cnt; ++p) ws = std::max(ws, hm::gmres_workspace_bytes(b->mem[p].n, b->mem[p].ncol, b->gm_restart));
    BCHK(b, hipMalloc(&b->d_gws, ws), "batch alloc");
  }
  if (ir) {
    BCHK(b, hipMemcpyAsync(b->d_res, res.data(), sizeof(res[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_lu32, lu32.data(), sizeof(lu32[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_convA, convA.data(), sizeof(convA[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_convR, convR.data(), sizeof(convR[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_convX, convX.data(), sizeof(convX[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_copyR, copyR.data(), sizeof(copyR[0]) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
  }
  if (b->ndeep) {
    BCHK(b, hipMemcpyAsync(b->d_deep, deep.data(), sizeof(int) * deep.size(), hipMemcpyHostToDevice, b->stream), "batch upload");
    BCHK(b, hipMemcpyAsync(b->d_deep_prefix, prefix.data(), sizeof(long long) * prefix.size(), hipMemcpyHostToDevice, b->stream), "batch upload");
  }
  // wave state needs depths for the classification
  for (int p = 0; p < cnt; ++p) {
    memset(&b->h_freq[p], 0, sizeof(hm::BatchFreq));
    b->h_freq[p].h = b->mem[p].h;
  }
  BCHK(b, hipMemcpyAsync(b->d_freq, b->h_freq, sizeof(hm::BatchFreq) * cnt, hipMemcpyHostToDevice, b->stream), "batch upload");
  // ---- classification of the finite-depth members' pairs into one list sorted
  // by key (regime, series length) and, within a key, by member; each
  // (key, member) run padded to a multiple of 64 so that no wave mixes members
  unsigned* d_hist = (unsigned*)at(o_hist);
  unsigned long long* d_cursor = (unsigned long long*)at(o_cursor);
  std::vector<std::vector<unsigned>> hist(cnt);
  for (int p = 0; p < cnt; ++p) {
    if (!(b->mem[p].h > 0.0)) continue;
    hist[p].assign(hm::kNumKeys, 0u);
    BCHK(b, hipMemsetAsync(d_hist, 0, sizeof(unsigned) * hm::kNumKeys, b->stream), "batch classify");
    BCHK(b, hm::batch_classify(b->stream, b->d_prob, b->d_freq, p, b->mem[p].nt(), b->mem[p].n, d_hist, nullptr, nullptr,
                               true), "batch classify");
    BCHK(b, hipMemcpyAsync(hist[p].data(), d_hist, sizeof(unsigned) * hm::kNumKeys, hipMemcpyDeviceToHost, b->stream), "batch classify");
    BCHK(b, hipStreamSynchronize(b->stream), "batch classify");
  }
  std::vector<std::vector<unsigned long long>> cur(cnt);
  for (int p = 0; p < cnt; ++p)
    if (!hist[p].empty()) cur[p].assign(hm::kNumKeys, 0ull);
  unsigned long long acc = 0, real = 0;
  std::vector<unsigned long long> key_start(hm::kNumKeys + 1, 0);
  for (int k = 0; k < hm::kNumKeys; ++k) {
    key_start[k] = acc;
    for (int p = 0; p < cnt; ++p) {
      if (hist[p].empty()) continue;
      cur[p][k] = acc;
      acc += ((unsigned long long)hist[p][k] + 63ull) & ~63ull;
      real += hist[p][k];
    }
  }
  key_start[hm::kNumKeys] = acc;
  if ((long long)real != finite_pairs || acc > code_cap) return b->fail(HM_ERR_STATE, "batch classification lost pairs");
  for (int r = 0; r < 4; ++r) {
    b->seg_off[r] = (long long)key_start[r * hm::kKeyStride];
    b->seg_cnt[r] = (long long)(key_start[(r + 1) * hm::kKeyStride] - key_start[r * hm::kKeyStride]);
  }
  BCHK(b, hipMemsetAsync(b->d_codes, 0xFF, 8 * (size_t)acc, b->stream), "batch classify");   // padding = no pair
  for (int p = 0; p < cnt; ++p) {
    if (hist[p].empty()) continue;
    BCHK(b, hipMemcpyAsync(d_cursor, cur[p].data(), 8 * hm::kNumKeys, hipMemcpyHostToDevice, b->stream), "batch classify");
    BCHK(b, hm::batch_classify(b->stream, b->d_prob, b->d_freq, p, b->mem[p].nt(), b->mem[p].n, nullptr, d_cursor,
                               b->d_codes, false), "batch classify");
    BCHK(b, hipStreamSynchronize(b->stream), "batch classify");
  }
  b->grid = (int)((b->deep_total + 255) / 256);
  for (int r = 0; r < 4; ++r) b->grid += (int)((b->seg_cnt[r] + 255) / 256);
  // ---- autotuning on the middle of the frequency range (1 rad/s): the series
  // variant of the assembly, then the K chunk of the grouped LU's GEMM tiles
  b->variant = 1;
  b->tuned_ms = 0.0;
  b->gemm_kb = 16;
  for (double& t : b->tuned_lu_ms) t = 0.0;
  if (autotune) {
    std::vector<double> w(cnt, 1.0);
    int rc = write_freq(b, w.data());
    if (rc) return rc;
    double best = 1e30;
    int bv = b->variant;
    for (int v = 0; v <= 2; ++v) {
      b->variant = v;
      double ms;
      if ((rc = time_assembly(b, &ms))) return rc;
      if (ms < best) { best = ms; bv = v; }
    }
    b->variant = bv;
    b->tuned_ms = best;
    if (!gm && (rc = tune_lu(b))) return rc;
  }
  // ---- graph
  if (b->use_graph) {
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch capture");
    int rc = issue_chain(b);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
    BCHK(b, e, "batch capture");
    b->graph = gr;
    BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "batch graph instantiate");
  }
  b->prepared = true;
  return HM_OK;
}

extern "C" int hm_batch_solve(hm_batch* b, const double* w1, double* amas, double* bdmp, double* exfc) {
  if (!b) return HM_ERR_ARG - 1;
  if (!b->prepared) return b->fail(HM_ERR_STATE, "hm_batch_solve before hm_batch_prepare");
  if (!w1) return b->fail(HM_ERR_ARG - 2, "w1 is NULL");
  int rc = write_freq(b, w1);
  if (rc) return rc;
  if (b->use_graph) BCHK(b, hipGraphLaunch(b->exec, b->stream), "batch graph launch");
  else if ((rc = issue_chain(b))) return rc;
  const int cnt = b->count();
  if (b->precision == HM_GMRES) {
    b->gm_stats.assign(cnt, hm::GmresStats{});
    const std::vector<hm::BatchProb>& pr = b->h_prob;
    for (int p = 0; p < cnt; ++p) {
      const Member& m = b->mem[p];
      hm::GmresArgs ga;
      ga.n = m.n;
      ga.ncol = m.ncol;
      ga.restart = b->gm_restart;
      ga.maxiter = b->gm_maxiter;
      ga.tol = b->gm_tol;
      ga.inner_tol = 0.5 * b->gm_tol;
      ga.A64 = pr[p].A;
      ga.lda64 = (int)pr[p].ld;
      ga.A32 = b->h_a32[p];
      ga.B = pr[p].R;
      ga.ldb = (int)pr[p].ld;
      ga.X = pr[p].X;
      BCHK(b, hm::gmres_solve(b->stream, ga, b->d_gws, &b->gm_stats[p]), "batch gmres");
    }
    BCHK(b, hm::batch_forces(b->stream, b->d_prob, b->d_freq, cnt, b->mmax, b->ncolmax), "batch forces");
  }
  // zero pivots
  std::vector<int> info(2 * (size_t)cnt);
  BCHK(b, hipMemcpyAsync(info.data(), b->d_info, sizeof(int) * 2 * cnt, hipMemcpyDeviceToHost, b->stream), "batch info");
  b->last_ratio = 0.0;
  b->extra_sweeps = 0;
  if (b->precision == HM_FP32_IR) {
    // refinement beyond the captured sweeps until every member's last correction
    // is below 1e-13 of its solution (at most 10 more)
    std::vector<double> ratio(cnt);
    for (;;) {
      BCHK(b, hm::correction_ratio_grouped(b->stream, b->d_convX, cnt, b->d_ratio), "batch refinement");
      BCHK(b, hipMemcpyAsync(ratio.data(), b->d_ratio, sizeof(double) * cnt, hipMemcpyDeviceToHost, b->stream), "batch refinement");
      BCHK(b, hipStreamSynchronize(b->stream), "batch refinement");
      double worst = 0.0;
      for (double r : ratio) worst = fmax(worst, r);
      b->last_ratio = worst;
      if (worst < 1e-13 || b->extra_sweeps >= 10) break;
      if ((rc = extra_sweep(b))) return rc;
      ++b->extra_sweeps;
    }
    if (b->extra_sweeps)
      BCHK(b, hm::batch_forces(b->stream, b->d_prob, b->d_freq, cnt, b->mmax, b->ncolmax), "batch forces");
  }
  BCHK(b, hipStreamSynchronize(b->stream), "batch solve");
  for (int p = 0; p < cnt; ++p)
    if (info[2 * p] > 0) return b->fail(info[2 * p], "member %d: zero pivot in column %d", p, info[2 * p]);
  // outputs, stacked by member
  const std::vector<hm::BatchProb>& prob = b->h_prob;
  for (int p = 0; p < cnt; ++p) {
    const Member& m = b->mem[p];
    const size_t mm = (size_t)m.m * m.m;
    if (amas) BCHK(b, hipMemcpyAsync(amas + b->off_amas[p], prob[p].amas, 8 * mm, hipMemcpyDeviceToHost, b->stream), "batch output");
    if (bdmp) BCHK(b, hipMemcpyAsync(bdmp + b->off_bdmp[p], prob[p].bdmp, 8 * mm, hipMemcpyDeviceToHost, b->stream), "batch output");
    if (exfc && m.nbeta)
      BCHK(b, hipMemcpyAsync(exfc + b->off_exfc[p], prob[p].exfc, 16 * (size_t)m.m * m.nbeta, hipMemcpyDeviceToHost, b->stream), "batch output");
  }
  BCHK(b, hipStreamSynchronize(b->stream), "batch output");
  return HM_OK;
}

extern "C" int hm_batch_set_gmres(hm_batch* b, int restart, double tol, int maxiter) {
  if (!b) return HM_ERR_ARG - 1;
  if (restart < 1 || restart > hm::kGmresMaxRestart)
    return b->fail(HM_ERR_ARG - 2, "restart must be in 1..%d", hm::kGmresMaxRestart);
  if (!(tol > 0.0 && tol < 1.0)) return b->fail(HM_ERR_ARG - 3, "tol must be in (0, 1)");
  if (maxiter < 1) return b->fail(HM_ERR_ARG - 4, "maxiter must be >= 1");
  b->gm_restart = restart;
  b->gm_tol = tol;
  b->gm_maxiter = maxiter;
  b->prepared = false;   // the workspace is sized for the restart length
  return HM_OK;
}

extern "C" int hm_batch_gmres_stats(hm_batch* b, int* max_iterations, int* max_cycles, int* all_converged,
                                    double* worst_rel_residual) {
  if (!b) return HM_ERR_ARG - 1;
  if (b->precision != HM_GMRES || b->gm_stats.empty())
    return b->fail(HM_ERR_STATE, "no HM_GMRES batch solve yet");
  int it = 0, cy = 0, ok = 1;
  double worst = 0.0;
  for (const hm::GmresStats& st : b->gm_stats) {
    it = std::max(it, st.iterations);
    cy = std::max(cy, st.cycles);
    ok = ok && st.converged;
    worst = std::max(worst, st.rel_residual);
  }
  if (max_iterations) *max_iterations = it;
  if (max_cycles) *max_cycles = cy;
  if (all_converged) *all_converged = ok;
  if (worst_rel_residual) *worst_rel_residual = worst;
  return HM_OK;
}

extern "C" int hm_batch_info(hm_batch* b, int* members, int* nmax, int* series_variant, int* grid_blocks,
                             double* tuned_assembly_ms, double* last_correction, int* extra_sweeps,
                             long long* device_bytes) {
  if (!b) return HM_ERR_ARG - 1;
  if (members) *members = b->count();
  if (nmax) *nmax = b->nmax;
  if (series_variant) *series_variant = b->variant;
  if (grid_blocks) *grid_blocks = b->grid;
  if (tuned_assembly_ms) *tuned_assembly_ms = b->tuned_ms;
  if (last_correction) *last_correction = b->last_ratio;
  if (extra_sweeps) *extra_sweeps = b->extra_sweeps;
  if (device_bytes) *device_bytes = (long long)b->pool_bytes;
  return HM_OK;
}
