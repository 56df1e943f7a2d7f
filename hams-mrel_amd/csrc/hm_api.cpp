// hm_api.cpp -- C-ABI of the boundary-integral hot path (include/hm_bem.h).
//
// A context owns one HIP stream and the HBM-resident state of one problem:
// geometry, the classified pair lists, the per-frequency wave state, the
// assembled matrices and the factors. The per-frequency chain runs entirely on
// the device; only the 6nb x 6nb added mass / damping and the excitation forces
// come back to the host.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <random>
#include <string>
#include <thread>
#include <algorithm>
#include <vector>

#include "hm_noalloc.hpp"
#include "hm_assembly.hpp"
#include "hm_bem.h"
#include "hm_device.hpp"
#include "hm_gmres.hpp"
#include "hm_hmatrix.hpp"
#include "hm_htree.hpp"
#include "hm_linalg.hpp"
#include "hm_surrogate.hpp"

namespace {

using hm::kNk;

// grow-only device buffer
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b ? b : 16);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

constexpr int kStages = 14;
enum Stage { kClassify = 0, kAssembly = 1, kRhs = 2, kIrr = 3, kLu = 4, kSolve = 5, kForces = 6,
             kAsmRegime0 = 7 /* + regime: Linton, Pidcock, eps-accelerated, eigenfunction */,
             kAsmSurrogate = 11, kAsmCross = 12, kAcaEntries = 13 };

}  // namespace

struct hm_ctx {
  hipStream_t stream = nullptr;
  std::string err;
  // constants
  double rho = 1025.0, g = 9.81, g_disp = 9.81, amp = 1.0, xw = 0.0, yw = 0.0;
  // geometry
  int n = 0, ni = 0, nbody = 0;
  // smallest horizontal gap between the bodies' bounding circles (panel centroids)
  double min_body_gap = 0.0;
  std::vector<int> body_start;
  DevBuf d_body, d_fp, d_verts, d_ncn, d_ds, d_pnsz, d_dxyz;
  bool have_geom = false;
  // body frames (hm_set_body_origins): the coordinates are local to each body
  // and d_org holds the horizontal placement of each field point's body
  DevBuf d_org;
  bool frames = false;
  // host copies of the per-panel inputs of the influence matrices (the
  // comparison of bodies for shared diagonal blocks)
  std::vector<double> fp_host, verts_host, ds_host, pnsz_host, dxyz_host;
  std::vector<int> ncn_host;
  // body_rep[b]: the first body whose hull panels are bitwise those of body b
  // (frames on); H-matrix diagonal blocks are assembled and factored once per
  // distinct body and copied to the others (block_share)
  std::vector<int> body_rep;
  bool block_share = true;
  DevBuf d_ublk;                // the distinct bodies (classification of the diagonal blocks)
  int n_ublk = 0;
  std::vector<double> origin_host;   // the placements (2 per body)
  std::vector<int> wp_start_host;    // the waterplane points' bodies (hm_set_body_origins)
  // the arguments of hm_set_geometry that the host copies above do not keep
  std::vector<double> geom_xyz;
  std::vector<int> geom_ncon;
  int geom_ntnd = 0;
  int device = 0;
  // hm_run_frequencies keeps `nworkers` frequencies in flight: this context and
  // nworkers - 1 clones of its configuration (own streams and buffers), each
  // solving the frequencies it takes in a host thread of its own. cfg_gen
  // counts the changes of geometry, depth, headings, constants and surrogate
  // parameters; the clones are rebuilt when it or a setting differs.
  int nworkers = 1;
  std::vector<hm_ctx*> workers;
  unsigned long long cfg_gen = 0, workers_gen = ~0ULL;
  std::string workers_sig;
  // cross approximation: process p's block equals process pshare[p]'s (equal
  // bodies at an equal offset, with block sharing on)
  DevBuf d_pshare, d_prun;           // the map and the list of the processes that run
  DevBuf d_evt, d_pmt;               // the cross approximation's per-frequency depth-factor tables
  std::vector<int> pshare_host;
  // share_sym: a process may also take the block of a process whose offset a
  // symmetry of the (one) body class maps onto its own, its rows and columns
  // permuted by that symmetry's panel map (psym_host[p]: 0 = plain copy, s =
  // map s of kSymMap); sym_ok[s]: map s takes the body's hull panels onto
  // themselves, sym_iperm[s m + k]: the panel that map s takes onto panel k
  bool share_sym = true;
  std::vector<int> psym_host, sym_ok, sym_iperm;
  DevBuf d_psym, d_sperm;
  bool pshare_ok = false;            // d_pshare holds the current geometry's map
  double hm_rank_mean_run = 0.0;     // mean rank of the processes that ran
  int n_pshared = 0;                 // processes that copy another's block
  bool sharing() const {
    if (!block_share || !hm_on) return false;
    for (int b = 0; b < nbody; ++b)
      if (body_rep[b] != b) return true;
    return false;
  }
  // depth + classification
  double h = 0.0;
  bool have_depth = false;
  DevBuf d_pairs, d_hist, d_cursor;
  std::vector<long long> hist_host;
  long long seg_off[4] = {0, 0, 0, 0}, seg_cnt[4] = {0, 0, 0, 0};
  // headings
  std::vector<double> beta;
  DevBuf d_beta;
  // frequency
  double w1 = 0.0, v = 0.0, wk = 0.0;
  double wvn[kNk];
  DevBuf d_wvn;
  bool have_freq = false;
  // system
  DevBuf d_A, d_S, d_W, d_R, d_C, d_D, d_X, d_A32, d_X32, d_Rw, d_ipiv, d_info, d_counter, d_norm;
  DevBuf d_amas, d_bdmp, d_exfc;
  bool assembled = false, factored = false, solved = false;
  int precision = HM_FP64;
  // GMRES (ZITER_SOLVE_MULTI parameters, IterativeSolvers.f90:56-58; the
  // block-diagonal preconditioner of SETUP_BLOCK_PRECONDITIONER :190)
  int gm_restart = 50, gm_maxiter = 500, gm_precond = -1;
  double gm_tol = 1e-6;
  DevBuf d_P32, d_pipiv, d_gwork;
  DevBuf d_luws;              // workspace of the cooperative LU panel kernel
  bool lu_coop = true;        // dense LU panels by one cooperative launch each where the grid fits
  void* luws() const { return lu_coop ? d_luws.p : nullptr; }
  DevBuf d_Pinv;              // explicit FP32 inverses of the preconditioner blocks
  DevBuf d_gjws;              // workspace of their Gauss-Jordan inversion (H-matrix form)
  DevBuf d_Pinv16;            // FP16 copy of d_Pinv (H-matrix form, pinv16_on)
  bool pinv16_on = true;      // the H-matrix form's preconditioner applied by FP16 MFMA GEMMs
  bool pinv16_ready = false;
  double gm_cycle_rtol = 1e-4;  // with the FP16 preconditioner: the Arnoldi target of a cycle (GmresArgs)
  bool pinv16_shared = false; // d_Pinv16 holds one block, shared by every body
  DevBuf d_pperm;             // the factored blocks' pivot permutations (with d_Pinv)
  bool pinv_ready = false;    // d_Pinv and d_pperm belong to the current d_P32 factors
  std::vector<int> gm_blocks;
  hm::GmresStats gm_stats;
  // learned surrogate of the evanescent sums (R/h >= 0.05)
  DevBuf d_sur, d_fb, d_fbcnt;
  bool sur_ready = false, sur_on = false;
  int sur_quant = 1;
  double sur_tol = 1e-4;
  long long sur_counts[3] = {0, 0, 0};
  // H-matrix diagonal blocks: PIDCOCK pairs assembled together with their mirrors
  bool pair_sym = true;
  // the frequency-independent halves of those pairs' PIDCOCK terms (built with
  // the classification; off when the device memory for it is not available)
  DevBuf d_pcache, d_ptab;
  bool pcache_on = false;
  hm::PidCache pcache() const {
    return pcache_on ? hm::PidCache{d_pcache.as<double>(), d_ptab.as<long long>()} : hm::PidCache{nullptr, nullptr};
  }
  // the infinite-depth branch of the surrogate (its own network and error map)
  DevBuf d_surd, d_surdcnt;
  bool surd_ready = false, surd_on = false;
  int surd_quant = 1;
  double surd_tol = 1e-4;
  static void* img_of(const DevBuf& b) { return static_cast<char*>(b.p) + hm::kSurImgOffset; }
  static hm::SurDev dev_of(const DevBuf& b, int quant) {
    const float* p = b.as<float>();
    return hm::SurDev{p, p + hm::kSurParams, p + 2 * hm::kSurParams, p + 2 * hm::kSurParams + hm::kSurOut, quant,
                      img_of(b)};
  }
  hm::SurDev surd_dev() const { return dev_of(d_surd, surd_quant); }
  void* sur_img() const { return static_cast<char*>(d_sur.p) + hm::kSurImgOffset; }
  hm::SurDev sur_dev() const {
    const float* b = d_sur.as<float>();
    return hm::SurDev{b, b + hm::kSurParams, b + 2 * hm::kSurParams, b + 2 * hm::kSurParams + hm::kSurOut, sur_quant,
                      sur_img()};
  }
  // H-matrix form (equal bodies, no waterplane points): dense diagonal blocks
  // (block b at b m (m+1) complex elements, leading dimension m, so that the
  // pair kernels' i + j ld addressing lands in it) and cross-approximated
  // off-diagonal blocks
  bool hm_on = false;
  double hm_eps = 1e-4, hm_eps_s = 1e-4;
  int hm_kmax = 50;
  DevBuf d_hmat, d_hD, d_hDS, d_hD32, d_hU32, d_hV32, d_htb, d_hts;
  DevBuf d_hUc[2], d_hVc[2];   // factors of A and S packed at their ranks (hmat_pack)
  DevBuf d_lrmeta[2];           // the packing's device metadata
  DevBuf d_rhsx, d_rhsy;        // the right-hand sides' nonzero columns per body and their products
  hm::LowRankPlan lr_plan[2];
  hm::LowRank hm_lr[2];
  int hm_kuse[2] = {0, 0};      // the largest rank of each matrix
  hm::HmatDev hmat{};
  // H-matrix built by hm_ziter_solve (ISOLV = 2) from an assembled matrix:
  // hd_nb clusters of hd_m rows
  bool hd_ready = false;
  int hd_nb = 0, hd_m = 0;
  // H-matrix over the geometric cluster tree (hm_hmatrix_build_tree, or the
  // assembly-aware form where the body clusters do not apply: one body, bodies
  // of unequal panel counts, interior waterplane points, or on request)
  hm::HTree htree;
  bool ht_ready = false;       // htree holds the H-matrix of the current system
  bool hm_tree = false;        // hm_set_hmatrix chose the cluster-tree form
  int ht_mode = 0, ht_leaf = 32;
  double ht_eta = 2.0;
  DevBuf d_htw;                // workspace of the tree products
  bool ht_from_build = false;  // htree came from hm_hmatrix_build_tree (serves hm_ziter_solve)
  std::vector<int> ht_blocks;  // its preconditioner blocks (empty: none)
  std::vector<double> xyz_host;   // panel centroids, x, y, z per panel
  int hm_unconverged = 0, hm_steps = 0, hm_rank_max = 0;
  double hm_rank_mean = 0.0;
  int bm() const { return n / nbody; }
  size_t hm_diag_elems() const { return (size_t)nbody * bm() * (bm() + 1); }
  // epsilon-accelerated regimes: 0 one pair per lane; lane groups of 8 (EIGENE)
  // and 16 (PIDCOCK) lanes: 1 (default); of 4 and 8 lanes: 2
  int series_variant = 1;
  // timing: event pairs recorded on the stream, resolved by hm_get_timing
  bool timing = false;
  double stage_ms[kStages] = {0};
  int stage_cnt[kStages] = {0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // H-matrix path: the cross approximation runs on a second stream while the
  // main stream assembles the diagonal blocks and factors (and inverts) the
  // block preconditioner
  hipStream_t side = nullptr;
  hipStream_t aca_stream = nullptr;   // the stream the current cross approximation runs on
  bool hm_overlap = true;             // cross approximation beside the diagonal blocks
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool side_pending = false;     // work on `side` the main stream has not joined yet
  bool prec_early = false;       // the preconditioner of the current assembly and settings is built
  DevBuf d_info2, d_counter2;    // its info words and counters
  // make the main stream wait for the side stream's work
  void join_side() {
    if (side_pending) {
      (void)hipEventRecord(ev_join, side);
      (void)hipStreamWaitEvent(stream, ev_join, 0);
      side_pending = false;
    }
  }
  struct Rec { int stage; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  int open_stage[kStages];
  hipEvent_t open_ev[kStages];

  int nt() const { return n + ni; }
  int m() const { return 6 * nbody; }
  int ncol() const { return 6 * nbody + (int)beta.size(); }
  hm::GeomDev geom() const {
    hm::GeomDev gd;
    gd.n = n; gd.ni = ni; gd.nt = n + ni; gd.nbody = nbody;
    gd.body_start = d_body.as<int>();
    gd.fp = d_fp.as<double>();
    gd.xyz_p = d_fp.as<double>();
    gd.verts = d_verts.as<double>();
    gd.ncn = d_ncn.as<int>();
    gd.ds = d_ds.as<double>();
    gd.pnsz = d_pnsz.as<double>();
    gd.dxyz = d_dxyz.as<double>();
    gd.org = frames ? d_org.as<double>() : nullptr;
    return gd;
  }
  hm::FreqDev freq() const {
    hm::FreqDev f;
    f.v = v; f.w1 = w1; f.wk = wk; f.h = h; f.rho = rho; f.g = g; f.amp = amp;
    f.wvn = d_wvn.as<double>();
    return f;
  }
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
    (void)hipGetLastError();   // do not let this failure resurface at a later launch check
    return fail(HM_ERR_HIP - (int)e, "%s: %s", what, hipGetErrorString(e));
  }
  hipEvent_t take_event() {
    hipEvent_t e = nullptr;
    if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
    else (void)hipEventCreate(&e);
    return e;
  }
  void stage_begin(int s, hipStream_t on = nullptr) {
    if (!timing) return;
    open_ev[s] = take_event();
    (void)hipEventRecord(open_ev[s], on ? on : stream);
  }
  int stage_end(int s, hipStream_t on = nullptr) {
    if (!timing) return HM_OK;
    hipEvent_t e = take_event();
    (void)hipEventRecord(e, on ? on : stream);
    recs.push_back({s, open_ev[s], e});
    return HM_OK;
  }
  int resolve_timing() {
    if (recs.empty()) return HM_OK;
    for (const Rec& r : recs) {   // stages may end on either stream
      hipError_t e = hipEventSynchronize(r.b);
      if (e != hipSuccess) return hip(e, "stage timing");
    }
    for (const Rec& r : recs) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      stage_ms[r.stage] += ms;
      stage_cnt[r.stage] += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    recs.clear();
    return HM_OK;
  }
};

#define HIPCHK(ctx, call, what)                              \
  do {                                                       \
    int _rc = (ctx)->hip((call), (what));                    \
    if (_rc != HM_OK) return _rc;                            \
  } while (0)

// ---------------------------------------------------------------------------
// DISPERSION, WavDynSubs.f90:40-127 (Newman 1990 third-order iteration)
// ---------------------------------------------------------------------------
extern "C" int hm_dispersion(double* wvn, int nrt, double w, double h, double g_disp) {
  if (!wvn) return HM_ERR_ARG - 1;
  if (nrt < 1) return HM_ERR_ARG - 2;
  if (!(h > 0.0)) return HM_ERR_ARG - 4;
  const double pi = hm::kPi;
  const double x = w * w * h / g_disp;
  double y;
  if (x > 0.0 && x <= 2.0) y = sqrt(x) * (0.9994 + 0.1701 * x + F32(0.0305) * x * x);
  else {
    const double t = x * exp(-2.0 * x);
    y = x + 2.0 * t - 6.0 * t * t;
  }
  if (fabs(y - x) < F32(1.e-10)) {
    wvn[0] = x / h;
  } else {
    double exx = 1.0;
    while (exx > 1.0e-10) {
      const double trial = y, dnm = trial * trial - x * x;
      const double fun = log((trial + x) / (trial - x)) / 2.0 - trial;
      const double dfun = -x / dnm - 1.0, d2fun = 2.0 * x * trial / (dnm * dnm);
      y = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
      exx = fabs(y - trial);
    }
    wvn[0] = y / h;
  }
  const double u = 3.0 * x / (7.0 + 3.0 * x);
  double t = 0.0159 + 0.1032 * u + 4.3152 * u * u - 2.8768 * u * u * u;
  for (int i = 2; i <= nrt; ++i) {
    double exx = 1.0;
    while (exx > 1.0e-10) {
      const double trial = t, yy = (i - 1) * pi - trial, dnm = yy * yy + x * x;
      const double fun = atan2(x, yy) - trial;
      const double dfun = x / dnm - 1.0, d2fun = 2.0 * x * trial / (dnm * dnm);
      t = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
      exx = fabs(t - trial);
    }
    wvn[i - 1] = ((i - 1) * pi - t) / h;
    t = t - pi * x / (x * x + pi * i * (pi * (i - 1) - t));
  }
  return HM_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int hm_create(hm_ctx** out) {
  if (!out) return HM_ERR_ARG - 1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return HM_ERR_HIP - (int)hipErrorNoDevice;
  hm_ctx* c = new hm_ctx();
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  // the fixed device array of hm_fingreen3d_batch_noalloc: its address on this
  // device (no allocation)
  if (e == hipSuccess) e = hm::noalloc_setup();
  if (e != hipSuccess) {
    delete c;
    return HM_ERR_HIP - (int)e;
  }
  *out = c;
  return HM_OK;
}

extern "C" int hm_destroy(hm_ctx* c) {
  if (!c) return HM_ERR_ARG - 1;
  for (hm_ctx* w : c->workers) hm_destroy(w);
  c->workers.clear();
  if (c->side) (void)hipStreamSynchronize(c->side);
  (void)hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->d_body, &c->d_org, &c->d_ublk, &c->d_pshare, &c->d_prun, &c->d_evt, &c->d_pmt, &c->d_fp, &c->d_verts, &c->d_ncn, &c->d_ds, &c->d_pnsz, &c->d_dxyz,
                    &c->d_pairs, &c->d_hist, &c->d_cursor, &c->d_beta, &c->d_wvn, &c->d_A, &c->d_S,
                    &c->d_W, &c->d_R, &c->d_C, &c->d_D, &c->d_X, &c->d_A32, &c->d_X32, &c->d_Rw,
                    &c->d_ipiv, &c->d_info, &c->d_counter, &c->d_norm, &c->d_amas, &c->d_bdmp, &c->d_exfc,
                    &c->d_P32, &c->d_pipiv, &c->d_gwork, &c->d_luws, &c->d_Pinv, &c->d_gjws, &c->d_Pinv16, &c->d_pperm, &c->d_sur, &c->d_fb, &c->d_fbcnt, &c->d_pcache, &c->d_ptab, &c->d_info2, &c->d_counter2,
                    &c->d_hmat, &c->d_hD, &c->d_hDS, &c->d_hD32, &c->d_hU32, &c->d_hV32, &c->d_htb, &c->d_hts,
                    &c->d_hUc[0], &c->d_hUc[1], &c->d_hVc[0], &c->d_hVc[1], &c->d_lrmeta[0], &c->d_lrmeta[1], &c->d_rhsx, &c->d_rhsy, &c->d_htw,
                    &c->d_surd, &c->d_surdcnt};
  for (DevBuf* b : bufs) b->release();
  c->htree.release();
  (void)c->resolve_timing();
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return HM_OK;
}

extern "C" const char* hm_error_string(const hm_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int hm_set_constants(hm_ctx* c, double rho, double g, double g_disp, double amp, double xw, double yw) {
  if (!c) return HM_ERR_ARG - 1;
  if (!(rho > 0)) return c->fail(HM_ERR_ARG - 2, "rho must be > 0");
  if (!(g > 0)) return c->fail(HM_ERR_ARG - 3, "g must be > 0");
  if (!(g_disp > 0)) return c->fail(HM_ERR_ARG - 4, "g_disp must be > 0");
  c->rho = rho; c->g = g; c->g_disp = g_disp; c->amp = amp; c->xw = xw; c->yw = yw;
  c->have_freq = false;
  ++c->cfg_gen;
  return HM_OK;
}

template <typename T>
static int upload(hm_ctx* c, DevBuf& b, const T* src, size_t count, const char* what) {
  HIPCHK(c, b.ensure(count * sizeof(T)), what);
  HIPCHK(c, hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream), what);
  return HM_OK;
}

// smallest horizontal gap between the bodies' bounding circles, from the
// global panel centroids in xyz_host
static void set_body_gap(hm_ctx* c) {
  const int nbody = c->nbody;
  const double* xyz_p = c->xyz_host.data();
  std::vector<double> cx(nbody, 0.0), cy(nbody, 0.0), rad(nbody, 0.0);
  for (int b = 0; b < nbody; ++b) {
    const int j0 = c->body_start[b], j1 = c->body_start[b + 1];
    for (int j = j0; j < j1; ++j) { cx[b] += xyz_p[3 * (size_t)j]; cy[b] += xyz_p[3 * (size_t)j + 1]; }
    cx[b] /= (j1 - j0);
    cy[b] /= (j1 - j0);
    for (int j = j0; j < j1; ++j)
      rad[b] = std::max(rad[b], std::hypot(xyz_p[3 * (size_t)j] - cx[b], xyz_p[3 * (size_t)j + 1] - cy[b]));
  }
  c->min_body_gap = nbody > 1 ? 1e300 : 0.0;
  for (int b1 = 0; b1 < nbody; ++b1)
    for (int b2 = b1 + 1; b2 < nbody; ++b2)
      c->min_body_gap = std::min(c->min_body_gap, std::hypot(cx[b1] - cx[b2], cy[b1] - cy[b2]) - rad[b1] - rad[b2]);
}

static int classify(hm_ctx* c);

extern "C" int hm_set_geometry(hm_ctx* c, int nbody, const int* body_start, int ntnd, const double* xyz,
                               const int* ncn, const int* ncon, const double* xyz_p, const double* ds,
                               const double* pnsz, const double* dxyz_p, int ni, const double* ixyz_p) {
  if (!c) return HM_ERR_ARG - 1;
  if (nbody < 1) return c->fail(HM_ERR_ARG - 2, "nbody must be >= 1");
  if (!body_start) return c->fail(HM_ERR_ARG - 3, "body_start is NULL");
  const int n = body_start[nbody];
  if (body_start[0] != 0 || n < 1) return c->fail(HM_ERR_ARG - 3, "body_start must start at 0 and end at n >= 1");
  for (int b = 0; b < nbody; ++b)
    if (body_start[b + 1] <= body_start[b]) return c->fail(HM_ERR_ARG - 3, "body %d has no panels", b);
  if (ntnd < 3 || !xyz) return c->fail(HM_ERR_ARG - 5, "nodes missing");
  if (!ncn || !ncon || !xyz_p || !ds || !pnsz || !dxyz_p) return c->fail(HM_ERR_ARG - 6, "panel arrays missing");
  if (ni < 0 || (ni > 0 && !ixyz_p)) return c->fail(HM_ERR_ARG - 12, "bad waterplane field points");
  if (n + ni >= 65536) return c->fail(HM_ERR_ARG - 3, "n + ni = %d exceeds the 16-bit pair encoding", n + ni);
  std::vector<double> verts((size_t)n * 12, 0.0);
  for (int j = 0; j < n; ++j) {
    if (ncn[j] != 3 && ncn[j] != 4) return c->fail(HM_ERR_ARG - 7, "panel %d has %d vertices", j + 1, ncn[j]);
    for (int v = 0; v < ncn[j]; ++v) {
      const int node = ncon[4 * j + v];
      if (node < 1 || node > ntnd) return c->fail(HM_ERR_ARG - 8, "panel %d vertex %d = %d out of range", j + 1, v + 1, node);
      for (int k = 0; k < 3; ++k) verts[12 * (size_t)j + 3 * v + k] = xyz[3 * (size_t)(node - 1) + k];
    }
  }
  std::vector<double> fp((size_t)(n + ni) * 3);
  memcpy(fp.data(), xyz_p, sizeof(double) * 3 * (size_t)n);
  if (ni) memcpy(fp.data() + 3 * (size_t)n, ixyz_p, sizeof(double) * 3 * (size_t)ni);
  c->n = n; c->ni = ni; c->nbody = nbody;
  c->body_start.assign(body_start, body_start + nbody + 1);
  c->xyz_host.assign(xyz_p, xyz_p + 3 * (size_t)n);
  set_body_gap(c);
  int rc;
  if ((rc = upload(c, c->d_body, body_start, nbody + 1, "geometry")) ||
      (rc = upload(c, c->d_fp, fp.data(), fp.size(), "geometry")) ||
      (rc = upload(c, c->d_verts, verts.data(), verts.size(), "geometry")) ||
      (rc = upload(c, c->d_ncn, ncn, n, "geometry")) || (rc = upload(c, c->d_ds, ds, n, "geometry")) ||
      (rc = upload(c, c->d_pnsz, pnsz, n, "geometry")) || (rc = upload(c, c->d_dxyz, dxyz_p, 6 * (size_t)n, "geometry")))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream), "geometry upload");
  c->fp_host = std::move(fp);
  c->verts_host = std::move(verts);
  c->geom_xyz.assign(xyz, xyz + 3 * (size_t)ntnd);
  c->geom_ncon.assign(ncon, ncon + 4 * (size_t)n);
  c->geom_ntnd = ntnd;
  c->wp_start_host.clear();
  ++c->cfg_gen;
  c->ncn_host.assign(ncn, ncn + n);
  c->ds_host.assign(ds, ds + n);
  c->pnsz_host.assign(pnsz, pnsz + n);
  c->dxyz_host.assign(dxyz_p, dxyz_p + 6 * (size_t)n);
  c->frames = false;
  c->pshare_ok = false;
  c->origin_host.assign(2 * (size_t)nbody, 0.0);
  c->body_rep.resize(nbody);
  for (int b = 0; b < nbody; ++b) c->body_rep[b] = b;
  c->have_geom = true;
  c->have_depth = false;
  c->hm_on = false;
  c->hm_tree = false;
  c->ht_ready = false;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

static int pshare_build(hm_ctx* c);

// body b's hull panels are bitwise those of body a: every input of the
// influence matrices' (a, a) block equals that of (b, b)
static bool same_body(const hm_ctx* c, int a, int b) {
  const int ja = c->body_start[a], jb = c->body_start[b], m = c->body_start[a + 1] - ja;
  if (c->body_start[b + 1] - jb != m) return false;
  auto eq = [&](const std::vector<double>& v, int w) {
    return memcmp(v.data() + (size_t)w * ja, v.data() + (size_t)w * jb, sizeof(double) * (size_t)w * m) == 0;
  };
  if (memcmp(c->ncn_host.data() + ja, c->ncn_host.data() + jb, sizeof(int) * (size_t)m) != 0) return false;
  if (!eq(c->fp_host, 3) || !eq(c->verts_host, 12) || !eq(c->ds_host, 1) || !eq(c->pnsz_host, 1)) return false;
  for (int j = 0; j < m; ++j)
    for (int k = 0; k < 3; ++k)
      if (memcmp(&c->dxyz_host[6 * (size_t)(ja + j) + k], &c->dxyz_host[6 * (size_t)(jb + j) + k], sizeof(double)))
        return false;
  return true;
}

extern "C" int hm_set_body_origins(hm_ctx* c, int nbody, const double* origin, const int* wp_start) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_geom) return c->fail(HM_ERR_STATE, "hm_set_body_origins before hm_set_geometry");
  if (nbody != c->nbody) return c->fail(HM_ERR_ARG - 2, "nbody = %d, the geometry has %d bodies", nbody, c->nbody);
  if (!origin) return c->fail(HM_ERR_ARG - 3, "origin is NULL");
  for (int k = 0; k < 2 * nbody; ++k)
    if (!std::isfinite(origin[k])) return c->fail(HM_ERR_ARG - 3, "origin[%d] is not finite", k);
  if (c->ni > 0) {
    if (!wp_start) return c->fail(HM_ERR_ARG - 4, "wp_start is NULL with %d waterplane points", c->ni);
    if (wp_start[0] != 0 || wp_start[nbody] != c->ni)
      return c->fail(HM_ERR_ARG - 4, "wp_start must run from 0 to ni = %d", c->ni);
    for (int b = 0; b < nbody; ++b)
      if (wp_start[b + 1] < wp_start[b]) return c->fail(HM_ERR_ARG - 4, "wp_start decreases at body %d", b);
  }
  const int n = c->n, nt = c->nt();
  std::vector<double> org(2 * (size_t)nt);
  for (int b = 0; b < nbody; ++b) {
    for (int j = c->body_start[b]; j < c->body_start[b + 1]; ++j) {
      org[2 * (size_t)j] = origin[2 * b];
      org[2 * (size_t)j + 1] = origin[2 * b + 1];
    }
    if (c->ni > 0)
      for (int p = wp_start[b]; p < wp_start[b + 1]; ++p) {
        org[2 * (size_t)(n + p)] = origin[2 * b];
        org[2 * (size_t)(n + p) + 1] = origin[2 * b + 1];
      }
  }
  int rc = upload(c, c->d_org, org.data(), org.size(), "body origins");
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream), "body origins");
  c->frames = true;
  c->origin_host.assign(origin, origin + 2 * nbody);
  c->wp_start_host.assign(nbody + 1, 0);
  if (c->ni > 0) c->wp_start_host.assign(wp_start, wp_start + nbody + 1);
  ++c->cfg_gen;
  c->pshare_ok = false;
  // global centroids for the body gaps and the cluster tree
  for (int j = 0; j < n; ++j) {
    c->xyz_host[3 * (size_t)j] = c->fp_host[3 * (size_t)j] + org[2 * (size_t)j];
    c->xyz_host[3 * (size_t)j + 1] = c->fp_host[3 * (size_t)j + 1] + org[2 * (size_t)j + 1];
  }
  set_body_gap(c);
  std::vector<int> u;
  for (int b = 0; b < nbody; ++b) {
    c->body_rep[b] = b;
    for (int a = 0; a < b; ++a)
      if (c->body_rep[a] == a && same_body(c, a, b)) { c->body_rep[b] = a; break; }
    if (c->body_rep[b] == b) u.push_back(b);
  }
  if ((rc = upload(c, c->d_ublk, u.data(), u.size(), "body origins"))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream), "body origins");
  c->n_ublk = (int)u.size();
  c->assembled = c->factored = c->solved = false;
  if (c->have_depth && c->h > 0.0) return classify(c);
  return HM_OK;
}

extern "C" int hm_get_body_classes(hm_ctx* c, int nbody, int* rep) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_geom) return c->fail(HM_ERR_STATE, "no geometry");
  if (nbody != c->nbody || !rep) return c->fail(HM_ERR_ARG - 2, "bad output array");
  for (int b = 0; b < nbody; ++b) rep[b] = c->body_rep[b];
  return HM_OK;
}

extern "C" int hm_set_precond_fp16(hm_ctx* c, int enable) {
  if (!c) return HM_ERR_ARG - 1;
  c->pinv16_on = enable != 0;
  c->pinv16_ready = false;
  c->prec_early = false;
  c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_cgemm_f16_check(hm_ctx* c, int m, int n, int k, const float* a, const float* b, float* out) {
  if (!c) return HM_ERR_ARG - 1;
  if (m < 1 || n < 1 || k < 1) return c->fail(HM_ERR_ARG - 2, "sizes must be >= 1");
  if (!a || !b || !out) return c->fail(HM_ERR_ARG - 5, "arrays missing");
  const size_t na = 2 * (size_t)m * k, nb = 2 * (size_t)k * n, nc = 2 * (size_t)m * n;
  const int kp = hm::h16_kp(k);
  DevBuf da, db, dc, ha, hb;
  int rc = 0;
  if ((rc = upload(c, da, a, na, "cgemm f16")) || (rc = upload(c, db, b, nb, "cgemm f16"))) return rc;
  HIPCHK(c, ha.ensure(2 * sizeof(_Float16) * (size_t)m * kp), "cgemm f16");
  HIPCHK(c, hb.ensure(2 * sizeof(_Float16) * (size_t)kp * n), "cgemm f16");
  HIPCHK(c, dc.ensure(sizeof(float) * nc), "cgemm f16");
  HIPCHK(c, hm::pack_a16(c->stream, da.as<float>(), m, 0, m, k, ha.as<_Float16>(), 1), "cgemm f16");
  HIPCHK(c, hm::pack_b16(c->stream, db.as<float>(), k, nullptr, 0, k, 1, n, hb.as<_Float16>(), 2LL * kp), "cgemm f16");
  HIPCHK(c, hm::cgemm_h16(c->stream, m, n, kp, ha.as<_Float16>(), 2LL * m * kp, hb.as<_Float16>(), 2LL * kp, 2LL * kp,
                          dc.as<float>(), m, (long long)m * n, 1), "cgemm f16");
  HIPCHK(c, hipMemcpyAsync(out, dc.p, sizeof(float) * nc, hipMemcpyDeviceToHost, c->stream), "cgemm f16");
  HIPCHK(c, hipStreamSynchronize(c->stream), "cgemm f16");
  return HM_OK;
}

extern "C" int hm_cgemm_3m_check(hm_ctx* c, int trans_a, int m, int n, int k, const double* a, const double* b,
                                 double* out) {
  if (!c) return HM_ERR_ARG - 1;
  if (trans_a != 0 && trans_a != 1) return c->fail(HM_ERR_ARG - 2, "trans_a must be 0 or 1");
  if (m < 1 || n < 1 || k < 1) return c->fail(HM_ERR_ARG - 3, "sizes must be >= 1");
  if (!a || !b || !out) return c->fail(HM_ERR_ARG - 6, "arrays missing");
  const size_t na = 2 * (size_t)m * k, nb = 2 * (size_t)k * n, nc = 2 * (size_t)m * n;
  DevBuf da, db, dc;
  int rc = 0;
  if ((rc = upload(c, da, a, na, "cgemm 3m")) || (rc = upload(c, db, b, nb, "cgemm 3m"))) return rc;
  HIPCHK(c, dc.ensure(sizeof(double) * nc), "cgemm 3m");
  HIPCHK(c, hm::cgemm_3m<double>(c->stream, trans_a == 1, m, n, k, 1.0, 0.0, da.as<double>(), trans_a ? k : m, 0,
                                 db.as<double>(), k, 0, 0.0, 0.0,
                                 dc.as<double>(), m, 0, 1), "cgemm 3m");
  HIPCHK(c, hipMemcpyAsync(out, dc.p, sizeof(double) * nc, hipMemcpyDeviceToHost, c->stream), "cgemm 3m");
  HIPCHK(c, hipStreamSynchronize(c->stream), "cgemm 3m");
  return HM_OK;
}

extern "C" int hm_get_block_sharing(hm_ctx* c, int* diag_blocks, int* aca_processes, double* mean_rank_run) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_geom) return c->fail(HM_ERR_STATE, "no geometry");
  const bool sh = c->sharing();
  if (sh) {
    int rc = pshare_build(c);
    if (rc) return rc;
  }
  if (diag_blocks) *diag_blocks = sh ? c->n_ublk : c->nbody;
  if (aca_processes) *aca_processes = hm::hmat_processes(c->nbody) - (sh ? c->n_pshared : 0);
  if (mean_rank_run) *mean_rank_run = c->hm_rank_mean_run;
  return HM_OK;
}

extern "C" int hm_set_block_sharing(hm_ctx* c, int enable) {
  if (!c) return HM_ERR_ARG - 1;
  if (enable < 0 || enable > 2) return c->fail(HM_ERR_ARG - 2, "enable must be 0, 1 or 2, got %d", enable);
  const bool was = c->block_share;
  c->block_share = enable != 0;
  if (c->share_sym != (enable == 2)) c->pshare_ok = false;
  c->share_sym = enable == 2;
  ++c->cfg_gen;
  c->assembled = c->factored = c->solved = false;
  if (was != c->block_share && c->hm_on && c->have_depth && c->h > 0.0) return classify(c);
  return HM_OK;
}

// classify the pairs by regime and series length (frequency independent)
static int classify(hm_ctx* c) {
  const hm::GeomDev gd = c->geom();
  const double h = c->h;
  // the H-matrix form assembles only the pairs of the diagonal blocks densely,
  // PIDCOCK pairs with their mirrors (the field points are the panel centroids)
  const int blk = c->hm_on ? c->bm() : 0;
  const bool sym = blk > 0 && c->pair_sym;
  // bodies whose hull panels equal an earlier body's share its diagonal block:
  // only the distinct bodies' blocks are classified (and assembled)
  const int* ublk = nullptr;
  int nblk = blk ? c->nbody : 0;
  if (blk && c->sharing()) {
    ublk = c->d_ublk.as<int>();
    nblk = c->n_ublk;
  }
  const long long np = blk ? (long long)nblk * blk * blk : (long long)c->nt() * c->n;
  for (int r = 0; r < 4; ++r) c->seg_off[r] = c->seg_cnt[r] = 0;
  c->stage_begin(kClassify);
  HIPCHK(c, c->d_hist.ensure(sizeof(unsigned) * hm::kNumKeys), "classify");
  HIPCHK(c, c->d_cursor.ensure(sizeof(unsigned long long) * hm::kNumKeys), "classify");
  HIPCHK(c, c->d_pairs.ensure(sizeof(uint32_t) * (size_t)np), "classify");
  HIPCHK(c, hipMemsetAsync(c->d_hist.p, 0, sizeof(unsigned) * hm::kNumKeys, c->stream), "classify");
  HIPCHK(c, hm::classify_pairs(c->stream, gd, h, c->d_hist.as<unsigned>(), nullptr, nullptr, true, blk, sym, ublk, nblk),
         "classify");
  std::vector<unsigned> hist(hm::kNumKeys);
  HIPCHK(c, hipMemcpyAsync(hist.data(), c->d_hist.p, sizeof(unsigned) * hm::kNumKeys, hipMemcpyDeviceToHost, c->stream), "classify");
  HIPCHK(c, hipStreamSynchronize(c->stream), "classify");
  // pairs assembled per key (a stored PIDCOCK pair is assembled with its mirror)
  c->hist_host.assign(hist.begin(), hist.end());
  if (sym)
    for (int k = hm::kPidcock * hm::kKeyStride; k < (hm::kEigene + 1) * hm::kKeyStride; ++k) c->hist_host[k] *= 2;
  std::vector<unsigned long long> cur(hm::kNumKeys);
  unsigned long long acc = 0;
  for (int k = 0; k < hm::kNumKeys; ++k) {
    cur[k] = acc;
    acc += hist[k];
  }
  // with mirrors, every stored PIDCOCK / EIGENE pair (i < j) stands for two
  const long long mir = (long long)(cur[(hm::kEigene + 1) * hm::kKeyStride] - cur[hm::kPidcock * hm::kKeyStride]);
  if ((long long)acc + (sym ? mir : 0) != np)
    return c->fail(HM_ERR_STATE, "pair classification lost pairs (%llu of %lld)", acc, np);
  for (int r = 0; r < 4; ++r) {
    c->seg_off[r] = (long long)cur[r * hm::kKeyStride];
    c->seg_cnt[r] = (long long)((r + 1 < 4 ? cur[(r + 1) * hm::kKeyStride] : acc) - cur[r * hm::kKeyStride]);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_cursor.p, cur.data(), sizeof(unsigned long long) * hm::kNumKeys, hipMemcpyHostToDevice, c->stream), "classify");
  HIPCHK(c, hm::classify_pairs(c->stream, gd, h, nullptr, c->d_cursor.as<unsigned long long>(), c->d_pairs.as<uint32_t>(),
                               false, blk, sym, ublk, nblk), "classify");
  // mixed path (mirrored pairs): cache the i pi / h halves of the PIDCOCK terms
  c->pcache_on = false;
  if (sym && c->seg_cnt[hm::kPidcock] > 0) {
    std::vector<long long> tab(2 * hm::kKeyStride, 0);
    long long rec = 0;
    for (int len = 0; len < hm::kKeyStride; ++len) {
      const int k = hm::kPidcock * hm::kKeyStride + len;
      tab[2 * len] = (long long)cur[k] - c->seg_off[hm::kPidcock];
      tab[2 * len + 1] = rec;
      rec += (long long)hist[k] * len;
    }
    const size_t bytes = 32 * (size_t)rec;
    if (c->d_pcache.ensure(bytes) == hipSuccess && c->d_ptab.ensure(sizeof(long long) * tab.size()) == hipSuccess) {
      HIPCHK(c, hipMemcpyAsync(c->d_ptab.p, tab.data(), sizeof(long long) * tab.size(), hipMemcpyHostToDevice,
                               c->stream), "classify");
      c->pcache_on = true;
      HIPCHK(c, hm::pidcock_cache_build(c->stream, gd, h, c->d_pairs.as<uint32_t>() + c->seg_off[hm::kPidcock],
                                        c->seg_cnt[hm::kPidcock], c->pcache()), "classify");
    } else {
      (void)hipGetLastError();              // not enough device memory: evaluate the terms in full
      c->d_pcache.release();
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "classify");
  return c->stage_end(kClassify);
}

extern "C" int hm_set_pair_symmetry(hm_ctx* c, int enable) {
  if (!c) return HM_ERR_ARG - 1;
  const bool was = c->pair_sym;
  c->pair_sym = enable != 0;
  c->assembled = c->factored = c->solved = false;
  if (was != c->pair_sym && c->hm_on && c->have_depth && c->h > 0.0) return classify(c);
  return HM_OK;
}

extern "C" int hm_set_depth(hm_ctx* c, double h) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_geom) return c->fail(HM_ERR_STATE, "hm_set_depth before hm_set_geometry");
  if (h == 0.0) return c->fail(HM_ERR_ARG - 2, "depth must be > 0 (finite) or < 0 (infinite)");
  c->h = h;
  ++c->cfg_gen;
  c->have_depth = true;
  c->have_freq = false;
  c->assembled = c->factored = c->solved = false;
  for (int r = 0; r < 4; ++r) c->seg_off[r] = c->seg_cnt[r] = 0;
  c->hist_host.assign(hm::kNumKeys, 0);
  if (h < 0.0) return HM_OK;
  return classify(c);
}

extern "C" int hm_set_headings(hm_ctx* c, int nbeta, const double* beta) {
  if (!c) return HM_ERR_ARG - 1;
  if (nbeta < 0 || (nbeta > 0 && !beta)) return c->fail(HM_ERR_ARG - 2, "bad headings");
  c->beta.assign(beta, beta + nbeta);
  ++c->cfg_gen;
  if (nbeta) {
    int rc = upload(c, c->d_beta, beta, nbeta, "headings");
    if (rc) return rc;
  } else {
    HIPCHK(c, c->d_beta.ensure(8), "headings");
  }
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_set_frequency(hm_ctx* c, double w1) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_depth) return c->fail(HM_ERR_STATE, "hm_set_frequency before hm_set_depth");
  if (!(w1 > 0.0)) return c->fail(HM_ERR_ARG - 2, "omega must be > 0");
  c->w1 = w1;
  c->v = w1 * w1 / c->g;
  if (c->h > 0.0) {
    hm_dispersion(c->wvn, kNk, w1, c->h, c->g_disp);
    c->wk = c->wvn[0];
  } else {
    c->wk = c->v;
    for (int k = 0; k < kNk; ++k) c->wvn[k] = 0.0;
  }
  int rc = upload(c, c->d_wvn, c->wvn, kNk, "frequency");
  if (rc) return rc;
  c->have_freq = true;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

// zero (which = 0) or infinite (which = 1) wave frequency: CalWaveProperts with
// SYBO = 1 (ImplementSubs.f90:45-70) sets V = WK = W1 = 0 or -1 and skips the
// dispersion roots, which the limits of the Green functions do not read
extern "C" int hm_set_frequency_limit(hm_ctx* c, int which) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_depth) return c->fail(HM_ERR_STATE, "hm_set_frequency_limit before hm_set_depth");
  if (which != 0 && which != 1) return c->fail(HM_ERR_ARG - 2, "which must be 0 (zero) or 1 (infinite frequency)");
  const double val = which ? -1.0 : 0.0;
  c->w1 = c->v = c->wk = val;
  for (int k = 0; k < kNk; ++k) c->wvn[k] = 0.0;
  int rc = upload(c, c->d_wvn, c->wvn, kNk, "frequency");
  if (rc) return rc;
  c->have_freq = true;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

// ---------------------------------------------------------------------------
// per-frequency chain
// ---------------------------------------------------------------------------
// the classified pairs (or the pair grid) into A and S (leading dimension ld);
// blk > 0: the diagonal blocks of blk x blk pairs only
static int assemble_pairs(hm_ctx* c, double* A, double* S, long long ld, int blk) {
  const hm::GeomDev gd = c->geom();
  const hm::FreqDev f = c->freq();
  const bool limit = fabs(c->v) < F32(1.e-8) || fabs(c->v + 1.0) < F32(1.e-8);
  if (c->h < 0.0 && !limit && blk == 0 && c->surd_on && c->surd_ready) {
    c->stage_begin(kAsmSurrogate);
    HIPCHK(c, hm::sur_assemble_deep(c->stream, gd, f, c->surd_dev(), (float)c->surd_tol, A, S, ld,
                                    c->d_surdcnt.as<unsigned long long>()), "assemble surrogate (deep)");
    c->stage_end(kAsmSurrogate);
  } else if (c->h < 0.0 || limit) {
    const bool share = blk > 0 && c->sharing();
    HIPCHK(c, hm::assemble_grid(c->stream, gd, f, c->h < 0.0, A, S, ld, blk, share ? c->d_ublk.as<int>() : nullptr,
                                share ? c->n_ublk : 0), "assemble");
  } else {
    const uint32_t* pairs = c->d_pairs.as<uint32_t>();
    const bool sur = c->sur_on && c->sur_ready;
    // the pair lists hold PIDCOCK and EIGENE pairs with their mirrors (classify)
    const bool sym = blk > 0 && c->pair_sym;
    for (int r = 0; r < (sur ? 2 : 4); ++r) {
      if (c->seg_cnt[r] == 0) continue;
      c->stage_begin(kAsmRegime0 + r);
      if (sym && (r == hm::kPidcock || r == hm::kEigene))
        HIPCHK(c, hm::assemble_sym(c->stream, r, gd, f, pairs + c->seg_off[r], c->seg_cnt[r], A, S, ld, nullptr,
                                   nullptr, r == hm::kPidcock ? c->pcache() : hm::PidCache{nullptr, nullptr}),
               "assemble");
      else
        HIPCHK(c, hm::assemble_regime(c->stream, r, gd, f, pairs + c->seg_off[r], c->seg_cnt[r], A, S, ld,
                                      c->series_variant), "assemble");
      c->stage_end(kAsmRegime0 + r);
    }
    if (sur) {
      // EIGENE segment: the network takes the pairs inside its validated cells
      // and marks them done; the plain series kernel runs over the segment with
      // that mask. EIGEN (R/h >= 0.5, at most 9 terms) stays plain.
      const long long ne = c->seg_cnt[2];
      if (ne > 0) {
        HIPCHK(c, c->d_fb.ensure(hm::sur_regular_bytes(ne, sym)), "surrogate alloc");
        const double kh = c->w1 * c->w1 * c->h / c->g_disp;
        hm::SurPartition part{};
        c->stage_begin(kAsmSurrogate);
        HIPCHK(c, hm::sur_regular_parts(c->stream, gd, f, kh, pairs + c->seg_off[2], ne, c->sur_dev(),
                                        (float)c->sur_tol, c->d_fb.p, c->d_fbcnt.as<unsigned long long>(), part, sym),
               "assemble surrogate");
        c->stage_end(kAsmSurrogate);
        c->stage_begin(kAsmRegime0 + 2);
        if (sym)
          HIPCHK(c, hm::assemble_sym(c->stream, 2, gd, f, part.pairs, ne, A, S, ld, part.mask, part.vreg), "assemble");
        else
          HIPCHK(c, hm::assemble_regime(c->stream, 2, gd, f, part.pairs, ne, A, S, ld, c->series_variant, part.mask,
                                        part.vreg), "assemble");
        c->stage_end(kAsmRegime0 + 2);
      }
      if (c->seg_cnt[3] > 0) {
        c->stage_begin(kAsmRegime0 + 3);
        HIPCHK(c, hm::assemble_regime(c->stream, 3, gd, f, pairs + c->seg_off[3], c->seg_cnt[3], A, S, ld,
                                      c->series_variant), "assemble");
        c->stage_end(kAsmRegime0 + 3);
      }
      c->sur_counts[1] += sym ? 2 * ne : ne;     // pairs assembled (mirrors included)
      c->sur_counts[2] += c->seg_cnt[3];
    }
  }
  return HM_OK;
}

static int precond_early(hm_ctx* c);

// the cross approximation's process map (hm::HmatDev::share): process (mat,
// b1, b2) takes the block of the first process of the same matrix whose bodies
// have the same representatives and whose placements differ by the same
// vector, computed as the entry kernels compute it (o_b1 - o_b2): its entries
// are the same bit for bit
// the horizontal maps of the square's symmetry group, (x, y) -> (a x + b y,
// c x + d y): identity, rotations by 90, 180, 270 degrees, reflections in the
// x axis, the y axis and the two diagonals
static const int kSymMap[8][4] = {{1, 0, 0, 1},  {0, -1, 1, 0}, {-1, 0, 0, -1}, {0, 1, -1, 0},
                                  {1, 0, 0, -1}, {-1, 0, 0, 1}, {0, 1, 1, 0},   {0, -1, -1, 0}};

// which maps of kSymMap take the hull panels of body 0 (local frame, the
// vertical axis through the body's origin) onto themselves: every panel onto
// one with the mapped centroid and normal and the same area and vertex count,
// to 1e-9 of the body's size. Only for one body class (every body equal).
static void find_symmetries(const hm_ctx* c, std::vector<int>& sym_ok, std::vector<int>& sym_iperm) {
  sym_ok.assign(8, 0);
  sym_iperm.clear();
  if (!c->frames || c->n_ublk != 1) return;
  const int j0 = c->body_start[0], m = c->body_start[1] - j0;
  if (m < 1) return;
  auto cen = [&](int j, int k) { return c->fp_host[3 * (size_t)(j0 + j) + k]; };
  auto nrm = [&](int j, int k) { return c->dxyz_host[6 * (size_t)(j0 + j) + k]; };
  double L = 0.0;
  for (int j = 0; j < m; ++j)
    for (int k = 0; k < 3; ++k) L = std::max(L, fabs(cen(j, k)));
  const double tol = 1e-9 * (L > 0.0 ? L : 1.0);
  // panels sorted by centroid x: the candidates of a mapped centroid are a range
  std::vector<int> ord(m);
  for (int j = 0; j < m; ++j) ord[j] = j;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return cen(a, 0) < cen(b, 0); });
  std::vector<double> xs(m);
  for (int j = 0; j < m; ++j) xs[j] = cen(ord[j], 0);
  sym_iperm.assign(8 * (size_t)m, -1);
  for (int j = 0; j < m; ++j) sym_iperm[j] = j;
  sym_ok[0] = 1;
  for (int sidx = 1; sidx < 8; ++sidx) {
    const int* q = kSymMap[sidx];
    int* ip = sym_iperm.data() + (size_t)sidx * m;
    bool ok = true;
    for (int i = 0; i < m && ok; ++i) {
      const double tx = q[0] * cen(i, 0) + q[1] * cen(i, 1), ty = q[2] * cen(i, 0) + q[3] * cen(i, 1);
      const double nx = q[0] * nrm(i, 0) + q[1] * nrm(i, 1), ny = q[2] * nrm(i, 0) + q[3] * nrm(i, 1);
      int hit = -1;
      for (auto it = std::lower_bound(xs.begin(), xs.end(), tx - tol); it != xs.end() && *it <= tx + tol; ++it) {
        const int k = ord[it - xs.begin()];
        if (fabs(cen(k, 1) - ty) <= tol && fabs(cen(k, 2) - cen(i, 2)) <= tol && fabs(nrm(k, 0) - nx) <= 1e-9 &&
            fabs(nrm(k, 1) - ny) <= 1e-9 && fabs(nrm(k, 2) - nrm(i, 2)) <= 1e-9 &&
            fabs(c->ds_host[j0 + k] - c->ds_host[j0 + i]) <= 1e-9 * c->ds_host[j0 + i] &&
            c->ncn_host[j0 + k] == c->ncn_host[j0 + i]) {
          hit = k;
          break;
        }
      }
      if (hit < 0 || ip[hit] >= 0) ok = false;
      else ip[hit] = i;
    }
    sym_ok[sidx] = ok ? 1 : 0;
  }
}

static void body_symmetries(hm_ctx* c) {
  if (c->share_sym) find_symmetries(c, c->sym_ok, c->sym_iperm);
  else {
    c->sym_ok.assign(8, 0);
    c->sym_iperm.clear();
  }
}

extern "C" int hm_get_body_symmetries(hm_ctx* c, int* ok) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_geom) return c->fail(HM_ERR_STATE, "no geometry");
  if (!ok) return c->fail(HM_ERR_ARG - 2, "ok is NULL");
  std::vector<int> so, ip;
  find_symmetries(c, so, ip);
  for (int k = 0; k < 8; ++k) ok[k] = so[k];
  return HM_OK;
}

static int pshare_build(hm_ctx* c) {
  if (c->pshare_ok) return HM_OK;
  const int nb = c->nbody, per = nb * (nb - 1), P = hm::hmat_processes(nb);
  body_symmetries(c);
  std::vector<int> sh(P), sy(P, 0);
  c->n_pshared = 0;
  int nsym = 0;
  double span = 0.0;
  for (int k = 0; k < 2 * nb; ++k) span = std::max(span, fabs(c->origin_host[k]));
  const double otol = 1e-12 * (span > 1.0 ? span : 1.0);
  for (int p = 0; p < P; ++p) {
    const int mat = p / per, r = p % per, b1 = r / (nb - 1), b2p = r % (nb - 1), b2 = b2p + (b2p >= b1 ? 1 : 0);
    const double dx = c->origin_host[2 * b1] - c->origin_host[2 * b2];
    const double dy = c->origin_host[2 * b1 + 1] - c->origin_host[2 * b2 + 1];
    sh[p] = p;
    for (int q = mat * per; q < p && sh[p] == p; ++q) {
      if (sh[q] != q) continue;
      const int rq = q % per, c1 = rq / (nb - 1), c2p = rq % (nb - 1), c2 = c2p + (c2p >= c1 ? 1 : 0);
      if (c->body_rep[c1] != c->body_rep[b1] || c->body_rep[c2] != c->body_rep[b2]) continue;
      const double ex = c->origin_host[2 * c1] - c->origin_host[2 * c2];
      const double ey = c->origin_host[2 * c1 + 1] - c->origin_host[2 * c2 + 1];
      if (memcmp(&ex, &dx, sizeof(double)) == 0 && memcmp(&ey, &dy, sizeof(double)) == 0) {
        sh[p] = q;
        break;
      }
      // a symmetry of the body that maps q's offset onto p's: p's block is
      // q's with rows and columns permuted by the symmetry's panel map
      for (int sidx = 1; sidx < (int)c->sym_ok.size(); ++sidx) {
        if (!c->sym_ok[sidx]) continue;
        const int* m4 = kSymMap[sidx];
        if (fabs(m4[0] * ex + m4[1] * ey - dx) <= otol && fabs(m4[2] * ex + m4[3] * ey - dy) <= otol) {
          sh[p] = q;
          sy[p] = sidx;
          ++nsym;
          break;
        }
      }
    }
    if (sh[p] != p) ++c->n_pshared;
  }
  c->pshare_host = sh;
  c->psym_host = nsym ? sy : std::vector<int>();
  std::vector<int> run;
  for (int p = 0; p < P; ++p)
    if (sh[p] == p) run.push_back(p);
  int rc = upload(c, c->d_pshare, sh.data(), sh.size(), "cross approximation map");
  if (rc) return rc;
  if ((rc = upload(c, c->d_prun, run.data(), run.size(), "cross approximation map"))) return rc;
  if (nsym) {
    if ((rc = upload(c, c->d_psym, sy.data(), sy.size(), "cross approximation map"))) return rc;
    if ((rc = upload(c, c->d_sperm, c->sym_iperm.data(), c->sym_iperm.size(), "cross approximation map"))) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "cross approximation map");
  c->pshare_ok = true;
  return HM_OK;
}

// matrix mat's factors packed at their ranks (rank: every process's rank)
static int pack_lowrank(hm_ctx* c, hipStream_t s, int mat, const std::vector<int>& rank) {
  const int nb = c->hmat.nb, m = c->hmat.m, per = nb * (nb - 1);
  hm::LowRankPlan& pl = c->lr_plan[mat];
  hm::lowrank_plan(nb, rank.data() + (size_t)mat * per, pl);
  const size_t cplx = 2 * sizeof(double);
  HIPCHK(c, c->d_hUc[mat].ensure(cplx * (size_t)nb * m * std::max(pl.ru, 1)), "hmatrix alloc");
  HIPCHK(c, c->d_hVc[mat].ensure(cplx * (size_t)nb * m * std::max(pl.rv, 1)), "hmatrix alloc");
  HIPCHK(c, c->d_lrmeta[mat].ensure(sizeof(int) * hm::lowrank_meta_ints(nb, pl.ru)), "hmatrix alloc");
  HIPCHK(c, hm::hmat_pack(s, c->hmat, mat, pl, c->d_lrmeta[mat].as<int>(), c->d_hUc[mat].as<double>(),
                          c->d_hVc[mat].as<double>(), &c->hm_lr[mat]), "hmatrix pack");
  return HM_OK;
}

// stride of the H-matrix form's diagonal blocks for hmat_apply: 0 when every
// body shares one block (the copies are bitwise the representative's)
static long long diag_stride(const hm_ctx* c, long long s) { return c->sharing() && c->n_ublk == 1 ? 0 : s; }

// complex entries of the packed U and V factors of matrix mat
static size_t lowrank_u_elems(const hm_ctx* c, int mat) { return (size_t)c->hmat.nb * c->hmat.m * std::max(c->hm_lr[mat].ru, 1); }
static size_t lowrank_v_elems(const hm_ctx* c, int mat) { return (size_t)c->hmat.nb * c->hmat.m * std::max(c->hm_lr[mat].rv, 1); }

// H-matrix diagonal blocks shared by bodies with equal hull panels: body b's
// block (elements [b stride, b stride + count) of `base`, `esize` bytes each)
// becomes a copy of its representative's
static int copy_shared_blocks(hm_ctx* c, hipStream_t s, void* base, long long stride, long long count, size_t esize) {
  if (!c->sharing()) return HM_OK;
  char* p = static_cast<char*>(base);
  for (int b = 0; b < c->nbody; ++b) {
    const int r = c->body_rep[b];
    if (r == b) continue;
    HIPCHK(c, hipMemcpyAsync(p + esize * (size_t)(b * stride), p + esize * (size_t)(r * stride), esize * (size_t)count,
                             hipMemcpyDeviceToDevice, s), "shared diagonal blocks");
  }
  return HM_OK;
}

// FP32 LU of the H-matrix form's preconditioner blocks (d_P32, one per body,
// pivots in d_pipiv): with shared diagonal blocks each distinct body's block is
// factored once and its factors, pivots and info word copied to the bodies
// that share it
static int factor_body_blocks(hm_ctx* c, hipStream_t s, int* info, unsigned* counter) {
  const int nb = c->nbody, m = c->bm();
  const long long sc = (long long)m * m;
  float* P = c->d_P32.as<float>();
  int* piv = c->d_pipiv.as<int>();
  if (!c->sharing()) {
    HIPCHK(c, hm::lu_factor<float>(s, P, m, sc, m, piv, info, counter, nb), "block preconditioner");
    return HM_OK;
  }
  for (int b = 0; b < nb; ++b)
    if (c->body_rep[b] == b)
      HIPCHK(c, hm::lu_factor<float>(s, P + 2 * b * sc, m, sc, m, piv + (size_t)b * m, info + b, counter + b, 1),
             "block preconditioner");
  int rc;
  if ((rc = copy_shared_blocks(c, s, P, sc, sc, 2 * sizeof(float))) ||
      (rc = copy_shared_blocks(c, s, piv, m, m, sizeof(int))) || (rc = copy_shared_blocks(c, s, info, 1, 1, sizeof(int))))
    return rc;
  return HM_OK;
}

// H-matrix form: diagonal blocks of A and S densely (the classified pairs of
// those blocks), the off-diagonal blocks by cross approximation, and the
// right-hand sides S W through the H-matrix product in FP64
static int assemble_hmatrix(hm_ctx* c) {
  const int n = c->n, nb = c->nbody, m = c->bm(), ncol = c->ncol(), kmax = c->hm_kmax;
  const size_t cplx = 2 * sizeof(double);
  const size_t tsz = cplx * (size_t)nb * (nb - 1) * kmax * ncol;
  HIPCHK(c, c->d_hD.ensure(cplx * c->hm_diag_elems()), "hmatrix alloc");
  HIPCHK(c, c->d_hDS.ensure(cplx * c->hm_diag_elems()), "hmatrix alloc");
  HIPCHK(c, c->d_hmat.ensure(hm::hmat_bytes(nb, m, kmax)), "hmatrix alloc");
  HIPCHK(c, c->d_htb.ensure(tsz), "hmatrix alloc");
  HIPCHK(c, c->d_hts.ensure(tsz), "hmatrix alloc");
  HIPCHK(c, c->d_W.ensure(cplx * (size_t)n * ncol), "assemble alloc W");
  HIPCHK(c, c->d_R.ensure(cplx * (size_t)n * ncol), "assemble alloc RHS");
  c->hmat = hm::hmat_carve(c->d_hmat.p, nb, m, kmax, c->hm_eps, c->hm_eps_s);
  if (c->sharing()) {
    int rc0 = pshare_build(c);
    if (rc0) return rc0;
    if (c->n_pshared > 0) {
      c->hmat.share = c->d_pshare.as<int>();
      c->hmat.run = c->d_prun.as<int>();
      c->hmat.nrun = hm::hmat_processes(nb) - c->n_pshared;
      if (!c->psym_host.empty()) {
        c->hmat.psym = c->d_psym.as<int>();
        c->hmat.sperm = c->d_sperm.as<int>();
      }
    }
  }
  const hm::GeomDev gd = c->geom();
  const hm::FreqDev f = c->freq();
  double* D = c->d_hD.as<double>();
  double* DS = c->d_hDS.as<double>();
  c->join_side();
  c->prec_early = false;
  c->stage_begin(kAssembly);
  // overlap (the default): the off-diagonal blocks by cross approximation on the
  // side stream (from the state the main stream has reached), the diagonal
  // blocks and the block preconditioner on the main stream: everything on the
  // main stream is enqueued before the cross approximation's host loop starts.
  // Without overlap the cross approximation follows on the main stream.
  hipStream_t acs = c->hm_overlap ? c->side : c->stream;
  if (c->hm_overlap) {
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream), "hmatrix");
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0), "hmatrix");
    c->side_pending = true;
  }
  int rc = assemble_pairs(c, D, DS, m, m);
  if (rc) return rc;
  // no PIDCOCK / LINTON entry between bodies whose bounding circles are >= 0.05 h apart
  c->hmat.narrow = c->h > 0.0 && c->min_body_gap > 0.05 * c->h * (1.0 + 1e-9) ? 1 : 0;
  c->stage_begin(kAsmCross, acs);
  // each entry-kernel launch of the cross approximation timed on its stream
  c->aca_stream = acs;
  hm::AcaTimer tm{c, [](void* u, int b) {
                    hm_ctx* cc = static_cast<hm_ctx*>(u);
                    if (b) cc->stage_begin(kAcaEntries, cc->aca_stream);
                    else cc->stage_end(kAcaEntries, cc->aca_stream);
                  }};
  // with the overlap, the steps a rank-kmax block needs are enqueued on the side
  // stream before the main stream's copies and preconditioner, so that the cross
  // approximation (the longer of the two) starts with the diagonal blocks; the
  // host then polls for processes still open
  int steps = 0;
  // finite depth: the evanescent modes' depth factors of every point and the
  // modes' normalisations, once per frequency, for the entries' direct sums
  hm::FreqDev fa = f;
  const bool limit = fabs(c->v) < F32(1.e-8) || fabs(c->v + 1.0) < F32(1.e-8);
  if (c->h > 0.0 && !limit) {
    HIPCHK(c, c->d_evt.ensure(2 * sizeof(double) * (size_t)kNk * c->nt()), "hmatrix alloc");
    HIPCHK(c, c->d_pmt.ensure(sizeof(double) * kNk), "hmatrix alloc");
    HIPCHK(c, hm::evan_table(acs, gd, f, c->d_evt.as<double>(), c->d_pmt.as<double>()), "cross approximation");
    fa.evt = c->d_evt.as<double>();
    fa.pmt = c->d_pmt.as<double>();
    fa.evt_ld = c->nt();
  }
  HIPCHK(c, hm::aca_begin(acs, gd, fa, c->hmat, c->hm_overlap ? kmax + 4 : 0, &steps, c->timing ? &tm : nullptr),
         "cross approximation");
  if ((rc = copy_shared_blocks(c, c->stream, D, m * (m + 1), m * m, 2 * sizeof(double))) ||
      (rc = copy_shared_blocks(c, c->stream, DS, m * (m + 1), m * m, 2 * sizeof(double))))
    return rc;
  if ((rc = precond_early(c))) return rc;
  HIPCHK(c, hm::aca_end(acs, gd, fa, c->hmat, &steps, &c->hm_unconverged, c->timing ? &tm : nullptr),
         "cross approximation");
  c->hm_steps = steps;
  c->stage_end(kAsmCross, acs);
  c->join_side();
  if ((rc = c->stage_end(kAssembly))) return rc;
  const int P = hm::hmat_processes(nb);
  std::vector<int> rank(P);
  HIPCHK(c, hipMemcpyAsync(rank.data(), c->hmat.rank, sizeof(int) * P, hipMemcpyDeviceToHost, c->stream), "hmatrix");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hmatrix");
  long long rs = 0;
  c->hm_rank_max = 0;
  for (int r : rank) { rs += r; c->hm_rank_max = r > c->hm_rank_max ? r : c->hm_rank_max; }
  c->hm_rank_mean = P ? (double)rs / P : 0.0;
  {
    const bool sh = c->hmat.share != nullptr && (int)c->pshare_host.size() == P;
    long long rr = 0, cnt = 0;
    for (int p = 0; p < P; ++p)
      if (!sh || c->pshare_host[p] == p) { rr += rank[p]; ++cnt; }
    c->hm_rank_mean_run = cnt ? (double)rr / cnt : 0.0;
  }
  // pack each matrix's factors at their ranks, so that the products run over
  // the ranks the blocks have instead of kmax
  for (int mat = 0; mat < 2; ++mat) {
    int ku = 0;
    for (int p = mat * P / 2; p < (mat + 1) * P / 2; ++p) ku = rank[p] > ku ? rank[p] : ku;
    c->hm_kuse[mat] = ku;
    int rc1 = pack_lowrank(c, c->stream, mat, rank);
    if (rc1) return rc1;
  }
  // right-hand sides RHS = S * W
  c->stage_begin(kRhs);
  HIPCHK(c, hm::build_w(c->stream, gd, f, c->d_beta.as<double>(), (int)c->beta.size(), c->xw, c->yw, c->d_W.as<double>(), n), "rhs");
  const bool rhs_sparse = diag_stride(c, 1) == 0 && ncol == 6 * nb + (int)c->beta.size();
  if (rhs_sparse) {
    // the diagonal blocks' product over the columns that are nonzero on each body
    const size_t ce = (size_t)m * nb * (6 + c->beta.size());
    HIPCHK(c, c->d_rhsx.ensure(cplx * ce), "rhs alloc");
    HIPCHK(c, c->d_rhsy.ensure(cplx * ce), "rhs alloc");
    HIPCHK(c, hm::hmat_diag_rhs(c->stream, nb, m, (int)c->beta.size(), DS, m, c->d_W.as<double>(), n,
                                c->d_R.as<double>(), n, c->d_rhsx.as<double>(), c->d_rhsy.as<double>()),
           "rhs diagonal blocks");
  }
  HIPCHK(c, hm::hmat_apply<double>(c->stream, nb, m, c->hm_lr[1], rhs_sparse ? nullptr : DS, m,
                                   diag_stride(c, (long long)m * (m + 1)),
                                   c->d_hUc[1].as<double>(), c->d_hVc[1].as<double>(), c->d_W.as<double>(), n,
                                   c->d_R.as<double>(), n, ncol, c->d_htb.as<double>(), c->d_hts.as<double>()),
         "rhs hmatrix product");
  if ((rc = c->stage_end(kRhs))) return rc;
  c->assembled = true;
  c->factored = c->solved = false;
  return HM_OK;
}

// precond 2 and -1 (the default): the preconditioner blocks applied as their
// explicit inverses (one batched GEMM per Krylov step); precond 1: by their
// triangular solves
static bool block_inverse(const hm_ctx* c) { return c->gm_precond == 2 || c->gm_precond < 0; }

// after the factorisation of the preconditioner blocks (c->gm_blocks), for
// block_inverse(c): their explicit FP32 inverses without the row interchanges
// (U^-1 L^-1) and their pivot permutations
static int precond_prepare(hm_ctx* c, hipStream_t s) {
  c->pinv_ready = false;
  c->pinv16_ready = false;
  if (c->gm_blocks.size() < 2 || !block_inverse(c)) return HM_OK;
  const int nblk = (int)c->gm_blocks.size() - 1;
  size_t tot = 0;
  for (int b = 0; b < nblk; ++b) {
    const size_t nb = (size_t)(c->gm_blocks[b + 1] - c->gm_blocks[b]);
    tot += nb * nb;
  }
  HIPCHK(c, c->d_pperm.ensure(sizeof(int) * (size_t)c->gm_blocks[nblk]), "preconditioner alloc");
  HIPCHK(c, c->d_Pinv.ensure(2 * sizeof(float) * tot), "preconditioner alloc");
  HIPCHK(c, hm::gmres_block_perm(s, nblk, c->gm_blocks.data(), c->d_pipiv.as<int>(), c->d_pperm.as<int>()),
         "preconditioner permutation");
  if (c->sharing() && c->gm_blocks == c->body_start) {
    // the H-matrix form's body blocks: one inverse per distinct body, copied
    const int m = c->bm();
    const long long sc = (long long)m * m;
    const int one[2] = {0, m};
    for (int b = 0; b < nblk; ++b)
      if (c->body_rep[b] == b)
        HIPCHK(c, hm::gmres_block_inverse(s, 1, one, c->d_P32.as<float>() + 2 * b * sc, c->d_Pinv.as<float>() + 2 * b * sc),
               "preconditioner inverse");
    int rc = copy_shared_blocks(c, s, c->d_Pinv.p, sc, sc, 2 * sizeof(float));
    if (rc) return rc;
  } else {
    HIPCHK(c, hm::gmres_block_inverse(s, nblk, c->gm_blocks.data(), c->d_P32.as<float>(), c->d_Pinv.as<float>()),
           "preconditioner inverse");
  }
  c->pinv_ready = true;
  return HM_OK;
}

// the H-matrix form's preconditioner as explicit inverses (precond 2 and -1):
// the FP32 copy of each distinct body's diagonal block inverted in place in
// d_Pinv by Gauss-Jordan elimination without row interchanges (the blocks are
// 2 pi I plus a compact part: no pivoting is needed, and D_b M_b^-1 is the
// identity to FP32 rounding as the Krylov steps assume), copied to the bodies
// that share it; the permutations are the identity
static int precond_gj(hm_ctx* c, hipStream_t s, int* info) {
  const int nb = c->nbody, m = c->bm();
  const long long sD = (long long)m * (m + 1), sc = (long long)m * m;
  const double* D = c->d_hD.as<double>();
  float* Pinv = c->d_Pinv.as<float>();
  const bool share = c->sharing();
  const int nw = share ? 1 : nb;
  HIPCHK(c, c->d_gjws.ensure(hm::gj_inverse_ws_bytes(m) * (size_t)nw), "preconditioner alloc");
  if (share) {
    for (int b = 0; b < nb; ++b) {
      if (c->body_rep[b] != b) continue;
      HIPCHK(c, hm::demote(s, D + 2 * b * sD, m, sD, Pinv + 2 * b * sc, m, sc, m, m, 1), "demote");
      HIPCHK(c, hm::gj_inverse_nopiv(s, Pinv + 2 * b * sc, m, sc, m, c->d_gjws.p, info + b, 1), "preconditioner inverse");
    }
    int rc = copy_shared_blocks(c, s, Pinv, sc, sc, 2 * sizeof(float));
    if (rc) return rc;
  } else {
    HIPCHK(c, hm::demote(s, D, m, sD, Pinv, m, sc, m, m, nb), "demote");
    HIPCHK(c, hm::gj_inverse_nopiv(s, Pinv, m, sc, m, c->d_gjws.p, info, nb), "preconditioner inverse");
  }
  HIPCHK(c, hm::fill_iota(s, c->d_pperm.as<int>(), c->n), "preconditioner permutation");
  c->pinv16_ready = false;
  if (c->pinv16_on) {
    // packed FP16 copies for the Arnoldi steps: one if every body shares one block
    const int np = share && c->n_ublk == 1 ? 1 : nb;
    HIPCHK(c, c->d_Pinv16.ensure(2 * sizeof(_Float16) * (size_t)np * m * hm::h16_kp(m)), "preconditioner alloc");
    HIPCHK(c, hm::pack_a16(s, Pinv, m, sc, m, m, c->d_Pinv16.as<_Float16>(), np), "preconditioner FP16 copy");
    c->pinv16_shared = np == 1 && nb > 1;
    c->pinv16_ready = true;
  }
  c->gm_blocks = c->body_start;
  c->pinv_ready = true;
  return HM_OK;
}

// H-matrix form, right after the diagonal blocks are assembled: their FP32
// copies, the block preconditioner's LU and (precond 2, -1) its inverses on the
// main stream, overlapping the cross approximation on the side stream (own info
// words and counters: hm_factor clears the main ones). Buffers are sized
// first, so that no allocation happens while the side stream runs.
static int precond_early(hm_ctx* c) {
  if (c->gm_precond == 0) return HM_OK;
  const int nb = c->nbody, m = c->bm();
  const long long sD = (long long)m * (m + 1), sc = (long long)m * m;
  HIPCHK(c, c->d_hD32.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
  HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
  HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)c->n), "factor alloc");
  HIPCHK(c, c->d_info2.ensure(sizeof(int) * (size_t)nb), "factor alloc");
  HIPCHK(c, c->d_counter2.ensure(sizeof(unsigned) * (size_t)nb), "factor alloc");
  if (block_inverse(c)) {
    HIPCHK(c, c->d_pperm.ensure(sizeof(int) * (size_t)c->n), "factor alloc");
    HIPCHK(c, c->d_Pinv.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
  }
  const double* D = c->d_hD.as<double>();
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemsetAsync(c->d_info2.p, 0, sizeof(int) * nb, s), "preconditioner");
  HIPCHK(c, hipMemsetAsync(c->d_counter2.p, 0, sizeof(unsigned) * nb, s), "preconditioner");
  HIPCHK(c, hm::demote(s, D, m, sD, c->d_hD32.as<float>(), m, sc, m, m, nb), "demote");
  int rc;
  if (block_inverse(c)) {
    if ((rc = precond_gj(c, s, c->d_info2.as<int>()))) return rc;
    c->prec_early = true;
    return HM_OK;
  }
  HIPCHK(c, hm::demote(s, D, m, sD, c->d_P32.as<float>(), m, sc, m, m, nb), "demote");
  if ((rc = factor_body_blocks(c, s, c->d_info2.as<int>(), c->d_counter2.as<unsigned>()))) return rc;
  c->gm_blocks = c->body_start;
  if ((rc = precond_prepare(c, s))) return rc;
  c->prec_early = true;
  return HM_OK;
}

// H-matrix form of hm_factor: FP32 copies of the diagonal blocks and of A's
// low-rank factors for the Krylov products, and the block-diagonal
// preconditioner (one block per body) factored in FP32
static int factor_hmatrix(hm_ctx* c) {
  const int nb = c->nbody, m = c->bm();
  const long long sD = (long long)m * (m + 1), sc = (long long)m * m;
  const size_t lu = lowrank_u_elems(c, 0), lv = lowrank_v_elems(c, 0);   // complex entries of A's packed U, V
  c->join_side();
  if (c->prec_early) {
    // the diagonal blocks were demoted, factored and inverted during the
    // assembly, beside the cross approximation: their info words, then A's factors
    c->prec_early = false;
    HIPCHK(c, hipMemcpyAsync(c->d_info.p, c->d_info2.p, sizeof(int) * nb, hipMemcpyDeviceToDevice, c->stream), "factor");
    HIPCHK(c, c->d_hU32.ensure(2 * sizeof(float) * lu), "factor alloc");
    HIPCHK(c, c->d_hV32.ensure(2 * sizeof(float) * lv), "factor alloc");
    HIPCHK(c, hm::demote(c->stream, c->d_hUc[0].as<double>(), m, 0, c->d_hU32.as<float>(), m, 0, m, (int)(lu / m), 1),
           "demote");
    HIPCHK(c, hm::demote(c->stream, c->d_hVc[0].as<double>(), m, 0, c->d_hV32.as<float>(), m, 0, m, (int)(lv / m), 1),
           "demote");
    return HM_OK;
  }
  HIPCHK(c, c->d_hD32.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
  HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
  HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)c->n), "factor alloc");
  HIPCHK(c, c->d_hU32.ensure(2 * sizeof(float) * lu), "factor alloc");
  HIPCHK(c, c->d_hV32.ensure(2 * sizeof(float) * lv), "factor alloc");
  const double* D = c->d_hD.as<double>();
  HIPCHK(c, hm::demote(c->stream, D, m, sD, c->d_hD32.as<float>(), m, sc, m, m, nb), "demote");
  HIPCHK(c, hm::demote(c->stream, D, m, sD, c->d_P32.as<float>(), m, sc, m, m, nb), "demote");
  HIPCHK(c, hm::demote(c->stream, c->d_hUc[0].as<double>(), m, 0, c->d_hU32.as<float>(), m, 0, m, (int)(lu / m), 1),
         "demote");
  HIPCHK(c, hm::demote(c->stream, c->d_hVc[0].as<double>(), m, 0, c->d_hV32.as<float>(), m, 0, m, (int)(lv / m), 1),
         "demote");
  if (c->gm_precond != 0 && block_inverse(c)) {
    HIPCHK(c, c->d_pperm.ensure(sizeof(int) * (size_t)c->n), "factor alloc");
    HIPCHK(c, c->d_Pinv.ensure(2 * sizeof(float) * (size_t)nb * sc), "factor alloc");
    HIPCHK(c, hipMemsetAsync(c->d_info.p, 0, sizeof(int) * nb, c->stream), "factor");
    return precond_gj(c, c->stream, c->d_info.as<int>());
  }
  int rc = factor_body_blocks(c, c->stream, c->d_info.as<int>(), c->d_counter.as<unsigned>());
  if (rc) return rc;
  c->gm_blocks = c->body_start;
  if (c->gm_precond == 0) c->gm_blocks.clear();
  return precond_prepare(c, c->stream);
}

extern "C" int hm_assemble(hm_ctx* c) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_freq) return c->fail(HM_ERR_STATE, "hm_assemble before hm_set_frequency");
  if (c->hm_on) return assemble_hmatrix(c);
  const int n = c->n, nt = c->nt(), ncol = c->ncol();
  const size_t cplx = 2 * sizeof(double);
  HIPCHK(c, c->d_A.ensure(cplx * (size_t)nt * n), "assemble alloc A");
  HIPCHK(c, c->d_S.ensure(cplx * (size_t)nt * n), "assemble alloc S");
  HIPCHK(c, c->d_W.ensure(cplx * (size_t)n * ncol), "assemble alloc W");
  HIPCHK(c, c->d_R.ensure(cplx * (size_t)nt * ncol), "assemble alloc RHS");
  const hm::GeomDev gd = c->geom();
  const hm::FreqDev f = c->freq();
  double* A = c->d_A.as<double>();
  double* S = c->d_S.as<double>();
  c->stage_begin(kAssembly);
  int rc = assemble_pairs(c, A, S, nt, 0);
  if (rc) return rc;
  rc = c->stage_end(kAssembly);
  if (rc) return rc;
  // right-hand sides RHS = S * W
  c->stage_begin(kRhs);
  HIPCHK(c, hm::build_w(c->stream, gd, f, c->d_beta.as<double>(), (int)c->beta.size(), c->xw, c->yw, c->d_W.as<double>(), n), "rhs");
  HIPCHK(c, hm::cgemm<double>(c->stream, false, nt, ncol, n, 1.0, 0.0, S, nt, 0, c->d_W.as<double>(), n, 0, 0.0, 0.0,
                              c->d_R.as<double>(), nt, 0, 1), "rhs gemm");
  if ((rc = c->stage_end(kRhs))) return rc;
  if (c->ni > 0) {
    // least squares through the normal equations with plain transpose
    c->stage_begin(kIrr);
    HIPCHK(c, c->d_C.ensure(cplx * (size_t)n * n), "irr alloc");
    HIPCHK(c, c->d_D.ensure(cplx * (size_t)n * ncol), "irr alloc");
    HIPCHK(c, hm::cgemm<double>(c->stream, true, n, n, nt, 1.0, 0.0, A, nt, 0, A, nt, 0, 0.0, 0.0, c->d_C.as<double>(), n, 0, 1), "irr A^T A");
    HIPCHK(c, hm::cgemm<double>(c->stream, true, n, ncol, nt, 1.0, 0.0, A, nt, 0, c->d_R.as<double>(), nt, 0, 0.0, 0.0, c->d_D.as<double>(), n, 0, 1), "irr A^T B");
    if ((rc = c->stage_end(kIrr))) return rc;
  }
  c->assembled = true;
  c->factored = c->solved = false;
  return HM_OK;
}

// the n x n system matrix and n x ncol right-hand side the solve works on
static double* sys_matrix(hm_ctx* c) { return c->ni > 0 ? c->d_C.as<double>() : c->d_A.as<double>(); }
static double* sys_rhs(hm_ctx* c) { return c->ni > 0 ? c->d_D.as<double>() : c->d_R.as<double>(); }
static int sys_ld_rhs(hm_ctx* c) { return c->ni > 0 ? c->n : c->nt(); }

// the cluster-tree H-matrix of the n x n matrix A (device, leading dimension
// lda) over the points (x[i], y[i], z[i])
static int build_tree(hm_ctx* c, const double* A, long long lda, int n, const std::vector<double>& x,
                      const std::vector<double>& y, const std::vector<double>& z) {
  hm::HTreeHost host;
  hm::htree_partition(n, x.data(), y.data(), z.data(), c->ht_leaf, c->ht_eta, host);
  c->ht_ready = false;
  HIPCHK(c, c->htree.build(c->stream, host, A, lda, c->hm_eps, c->hm_kmax), "cluster-tree H-matrix");
  const hm::HTreeStats& st = c->htree.stats();
  c->hm_unconverged = st.fallback_dense;
  c->hm_steps = st.steps;
  c->hm_rank_max = st.max_rank;
  c->hm_rank_mean = st.mean_rank;
  c->ht_ready = true;
  return HM_OK;
}

// GMRES products through the cluster-tree H-matrix (FP32 for the Krylov steps,
// FP64 for the residuals); the workspace is sized for ncol columns
static int tree_ops(hm_ctx* c, hm::GmresArgs& ga, int n, int ncol) {
  HIPCHK(c, c->d_htw.ensure(c->htree.workspace_bytes(ncol, sizeof(double))), "tree product alloc");
  const hm::HTree* t = &c->htree;
  void* w = c->d_htw.p;
  ga.op32 = [=](hipStream_t st, const float* Z, float* W) { return t->apply<float>(st, Z, n, W, n, ncol, w); };
  ga.op64 = [=](hipStream_t st, const double* Z, double* W) { return t->apply<double>(st, Z, n, W, n, ncol, w); };
  return HM_OK;
}

extern "C" int hm_factor(hm_ctx* c, int precision) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->assembled) return c->fail(HM_ERR_STATE, "hm_factor before hm_assemble");
  c->join_side();
  if (precision != HM_FP64 && precision != HM_FP32_IR && precision != HM_GMRES)
    return c->fail(HM_ERR_ARG - 2, "unknown precision %d", precision);
  const int n = c->n;
  HIPCHK(c, c->d_ipiv.ensure(sizeof(int) * (size_t)n), "factor alloc");
  // one info word and one counter per matrix of a batched factorisation (the
  // block preconditioner factors the bodies' blocks as a batch)
  const size_t nmat = (size_t)(c->nbody > 1 ? c->nbody : 1);
  HIPCHK(c, c->d_info.ensure(sizeof(int) * nmat), "factor alloc");
  HIPCHK(c, c->d_counter.ensure(sizeof(unsigned) * nmat), "factor alloc");
  HIPCHK(c, hipMemsetAsync(c->d_info.p, 0, sizeof(int) * nmat, c->stream), "factor");
  HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, sizeof(unsigned) * nmat, c->stream), "factor");
  // the factorisation overwrites the preconditioner an earlier H-matrix build set up
  c->hd_ready = false;
  c->ht_from_build = false;
  if (c->hm_tree && precision != HM_GMRES) return c->fail(HM_ERR_ARG - 2, "the H-matrix form solves with HM_GMRES only");
  c->stage_begin(kLu);
  if (c->hm_on) {
    if (precision != HM_GMRES) return c->fail(HM_ERR_ARG - 2, "the H-matrix form solves with HM_GMRES only");
    int rc = factor_hmatrix(c);
    if (rc) return rc;
  } else if (precision == HM_FP64) {
    HIPCHK(c, c->d_luws.ensure(hm::lu_coop_ws_bytes(n)), "factor alloc");
    HIPCHK(c, hm::lu_factor<double>(c->stream, sys_matrix(c), n, 0, n, c->d_ipiv.as<int>(), c->d_info.as<int>(),
                                    c->d_counter.as<unsigned>(), 1, c->luws()), "lu fp64");
  } else if (precision == HM_FP32_IR) {
    HIPCHK(c, c->d_A32.ensure(2 * sizeof(float) * (size_t)n * n), "factor alloc");
    HIPCHK(c, hm::demote(c->stream, sys_matrix(c), n, 0, c->d_A32.as<float>(), n, 0, n, n, 1), "demote");
    HIPCHK(c, c->d_luws.ensure(hm::lu_coop_ws_bytes(n)), "factor alloc");
    HIPCHK(c, hm::lu_factor<float>(c->stream, c->d_A32.as<float>(), n, 0, n, c->d_ipiv.as<int>(), c->d_info.as<int>(),
                                   c->d_counter.as<unsigned>(), 1, c->luws()), "lu fp32");
  } else {
    // GMRES: FP32 copy for the Krylov products (the cluster-tree form keeps its
    // own FP32 factors); block-diagonal preconditioner (one block per body) by
    // default for several bodies, none for one body
    if (!c->hm_tree) {
      HIPCHK(c, c->d_A32.ensure(2 * sizeof(float) * (size_t)n * n), "factor alloc");
      HIPCHK(c, hm::demote(c->stream, sys_matrix(c), n, 0, c->d_A32.as<float>(), n, 0, n, n, 1), "demote");
    }
    const bool blocks = c->gm_precond < 0 ? c->nbody > 1 : c->gm_precond >= 1;
    c->gm_blocks.clear();
    if (blocks) {
      c->gm_blocks = c->body_start;
      size_t tot = 0;
      for (int b = 0; b < c->nbody; ++b) {
        const size_t nb = (size_t)(c->body_start[b + 1] - c->body_start[b]);
        tot += nb * nb;
      }
      HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * tot), "factor alloc");
      HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)n), "factor alloc");
      HIPCHK(c, hm::gmres_block_precond_setup(c->stream, sys_matrix(c), n, c->nbody, c->gm_blocks.data(),
                                              c->d_P32.as<float>(), c->d_pipiv.as<int>(), c->d_info.as<int>(),
                                              c->d_counter.as<unsigned>()), "block preconditioner");
    }
    if (c->hm_tree) {
      // the H-matrix of the system matrix (A, or C = A^T A with waterplane rows)
      // over the hull panels' centroids
      std::vector<double> x(n), y(n), z(n);
      for (int i = 0; i < n; ++i) {
        x[i] = c->xyz_host[3 * (size_t)i];
        y[i] = c->xyz_host[3 * (size_t)i + 1];
        z[i] = c->xyz_host[3 * (size_t)i + 2];
      }
      int rc = build_tree(c, sys_matrix(c), n, n, x, y, z);
      if (rc != HM_OK) return rc;
    }
    int rc = precond_prepare(c, c->stream);
    if (rc != HM_OK) return rc;
  }
  std::vector<int> infos(nmat, 0);
  HIPCHK(c, hipMemcpyAsync(infos.data(), c->d_info.p, sizeof(int) * nmat, hipMemcpyDeviceToHost, c->stream), "factor");
  HIPCHK(c, hipStreamSynchronize(c->stream), "factor");
  int info = infos[0];
  if (precision == HM_GMRES && !c->gm_blocks.empty())
    for (size_t b = 0; b < nmat && !info; ++b)
      if (infos[b] > 0) info = c->body_start[b] + infos[b];   // zero pivot of a preconditioner block
  int rc = c->stage_end(kLu);
  if (rc) return rc;
  c->precision = precision;
  c->factored = true;
  c->solved = false;
  if (info > 0) return c->fail(info, "zero pivot in column %d", info);
  return HM_OK;
}

// per-column max(|re|,|im|) of an n x ncol block
template <typename T>
__global__ void k_colmax_t(const T* X, int ldx, int n, double* out) {
  __shared__ double s[256];
  const int c = blockIdx.x;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const long long off = (long long)i + (long long)c * ldx;
    m = fmax(m, fmax(fabs((double)X[2 * off]), fabs((double)X[2 * off + 1])));
  }
  s[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (!threadIdx.x) out[c] = s[0];
}
#define k_colmax k_colmax_t<double>
#define k_colmax_f k_colmax_t<float>

constexpr int kMaxRefine = 30;
constexpr double kRefineTol = 1e-13;

extern "C" int hm_solve(hm_ctx* c, int* iters) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->factored) return c->fail(HM_ERR_STATE, "hm_solve before hm_factor");
  const int n = c->n, ncol = c->ncol();
  const size_t cplx = 2 * sizeof(double);
  HIPCHK(c, c->d_X.ensure(cplx * (size_t)n * ncol), "solve alloc");
  double* X = c->d_X.as<double>();
  const double* B = sys_rhs(c);
  const int ldb = sys_ld_rhs(c);
  int it = 0;
  c->stage_begin(kSolve);
  HIPCHK(c, hipMemcpy2DAsync(X, cplx * n, B, cplx * ldb, cplx * n, ncol, hipMemcpyDeviceToDevice, c->stream), "solve");
  if (c->precision == HM_FP64) {
    HIPCHK(c, hm::lu_solve<double>(c->stream, sys_matrix(c), n, 0, n, c->d_ipiv.as<int>(), X, n, 0, ncol, 1), "getrs fp64");
  } else if (c->precision == HM_GMRES) {
    hm::GmresArgs ga;
    ga.n = n;
    ga.ncol = ncol;
    ga.restart = c->gm_restart;
    ga.maxiter = c->gm_maxiter;
    ga.tol = c->gm_tol;
    ga.inner_tol = 0.5 * c->gm_tol;
    if (c->hm_on) {
      const int nb = c->nbody, m = c->bm();
      const hm::LowRank k = c->hm_lr[0];
      const long long sD = diag_stride(c, (long long)m * (m + 1)), sc = diag_stride(c, (long long)m * m);
      const float* D32 = c->d_hD32.as<float>();
      const float* U32 = c->d_hU32.as<float>();
      const float* V32 = c->d_hV32.as<float>();
      const double* D = c->d_hD.as<double>();
      const double* U = c->d_hUc[0].as<double>();
      const double* V = c->d_hVc[0].as<double>();
      double* tb = c->d_htb.as<double>();
      double* ts = c->d_hts.as<double>();
      ga.op32 = [=](hipStream_t s, const float* Z, float* W) {
        return hm::hmat_apply<float>(s, nb, m, k, D32, m, sc, U32, V32, Z, n, W, n, ncol, (float*)tb, (float*)ts);
      };
      // the preconditioner blocks are the FP32 LU of D32's blocks (factor_hmatrix)
      ga.op32_off = [=](hipStream_t s, const float* Z, float* W) {
        return hm::hmat_apply<float>(s, nb, m, k, nullptr, m, sc, U32, V32, Z, n, W, n, ncol, (float*)tb, (float*)ts);
      };
      ga.op64 = [=](hipStream_t s, const double* Z, double* W) {
        return hm::hmat_apply<double>(s, nb, m, k, D, m, sD, U, V, Z, n, W, n, ncol, tb, ts);
      };
    } else if (c->hm_tree) {
      if (!c->ht_ready) return c->fail(HM_ERR_STATE, "the cluster-tree H-matrix was released; factor again");
      int rc = tree_ops(c, ga, n, ncol);
      if (rc != HM_OK) return rc;
    } else {
      ga.A64 = sys_matrix(c);
      ga.lda64 = n;
      ga.A32 = c->d_A32.as<float>();
    }
    ga.B = B;
    ga.ldb = ldb;
    ga.X = X;
    if (!c->gm_blocks.empty()) {
      ga.nblk = (int)c->gm_blocks.size() - 1;
      ga.blk_start = c->gm_blocks.data();
      ga.P = c->d_P32.as<float>();
      ga.pipiv = c->d_pipiv.as<int>();
      if (c->pinv_ready && block_inverse(c)) {
        ga.pperm = c->d_pperm.as<int>();
        ga.Pinv = c->d_Pinv.as<float>();
        if (c->hm_on && c->pinv16_on && c->pinv16_ready) {
          ga.Pinv16 = c->d_Pinv16.as<_Float16>();
          ga.pinv16_shared = c->pinv16_shared;
          ga.cycle_rtol = c->gm_cycle_rtol;
        }
        // the H-matrix form's one shared block: its inverse copied to every body
        ga.pinv_shared = c->hm_on && c->gm_blocks == c->body_start && c->sharing() && c->n_ublk == 1 &&
                         c->nbody > 1;
      }
    }
    HIPCHK(c, c->d_gwork.ensure(hm::gmres_workspace_bytes(n, ncol, c->gm_restart, ga.Pinv16 ? ga.nblk : 0)),
           "gmres alloc");
    HIPCHK(c, hm::gmres_solve(c->stream, ga, c->d_gwork.p, &c->gm_stats), "gmres");
    it = c->gm_stats.iterations;
  } else {
    // FP32 solves of the FP64 residual; stop when the last correction is below
    // 1e-13 of the solution in every column
    HIPCHK(c, c->d_X32.ensure(2 * sizeof(float) * (size_t)n * ncol), "solve alloc");
    HIPCHK(c, c->d_Rw.ensure(cplx * (size_t)n * ncol), "solve alloc");
    HIPCHK(c, c->d_norm.ensure(sizeof(double) * 2 * ncol), "solve alloc");
    float* X32 = c->d_X32.as<float>();
    double* Rw = c->d_Rw.as<double>();
    double* nrm = c->d_norm.as<double>();
    std::vector<double> hn(2 * (size_t)ncol);
    const double* A = sys_matrix(c);
    for (it = 0; it < kMaxRefine;) {
      const double* res = it == 0 ? X : Rw;       // X holds B before the first pass
      HIPCHK(c, hm::demote(c->stream, res, n, 0, X32, n, 0, n, ncol, 1), "ir demote");
      HIPCHK(c, hm::lu_solve<float>(c->stream, c->d_A32.as<float>(), n, 0, n, c->d_ipiv.as<int>(), X32, n, 0, ncol, 1), "getrs fp32");
      HIPCHK(c, hm::promote_add(c->stream, X32, n, 0, X, n, 0, n, ncol, 1, it == 0), "ir update");
      ++it;
      hipLaunchKernelGGL(k_colmax_f, dim3(ncol), dim3(256), 0, c->stream, X32, n, n, nrm + ncol);
      hipLaunchKernelGGL(k_colmax, dim3(ncol), dim3(256), 0, c->stream, X, n, n, nrm);
      HIPCHK(c, hipGetLastError(), "ir norm");
      HIPCHK(c, hipMemcpyAsync(hn.data(), nrm, sizeof(double) * 2 * ncol, hipMemcpyDeviceToHost, c->stream), "ir norm");
      HIPCHK(c, hipStreamSynchronize(c->stream), "ir norm");
      double worst = 0.0;
      for (int k = 0; k < ncol; ++k) worst = fmax(worst, hn[ncol + k] / (hn[k] > 0.0 ? hn[k] : 1.0));
      if (it > 1 && worst < kRefineTol) break;
      // residual R = B - A X in FP64
      HIPCHK(c, hipMemcpy2DAsync(Rw, cplx * n, B, cplx * ldb, cplx * n, ncol, hipMemcpyDeviceToDevice, c->stream), "ir residual");
      HIPCHK(c, hm::cgemm<double>(c->stream, false, n, ncol, n, -1.0, 0.0, A, n, 0, X, n, 0, 1.0, 0.0, Rw, n, 0, 1), "ir residual");
    }
  }
  int rc = c->stage_end(kSolve);
  if (rc) return rc;
  if (iters) *iters = it;
  c->solved = true;
  return HM_OK;
}

extern "C" int hm_set_gmres(hm_ctx* c, int restart, double tol, int maxiter, int precond) {
  if (!c) return HM_ERR_ARG - 1;
  if (restart < 1 || restart > hm::kGmresMaxRestart)
    return c->fail(HM_ERR_ARG - 2, "restart must be in 1..%d", hm::kGmresMaxRestart);
  if (!(tol > 0.0 && tol < 1.0)) return c->fail(HM_ERR_ARG - 3, "tol must be in (0, 1)");
  if (maxiter < 1) return c->fail(HM_ERR_ARG - 4, "maxiter must be >= 1");
  if (precond < -1 || precond > 2) return c->fail(HM_ERR_ARG - 5, "precond must be -1, 0, 1 or 2");
  c->gm_restart = restart;
  c->gm_tol = tol;
  c->gm_maxiter = maxiter;
  c->gm_precond = precond;
  c->factored = c->solved = false;
  c->pinv_ready = false;
  c->pinv16_ready = false;
  c->prec_early = false;
  return HM_OK;
}

extern "C" int hm_set_hmatrix_overlap(hm_ctx* c, int enable) {
  if (!c) return HM_ERR_ARG - 1;
  c->join_side();
  c->hm_overlap = enable != 0;
  c->prec_early = false;
  return HM_OK;
}

extern "C" int hm_set_hmatrix(hm_ctx* c, int enable, double eps, double eps_s, int kmax) {
  if (!c) return HM_ERR_ARG - 1;
  if (!(eps > 0.0 && eps < 1.0)) return c->fail(HM_ERR_ARG - 3, "eps must be in (0, 1)");
  if (!(eps_s > 0.0 && eps_s < 1.0)) return c->fail(HM_ERR_ARG - 4, "eps_s must be in (0, 1)");
  if (kmax < 1) return c->fail(HM_ERR_ARG - 5, "kmax must be >= 1");
  bool tree = false;
  if (enable) {
    if (!c->have_geom) return c->fail(HM_ERR_STATE, "hm_set_hmatrix before hm_set_geometry");
    // body clusters (assembly-aware cross approximation) for two or more bodies
    // of equal panel counts without waterplane points; the cluster tree over the
    // assembled system otherwise, or when hm_set_cluster_tree asks for it
    const int m = c->n / c->nbody;
    bool bodies = c->ht_mode == 0 && c->nbody >= 2 && c->ni == 0 && kmax <= m;
    for (int b = 0; bodies && b <= c->nbody; ++b) bodies = c->body_start[b] == b * m;
    tree = !bodies;
  }
  c->join_side();
  // the tolerances and rank cap also serve hm_hmatrix_build and hm_hmatrix_build_tree
  c->hm_eps = eps;
  c->hm_eps_s = eps_s;
  c->hm_kmax = kmax;
  const bool was = c->hm_on;
  c->hm_on = enable != 0 && !tree;
  c->hm_tree = enable != 0 && tree;
  if (c->hm_tree || !c->ht_from_build) {
    c->ht_ready = false;
    c->ht_from_build = false;
  }
  c->assembled = c->factored = c->solved = false;
  c->hm_unconverged = c->hm_steps = c->hm_rank_max = 0;
  c->hm_rank_mean = 0.0;
  if (was != c->hm_on && c->have_depth && c->h > 0.0) return classify(c);
  return HM_OK;
}

extern "C" int hm_set_cluster_tree(hm_ctx* c, int mode, int leaf, double eta) {
  if (!c) return HM_ERR_ARG - 1;
  if (mode != 0 && mode != 1) return c->fail(HM_ERR_ARG - 2, "mode must be 0 (body clusters where they apply) or 1 (tree)");
  if (leaf < 2 || leaf > 64) return c->fail(HM_ERR_ARG - 3, "leaf must be in 2..64");
  if (!(eta > 0.0)) return c->fail(HM_ERR_ARG - 4, "eta must be > 0");
  c->ht_mode = mode;
  c->ht_leaf = leaf;
  c->ht_eta = eta;
  return HM_OK;
}

extern "C" int hm_get_cluster_tree_stats(hm_ctx* c, int* clusters, int* leaves, int* dense_blocks, int* lowrank_blocks,
                                         int* dense_fallbacks) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->ht_ready) return c->fail(HM_ERR_STATE, "no cluster-tree H-matrix is built");
  const hm::HTreeStats& st = c->htree.stats();
  if (clusters) *clusters = st.clusters;
  if (leaves) *leaves = st.leaves;
  if (dense_blocks) *dense_blocks = st.dense_blocks;
  if (lowrank_blocks) *lowrank_blocks = st.lowrank_blocks;
  if (dense_fallbacks) *dense_fallbacks = st.fallback_dense;
  return HM_OK;
}

extern "C" int hm_get_hmatrix_stats(hm_ctx* c, int* unconverged, int* steps, double* mean_rank, int* max_rank,
                                    double* stored_fraction) {
  if (!c) return HM_ERR_ARG - 1;
  if (unconverged) *unconverged = c->hm_unconverged;
  if (steps) *steps = c->hm_steps;
  if (mean_rank) *mean_rank = c->hm_rank_mean;
  if (max_rank) *max_rank = c->hm_rank_max;
  if (stored_fraction) {
    // entries held per matrix (dense diagonal blocks + U and V at the actual
    // ranks) over the n^2 of the dense matrix
    const double m = c->nbody ? (double)c->n / c->nbody : 0.0;
    const double off = (double)c->nbody * (c->nbody - 1) * 2.0 * m * c->hm_rank_mean;
    *stored_fraction = c->hm_on && c->n ? (c->nbody * m * m + off) / ((double)c->n * c->n) : 1.0;
    if (!c->hm_on && !c->hd_ready && c->ht_ready) *stored_fraction = c->htree.stats().stored_fraction;
  }
  return HM_OK;
}

extern "C" int hm_hmatrix_apply(hm_ctx* c, int mat, int ncol, const double* x, double* y) {
  if (!c) return HM_ERR_ARG - 1;
  const bool from_assembly = c->hm_on && c->assembled;
  if (!from_assembly && !c->hd_ready && c->ht_ready) {
    // the cluster-tree H-matrix (of A, or of C = A^T A with waterplane rows)
    if (mat != 0) return c->fail(HM_ERR_ARG - 2, "the cluster-tree H-matrix has only mat 0");
    if (ncol < 1) return c->fail(HM_ERR_ARG - 3, "bad ncol %d", ncol);
    if (!x || !y) return c->fail(HM_ERR_ARG - 4, "x or y is NULL");
    const int n = c->htree.n();
    const size_t bytes = 2 * sizeof(double) * (size_t)n * ncol;
    DevBuf dx, dy;
    int rc = HM_OK;
    if ((rc = c->hip(dx.ensure(bytes), "hmatrix apply")) == HM_OK && (rc = c->hip(dy.ensure(bytes), "hmatrix apply")) == HM_OK &&
        (rc = c->hip(c->d_htw.ensure(c->htree.workspace_bytes(ncol, sizeof(double))), "hmatrix apply")) == HM_OK &&
        (rc = c->hip(hipMemcpyAsync(dx.p, x, bytes, hipMemcpyHostToDevice, c->stream), "hmatrix apply")) == HM_OK &&
        (rc = c->hip(c->htree.apply<double>(c->stream, dx.as<double>(), n, dy.as<double>(), n, ncol, c->d_htw.p),
                     "hmatrix apply")) == HM_OK &&
        (rc = c->hip(hipMemcpyAsync(y, dy.p, bytes, hipMemcpyDeviceToHost, c->stream), "hmatrix apply")) == HM_OK)
      rc = c->hip(hipStreamSynchronize(c->stream), "hmatrix apply");
    dx.release();
    dy.release();
    return rc;
  }
  if (!from_assembly && !c->hd_ready)
    return c->fail(HM_ERR_STATE, "hm_hmatrix_apply needs an assembled H-matrix form or one built by hm_hmatrix_build(_tree)");
  if (mat != 0 && mat != 1) return c->fail(HM_ERR_ARG - 2, "mat must be 0 (A) or 1 (S)");
  if (!from_assembly && mat != 0) return c->fail(HM_ERR_ARG - 2, "an H-matrix built from a matrix has only mat 0");
  if (ncol < 1 || (from_assembly && ncol > c->ncol())) return c->fail(HM_ERR_ARG - 3, "bad ncol %d", ncol);
  if (!x || !y) return c->fail(HM_ERR_ARG - 4, "x or y is NULL");
  const int nb = from_assembly ? c->nbody : c->hd_nb, m = from_assembly ? c->bm() : c->hd_m, n = nb * m;
  const size_t tneed = 2 * sizeof(double) * (size_t)nb * std::max(1, std::max(c->hm_lr[mat].ru, c->hm_lr[mat].rv)) * ncol;
  if (c->d_htb.bytes < tneed || c->d_hts.bytes < tneed) {
    int e = c->hip(c->d_htb.ensure(tneed), "hmatrix apply");
    if (e == HM_OK) e = c->hip(c->d_hts.ensure(tneed), "hmatrix apply");
    if (e) return e;
  }
  const size_t bytes = 2 * sizeof(double) * (size_t)n * ncol;
  DevBuf dx, dy;
  int rc = HM_OK;
  if ((rc = c->hip(dx.ensure(bytes), "hmatrix apply")) == HM_OK && (rc = c->hip(dy.ensure(bytes), "hmatrix apply")) == HM_OK &&
      (rc = c->hip(hipMemcpyAsync(dx.p, x, bytes, hipMemcpyHostToDevice, c->stream), "hmatrix apply")) == HM_OK &&
      (rc = c->hip(hm::hmat_apply<double>(c->stream, nb, m, c->hm_lr[mat],
                                          mat ? c->d_hDS.as<double>() : c->d_hD.as<double>(), m,
                                          from_assembly ? diag_stride(c, (long long)m * (m + 1)) : (long long)m * (m + 1),
                                          c->d_hUc[mat].as<double>(), c->d_hVc[mat].as<double>(), dx.as<double>(), n,
                                          dy.as<double>(), n, ncol, c->d_htb.as<double>(), c->d_hts.as<double>()),
                   "hmatrix apply")) == HM_OK &&
      (rc = c->hip(hipMemcpyAsync(y, dy.p, bytes, hipMemcpyDeviceToHost, c->stream), "hmatrix apply")) == HM_OK)
    rc = c->hip(hipStreamSynchronize(c->stream), "hmatrix apply");
  dx.release();
  dy.release();
  return rc;
}

// H-matrix of an assembled host matrix (HMATRIX_BUILD(N, AMAT, XYZ_P),
// HMatrix.f90:135): clusters = the nblocks equal blocks given, diagonal blocks
// dense, the rest by cross approximation of the matrix's entries; the FP32
// copies and the block-diagonal preconditioner for GMRES are set up with it
extern "C" int hm_hmatrix_build(hm_ctx* c, int n, const double* a, int lda, int nblocks, const int* block_sizes) {
  if (!c) return HM_ERR_ARG - 1;
  c->join_side();
  c->prec_early = false;
  if (n < 1) return c->fail(HM_ERR_ARG - 2, "n must be >= 1");
  if (!a || lda < n) return c->fail(HM_ERR_ARG - 3, "bad matrix");
  if (nblocks < 2 || !block_sizes) return c->fail(HM_ERR_ARG - 5, "the H-matrix needs >= 2 cluster blocks");
  const int m = block_sizes[0];
  for (int k = 0; k < nblocks; ++k)
    if (block_sizes[k] != m) return c->fail(HM_ERR_ARG - 6, "the H-matrix needs cluster blocks of equal size");
  if ((long long)m * nblocks != n) return c->fail(HM_ERR_ARG - 6, "block sizes sum to %lld, not n = %d", (long long)m * nblocks, n);
  if (c->hm_kmax > m) return c->fail(HM_ERR_STATE, "kmax %d exceeds the block size %d", c->hm_kmax, m);
  // the context's matrix buffers are reused: any assembled state is gone
  c->assembled = c->factored = c->solved = false;
  c->hd_ready = false;
  c->ht_ready = c->ht_from_build = false;
  const size_t cplx = 2 * sizeof(double);
  hipStream_t s = c->stream;
  const int nb = nblocks, kmax = c->hm_kmax;
  const long long sD = (long long)m * (m + 1), sc = (long long)m * m;
  HIPCHK(c, c->d_A.ensure(cplx * (size_t)n * n), "hmatrix alloc");
  HIPCHK(c, hipMemcpy2DAsync(c->d_A.p, cplx * n, a, cplx * lda, cplx * n, n, hipMemcpyHostToDevice, s), "hmatrix upload");
  HIPCHK(c, c->d_info.ensure(sizeof(int) * nb), "hmatrix alloc");
  HIPCHK(c, c->d_counter.ensure(sizeof(unsigned) * nb), "hmatrix alloc");
  HIPCHK(c, hipMemsetAsync(c->d_info.p, 0, sizeof(int) * nb, s), "hmatrix");
  HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, sizeof(unsigned) * nb, s), "hmatrix");
  HIPCHK(c, c->d_hmat.ensure(hm::hmat_bytes(nb, m, kmax, 1)), "hmatrix alloc");
  HIPCHK(c, c->d_hD.ensure(cplx * (size_t)nb * sD), "hmatrix alloc");
  c->hmat = hm::hmat_carve(c->d_hmat.p, nb, m, kmax, c->hm_eps, c->hm_eps, 1, c->d_A.as<double>(), n);
  for (int k = 0; k < nb; ++k)
    HIPCHK(c, hipMemcpy2DAsync(c->d_hD.as<double>() + 2 * (size_t)k * sD, cplx * m,
                               c->d_A.as<double>() + 2 * ((size_t)k * m + (size_t)k * m * n), cplx * n, cplx * m, m,
                               hipMemcpyDeviceToDevice, s), "hmatrix diagonal blocks");
  hm::GeomDev g0{};
  hm::FreqDev f0{};
  HIPCHK(c, hm::aca_build(s, g0, f0, c->hmat, &c->hm_unconverged, &c->hm_steps), "cross approximation");
  const int P = hm::hmat_processes(nb, 1);
  std::vector<int> rank(P);
  HIPCHK(c, hipMemcpyAsync(rank.data(), c->hmat.rank, sizeof(int) * P, hipMemcpyDeviceToHost, s), "hmatrix");
  HIPCHK(c, hipStreamSynchronize(s), "hmatrix");
  int ku = 0;
  long long rs = 0;
  for (int r : rank) { ku = r > ku ? r : ku; rs += r; }
  c->hm_kuse[0] = ku;
  c->hm_rank_max = ku;
  c->hm_rank_mean = P ? (double)rs / P : 0.0;
  {
    const bool sh = c->hmat.share != nullptr && (int)c->pshare_host.size() == P;
    long long rr = 0, cnt = 0;
    for (int p = 0; p < P; ++p)
      if (!sh || c->pshare_host[p] == p) { rr += rank[p]; ++cnt; }
    c->hm_rank_mean_run = cnt ? (double)rr / cnt : 0.0;
  }
  {
    int rc1 = pack_lowrank(c, s, 0, rank);
    if (rc1) return rc1;
  }
  const size_t lu = lowrank_u_elems(c, 0), lv = lowrank_v_elems(c, 0);
  HIPCHK(c, c->d_hD32.ensure(2 * sizeof(float) * (size_t)nb * sc), "hmatrix alloc");
  HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * (size_t)nb * sc), "hmatrix alloc");
  HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)n), "hmatrix alloc");
  HIPCHK(c, c->d_hU32.ensure(2 * sizeof(float) * lu), "hmatrix alloc");
  HIPCHK(c, c->d_hV32.ensure(2 * sizeof(float) * lv), "hmatrix alloc");
  const double* D = c->d_hD.as<double>();
  HIPCHK(c, hm::demote(s, D, m, sD, c->d_hD32.as<float>(), m, sc, m, m, nb), "demote");
  HIPCHK(c, hm::demote(s, D, m, sD, c->d_P32.as<float>(), m, sc, m, m, nb), "demote");
  HIPCHK(c, hm::demote(s, c->d_hUc[0].as<double>(), m, 0, c->d_hU32.as<float>(), m, 0, m, (int)(lu / m), 1), "demote");
  HIPCHK(c, hm::demote(s, c->d_hVc[0].as<double>(), m, 0, c->d_hV32.as<float>(), m, 0, m, (int)(lv / m), 1), "demote");
  HIPCHK(c, hm::lu_factor<float>(s, c->d_P32.as<float>(), m, sc, m, c->d_pipiv.as<int>(), c->d_info.as<int>(),
                                 c->d_counter.as<unsigned>(), nb), "block preconditioner");
  std::vector<int> infos(nb, 0);
  HIPCHK(c, hipMemcpyAsync(infos.data(), c->d_info.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s), "hmatrix");
  HIPCHK(c, hipStreamSynchronize(s), "hmatrix");
  for (int k = 0; k < nb; ++k)
    if (infos[k] > 0) return c->fail(k * m + infos[k], "zero pivot in diagonal block %d", k + 1);
  c->pinv_ready = false;   // new factors: their permutations (and inverses) follow at the first solve
  c->hd_nb = nb;
  c->hd_m = m;
  c->hd_ready = true;
  return HM_OK;
}

extern "C" int hm_hmatrix_release(hm_ctx* c) {
  if (!c) return HM_ERR_ARG - 1;
  c->join_side();
  c->hd_ready = false;
  if (c->ht_ready) {
    (void)hipStreamSynchronize(c->stream);
    c->htree.release();
    c->ht_ready = false;
    c->ht_from_build = false;
    if (c->hm_tree) c->factored = c->solved = false;   // the solve needs it: factor again
  }
  return HM_OK;
}

// the cluster tree and block partition of hm_hmatrix_build_tree, on the host
// only: perm (n, 0-based cluster position -> point) and 5 ints per block (rs,
// nr, cs, nc, admissible); blocks = NULL asks for the count only
extern "C" int hm_cluster_tree_partition(int n, const double* xyz_p, int ldxyz, int leaf, double eta, int* perm,
                                         int max_blocks, int* nblocks, int* blocks) {
  if (n < 1) return HM_ERR_ARG - 1;
  if (!xyz_p || ldxyz < n) return HM_ERR_ARG - 2;
  if (leaf < 2 || leaf > 64) return HM_ERR_ARG - 4;
  if (!(eta > 0.0)) return HM_ERR_ARG - 5;
  if (!perm || !nblocks) return HM_ERR_ARG - 6;
  hm::HTreeHost t;
  hm::htree_partition(n, xyz_p, xyz_p + ldxyz, xyz_p + 2 * (size_t)ldxyz, leaf, eta, t);
  std::copy(t.perm.begin(), t.perm.end(), perm);
  *nblocks = (int)t.blocks.size();
  if (!blocks) return HM_OK;
  if ((int)t.blocks.size() > max_blocks) return HM_ERR_ARG - 7;
  for (size_t k = 0; k < t.blocks.size(); ++k) {
    const hm::HTreeBlock& b = t.blocks[k];
    int* o = blocks + 5 * k;
    o[0] = b.rs; o[1] = b.nr; o[2] = b.cs; o[3] = b.nc; o[4] = b.lowrank;
  }
  return HM_OK;
}

// HMATRIX_BUILD(N, AMAT, XYZ_P) (HMatrix.f90:135) with the reference's cluster
// tree and block partition over the points XYZ_P (column-major n x 3, leading
// dimension ldxyz), the admissible blocks compressed by cross approximation of
// the matrix's entries (tolerance and rank cap of hm_set_hmatrix, leaf size
// and eta of hm_set_cluster_tree); with nblocks > 0 also the block-diagonal
// FP32 preconditioner of SETUP_BLOCK_PRECONDITIONER (IterativeSolvers.f90:190)
// over those blocks, nblocks = 0 none (the reference's single-body case)
extern "C" int hm_hmatrix_build_tree(hm_ctx* c, int n, const double* a, int lda, const double* xyz_p, int ldxyz,
                                     int nblocks, const int* block_sizes) {
  if (!c) return HM_ERR_ARG - 1;
  c->join_side();
  c->prec_early = false;
  if (n < 1) return c->fail(HM_ERR_ARG - 2, "n must be >= 1");
  if (!a || lda < n) return c->fail(HM_ERR_ARG - 3, "bad matrix");
  if (!xyz_p || ldxyz < n) return c->fail(HM_ERR_ARG - 5, "bad coordinates");
  if (nblocks < 0 || (nblocks > 0 && !block_sizes)) return c->fail(HM_ERR_ARG - 7, "bad preconditioner blocks");
  std::vector<int> start(1, 0);
  for (int k = 0; k < nblocks; ++k) {
    if (block_sizes[k] < 1) return c->fail(HM_ERR_ARG - 8, "block %d is empty", k + 1);
    start.push_back(start.back() + block_sizes[k]);
  }
  if (nblocks > 0 && start.back() != n) return c->fail(HM_ERR_ARG - 8, "block sizes sum to %d, not n = %d", start.back(), n);
  // the context's matrix buffers are reused: any assembled state is gone
  c->assembled = c->factored = c->solved = false;
  c->hd_ready = c->ht_ready = c->ht_from_build = false;
  const size_t cplx = 2 * sizeof(double);
  hipStream_t s = c->stream;
  HIPCHK(c, c->d_A.ensure(cplx * (size_t)n * n), "hmatrix alloc");
  HIPCHK(c, hipMemcpy2DAsync(c->d_A.p, cplx * n, a, cplx * lda, cplx * n, n, hipMemcpyHostToDevice, s), "hmatrix upload");
  c->ht_blocks.clear();
  if (nblocks > 0) {
    const size_t nmat = (size_t)nblocks;
    HIPCHK(c, c->d_info.ensure(sizeof(int) * nmat), "hmatrix alloc");
    HIPCHK(c, c->d_counter.ensure(sizeof(unsigned) * nmat), "hmatrix alloc");
    HIPCHK(c, hipMemsetAsync(c->d_info.p, 0, sizeof(int) * nmat, s), "hmatrix");
    HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, sizeof(unsigned) * nmat, s), "hmatrix");
    size_t tot = 0;
    for (int k = 0; k < nblocks; ++k) tot += (size_t)block_sizes[k] * block_sizes[k];
    HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * tot), "hmatrix alloc");
    HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)n), "hmatrix alloc");
    HIPCHK(c, hm::gmres_block_precond_setup(s, c->d_A.as<double>(), n, nblocks, start.data(), c->d_P32.as<float>(),
                                            c->d_pipiv.as<int>(), c->d_info.as<int>(), c->d_counter.as<unsigned>()),
           "block preconditioner");
    std::vector<int> infos(nmat, 0);
    HIPCHK(c, hipMemcpyAsync(infos.data(), c->d_info.p, sizeof(int) * nmat, hipMemcpyDeviceToHost, s), "hmatrix");
    HIPCHK(c, hipStreamSynchronize(s), "hmatrix");
    for (size_t k = 0; k < nmat; ++k)
      if (infos[k] > 0) return c->fail(start[k] + infos[k], "zero pivot in preconditioner block %zu", k + 1);
    c->ht_blocks = start;
  }
  std::vector<double> x(xyz_p, xyz_p + n), y(xyz_p + ldxyz, xyz_p + ldxyz + n), z(xyz_p + 2 * (size_t)ldxyz, xyz_p + 2 * (size_t)ldxyz + n);
  int rc = build_tree(c, c->d_A.as<double>(), n, n, x, y, z);
  if (rc != HM_OK) return rc;
  c->pinv_ready = false;   // new factors: their permutations (and inverses) follow at the first solve
  c->ht_from_build = true;
  return HM_OK;
}

// ZITER_SOLVE_MULTI(ISOLV, N, AMAT, B, X, NRHS, INFO) (IterativeSolvers.f90:574)
// on an assembled host matrix. isolv = 1: dense GMRES with the block-diagonal
// preconditioner of SETUP_BLOCK_PRECONDITIONER(N, AMAT, NBLOCKS, BLOCK_SIZES)
// (:190; nblocks = 0: none); isolv = 2: GMRES on the H-matrix of the last
// hm_hmatrix_build (as ZGMRES_SOLVE uses HMATVEC once HMATRIX_BUILD has run,
// :325-343), which also supplies the preconditioner; a and the blocks are then
// not read.
extern "C" int hm_ziter_solve(hm_ctx* c, int isolv, int n, const double* a, int lda, const double* b, int ldb,
                              double* x, int ldx, int nrhs, int nblocks, const int* block_sizes, int* info) {
  if (!c) return HM_ERR_ARG - 1;
  c->join_side();
  c->prec_early = false;
  if (isolv != 1 && isolv != 2) return c->fail(HM_ERR_ARG - 2, "isolv must be 1 (dense) or 2 (H-matrix)");
  if (n < 1) return c->fail(HM_ERR_ARG - 3, "n must be >= 1");
  if (isolv == 1 && (!a || lda < n)) return c->fail(HM_ERR_ARG - 5, "bad matrix");
  if (!b || ldb < n) return c->fail(HM_ERR_ARG - 7, "bad right-hand sides");
  if (!x || ldx < n) return c->fail(HM_ERR_ARG - 8, "bad solution array");
  if (nrhs < 1) return c->fail(HM_ERR_ARG - 10, "nrhs must be >= 1");
  const bool tree = isolv == 2 && !c->hd_ready && c->ht_from_build;
  if (isolv == 2) {
    if (!c->hd_ready && !tree) return c->fail(HM_ERR_STATE, "isolv = 2 before hm_hmatrix_build or hm_hmatrix_build_tree");
    const int hn = tree ? c->htree.n() : c->hd_nb * c->hd_m;
    if (hn != n) return c->fail(HM_ERR_ARG - 3, "n = %d differs from the H-matrix's %d", n, hn);
  }
  std::vector<int> start(1, 0);
  if (isolv == 1) {
    if (nblocks < 0 || (nblocks > 0 && !block_sizes)) return c->fail(HM_ERR_ARG - 11, "bad preconditioner blocks");
    for (int k = 0; k < nblocks; ++k) {
      if (block_sizes[k] < 1) return c->fail(HM_ERR_ARG - 12, "block %d is empty", k + 1);
      start.push_back(start.back() + block_sizes[k]);
    }
    if (nblocks > 0 && start.back() != n) return c->fail(HM_ERR_ARG - 12, "block sizes sum to %d, not n = %d", start.back(), n);
    c->hd_ready = false;
    if (c->ht_from_build) {
      c->ht_from_build = false;   // its preconditioner factors are about to be overwritten
      c->ht_ready = false;
    }
  }
  // both forms overwrite the context's right-hand sides and solution (d_R, d_X):
  // the per-frequency chain must assemble again before hm_solve / hm_forces
  c->assembled = c->factored = c->solved = false;
  const size_t cplx = 2 * sizeof(double);
  hipStream_t s = c->stream;
  HIPCHK(c, c->d_R.ensure(cplx * (size_t)n * nrhs), "ziter alloc");
  HIPCHK(c, c->d_X.ensure(cplx * (size_t)n * nrhs), "ziter alloc");
  HIPCHK(c, hipMemcpy2DAsync(c->d_R.p, cplx * n, b, cplx * ldb, cplx * n, nrhs, hipMemcpyHostToDevice, s), "ziter upload");
  hm::GmresArgs ga;
  ga.n = n;
  ga.ncol = nrhs;
  ga.restart = c->gm_restart;
  ga.maxiter = c->gm_maxiter;
  ga.tol = c->gm_tol;
  ga.inner_tol = 0.5 * c->gm_tol;
  ga.B = c->d_R.as<double>();
  ga.ldb = n;
  ga.X = c->d_X.as<double>();
  c->gm_blocks.clear();
  if (isolv == 1) {
    HIPCHK(c, c->d_A.ensure(cplx * (size_t)n * n), "ziter alloc");
    HIPCHK(c, hipMemcpy2DAsync(c->d_A.p, cplx * n, a, cplx * lda, cplx * n, n, hipMemcpyHostToDevice, s), "ziter upload");
    HIPCHK(c, c->d_A32.ensure(2 * sizeof(float) * (size_t)n * n), "ziter alloc");
    HIPCHK(c, hm::demote(s, c->d_A.as<double>(), n, 0, c->d_A32.as<float>(), n, 0, n, n, 1), "demote");
    ga.A64 = c->d_A.as<double>();
    ga.lda64 = n;
    ga.A32 = c->d_A32.as<float>();
    if (nblocks > 0) {
      const size_t nmat = (size_t)nblocks;
      HIPCHK(c, c->d_info.ensure(sizeof(int) * nmat), "ziter alloc");
      HIPCHK(c, c->d_counter.ensure(sizeof(unsigned) * nmat), "ziter alloc");
      HIPCHK(c, hipMemsetAsync(c->d_info.p, 0, sizeof(int) * nmat, s), "ziter");
      HIPCHK(c, hipMemsetAsync(c->d_counter.p, 0, sizeof(unsigned) * nmat, s), "ziter");
      c->gm_blocks = start;
      size_t tot = 0;
      for (int k = 0; k < nblocks; ++k) tot += (size_t)block_sizes[k] * block_sizes[k];
      HIPCHK(c, c->d_P32.ensure(2 * sizeof(float) * tot), "ziter alloc");
      HIPCHK(c, c->d_pipiv.ensure(sizeof(int) * (size_t)n), "ziter alloc");
      HIPCHK(c, hm::gmres_block_precond_setup(s, c->d_A.as<double>(), n, nblocks, c->gm_blocks.data(),
                                              c->d_P32.as<float>(), c->d_pipiv.as<int>(), c->d_info.as<int>(),
                                              c->d_counter.as<unsigned>()), "block preconditioner");
      std::vector<int> infos(nmat, 0);
      HIPCHK(c, hipMemcpyAsync(infos.data(), c->d_info.p, sizeof(int) * nmat, hipMemcpyDeviceToHost, s), "ziter");
      HIPCHK(c, hipStreamSynchronize(s), "ziter");
      for (size_t k = 0; k < nmat; ++k)
        if (infos[k] > 0) return c->fail(start[k] + infos[k], "zero pivot in preconditioner block %zu", k + 1);
      int rc = precond_prepare(c, s);
      if (rc != HM_OK) return rc;
    }
  } else if (tree) {
    int rc = tree_ops(c, ga, n, nrhs);
    if (rc != HM_OK) return rc;
    c->gm_blocks = c->ht_blocks;
    if (!c->gm_blocks.empty() && !c->pinv_ready) {
      rc = precond_prepare(c, s);
      if (rc != HM_OK) return rc;
    }
  } else {
    const int nb = c->hd_nb, m = c->hd_m;
    const hm::LowRank ku = c->hm_lr[0];
    const long long sD = (long long)m * (m + 1), sc = (long long)m * m;
    const size_t tsz = cplx * (size_t)nb * std::max(1, std::max(ku.ru, ku.rv)) * nrhs;
    HIPCHK(c, c->d_htb.ensure(tsz), "ziter alloc");
    HIPCHK(c, c->d_hts.ensure(tsz), "ziter alloc");
    const float* D32 = c->d_hD32.as<float>();
    const float* U32 = c->d_hU32.as<float>();
    const float* V32 = c->d_hV32.as<float>();
    const double* D = c->d_hD.as<double>();
    const double* U = c->d_hUc[0].as<double>();
    const double* V = c->d_hVc[0].as<double>();
    double* tb = c->d_htb.as<double>();
    double* ts = c->d_hts.as<double>();
    const int ncol = nrhs;
    ga.op32 = [=](hipStream_t st, const float* Z, float* W) {
      return hm::hmat_apply<float>(st, nb, m, ku, D32, m, sc, U32, V32, Z, n, W, n, ncol, (float*)tb, (float*)ts);
    };
    // the preconditioner blocks are the FP32 LU of D32's blocks (hm_hmatrix_build)
    ga.op32_off = [=](hipStream_t st, const float* Z, float* W) {
      return hm::hmat_apply<float>(st, nb, m, ku, nullptr, m, sc, U32, V32, Z, n, W, n, ncol, (float*)tb, (float*)ts);
    };
    ga.op64 = [=](hipStream_t st, const double* Z, double* W) {
      return hm::hmat_apply<double>(st, nb, m, ku, D, m, sD, U, V, Z, n, W, n, ncol, tb, ts);
    };
    for (int k = 0; k <= nb; ++k) start.push_back(k * m);
    start.erase(start.begin());
    c->gm_blocks = start;
    if (!c->pinv_ready) {
      int rc = precond_prepare(c, s);
      if (rc != HM_OK) return rc;
    }
  }
  if (!c->gm_blocks.empty()) {
    ga.nblk = (int)c->gm_blocks.size() - 1;
    ga.blk_start = c->gm_blocks.data();
    ga.P = c->d_P32.as<float>();
    ga.pipiv = c->d_pipiv.as<int>();
    if (c->pinv_ready && block_inverse(c)) {
      ga.pperm = c->d_pperm.as<int>();
      ga.Pinv = c->d_Pinv.as<float>();
    }
  }
  HIPCHK(c, c->d_gwork.ensure(hm::gmres_workspace_bytes(n, nrhs, c->gm_restart)), "gmres alloc");
  HIPCHK(c, hm::gmres_solve(s, ga, c->d_gwork.p, &c->gm_stats), "gmres");
  HIPCHK(c, hipMemcpy2DAsync(x, cplx * ldx, c->d_X.p, cplx * n, cplx * n, nrhs, hipMemcpyDeviceToHost, s), "ziter download");
  HIPCHK(c, hipStreamSynchronize(s), "ziter");
  if (info) *info = c->gm_stats.converged ? 0 : 1;
  return HM_OK;
}

extern "C" int hm_set_lu_panel(hm_ctx* c, int cooperative) {
  if (!c) return HM_ERR_ARG - 1;
  if (cooperative != 0 && cooperative != 1) return c->fail(HM_ERR_ARG - 2, "cooperative must be 0 or 1");
  c->lu_coop = cooperative != 0;
  return HM_OK;
}

extern "C" int hm_set_series_variant(hm_ctx* c, int variant) {
  if (!c) return HM_ERR_ARG - 1;
  if (variant < 0 || variant > 2) return c->fail(HM_ERR_ARG - 2, "variant must be 0, 1 or 2");
  c->series_variant = variant;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_get_gmres_stats(hm_ctx* c, int* iters, int* cycles, int* converged, double* rel_residual) {
  if (!c) return HM_ERR_ARG - 1;
  if (iters) *iters = c->gm_stats.iterations;
  if (cycles) *cycles = c->gm_stats.cycles;
  if (converged) *converged = c->gm_stats.converged;
  if (rel_residual) *rel_residual = c->gm_stats.rel_residual;
  return HM_OK;
}

// ---------------------------------------------------------------------------
// learned surrogate of the evanescent sums (hm_surrogate.hpp)
// ---------------------------------------------------------------------------
namespace {

// Kh table (h = 1) with the 200 dispersion roots of each entry
void kh_table(int nkh, uint64_t seed, double g_disp, std::vector<double>& kh, std::vector<double>& wvn) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  kh.resize(nkh);
  wvn.resize((size_t)nkh * kNk);
  const double l0 = log(hm::kSurKhMin), l1 = log(hm::kSurKhMax);
  for (int i = 0; i < nkh; ++i) {
    const double k = exp(l0 + (i + u(rng)) / nkh * (l1 - l0));
    const double w = sqrt(k * g_disp);
    kh[i] = w * w / g_disp;
    hm_dispersion(wvn.data() + (size_t)kNk * i, kNk, w, 1.0, g_disp);
  }
}

struct TrainBufs {
  DevBuf kh, wvn, X, Y, M, cell, Z, A, D, Yh, grad, m, v, loss, U, bits, work;
};

}  // namespace

extern "C" int hm_surrogate_blob_size(void) { return hm::kSurBlob; }

// training of the finite-depth network (deep = false) or of the
// infinite-depth branch (deep = true): the same network, optimiser and error
// map, on samples of the evanescent sums or of the Havelock function
static int train_surrogate(hm_ctx* c, bool deep, long long nsamples, int epochs, int batch, double lr,
                           unsigned long long seed, int quant, double* stats) {
  if (!c) return HM_ERR_ARG - 1;
  if (nsamples < 1024) return c->fail(HM_ERR_ARG - 2, "need at least 1024 samples");
  if (epochs < 0) return c->fail(HM_ERR_ARG - 3, "epochs must be >= 0");
  if (batch < 64 || batch > 65536) return c->fail(HM_ERR_ARG - 4, "batch must be in 64..65536");
  if (!(lr > 0.0)) return c->fail(HM_ERR_ARG - 5, "lr must be > 0");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->stream;
  TrainBufs tb;
  auto release = [&]() {
    DevBuf* all[] = {&tb.kh, &tb.wvn, &tb.X, &tb.Y, &tb.M, &tb.cell, &tb.Z, &tb.A, &tb.D, &tb.Yh, &tb.grad, &tb.m,
                     &tb.v, &tb.loss, &tb.U, &tb.bits, &tb.work};
    for (DevBuf* b : all) b->release();
  };
  struct Guard { std::function<void()> f; ~Guard() { f(); } } guard{release};
  const int nkh = 2048;
  std::vector<double> kh, wvn;
  kh_table(nkh, seed ^ 0x5eedull, c->g_disp, kh, wvn);
  const long long n = nsamples;
  const long long nval = n / 5 > 200000 ? n / 5 : 200000;
  const long long nmax = n > nval ? n : nval;
  int rc;
  HIPCHK(c, tb.kh.ensure(sizeof(double) * nkh), "surrogate alloc");
  HIPCHK(c, tb.wvn.ensure(sizeof(double) * wvn.size()), "surrogate alloc");
  HIPCHK(c, tb.X.ensure(4 * sizeof(float) * (size_t)nmax), "surrogate alloc");
  HIPCHK(c, tb.Y.ensure(3 * sizeof(float) * (size_t)nmax), "surrogate alloc");
  HIPCHK(c, hipMemcpyAsync(tb.kh.p, kh.data(), sizeof(double) * nkh, hipMemcpyHostToDevice, s), "surrogate");
  HIPCHK(c, hipMemcpyAsync(tb.wvn.p, wvn.data(), sizeof(double) * wvn.size(), hipMemcpyHostToDevice, s), "surrogate");
  if (deep)
    HIPCHK(c, hm::sur_generate_deep(s, n, seed, tb.X.as<float>(), tb.Y.as<float>(), nullptr, nullptr),
           "surrogate samples");
  else
    HIPCHK(c, hm::sur_generate(s, n, seed, nkh, tb.kh.as<double>(), tb.wvn.as<double>(), tb.X.as<float>(),
                               tb.Y.as<float>(), nullptr, nullptr), "surrogate samples");
  // output scales: RMS of each scaled sum
  std::vector<float> yh((size_t)n * 3);
  HIPCHK(c, hipMemcpyAsync(yh.data(), tb.Y.p, sizeof(float) * yh.size(), hipMemcpyDeviceToHost, s), "surrogate");
  HIPCHK(c, hipStreamSynchronize(s), "surrogate samples");
  float yscale[3];
  for (int o = 0; o < 3; ++o) {
    double acc = 0.0;
    for (long long i = 0; i < n; ++i) acc += (double)yh[3 * i + o] * yh[3 * i + o];
    yscale[o] = (float)sqrt(acc / (double)n);
    if (!(yscale[o] > 0.0f)) yscale[o] = 1.0f;
  }
  // Glorot-uniform initial weights, zero biases
  std::vector<float> p(hm::kSurParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fo, int fi) {
    std::uniform_real_distribution<float> u(-sqrtf(6.0f / (fi + fo)), sqrtf(6.0f / (fi + fo)));
    for (int k = 0; k < fo * fi; ++k) p[off + k] = u(rng);
  };
  fill(hm::kSurOffW0, hm::kSurH, hm::kSurIn);
  for (int l = 0; l < 3; ++l) fill(hm::kSurOffW1 + l * hm::kSurHiddenStride, hm::kSurH, hm::kSurH);
  fill(hm::kSurOffW4, hm::kSurOut, hm::kSurH);
  DevBuf& buf = deep ? c->d_surd : c->d_sur;
  HIPCHK(c, buf.ensure(hm::kSurBufBytes), "surrogate alloc");
  float* P = buf.as<float>();
  float* Q = P + hm::kSurParams;
  float* YS = Q + hm::kSurParams;
  float* MAP = YS + hm::kSurOut;
  HIPCHK(c, hipMemcpyAsync(P, p.data(), sizeof(float) * hm::kSurParams, hipMemcpyHostToDevice, s), "surrogate");
  HIPCHK(c, hipMemcpyAsync(YS, yscale, sizeof yscale, hipMemcpyHostToDevice, s), "surrogate");
  const int B = batch;
  HIPCHK(c, tb.Z.ensure(4 * sizeof(float) * hm::kSurH * (size_t)B), "surrogate alloc");
  HIPCHK(c, tb.A.ensure(4 * sizeof(float) * hm::kSurH * (size_t)B), "surrogate alloc");
  HIPCHK(c, tb.D.ensure(sizeof(float) * (4 * hm::kSurH + 3) * (size_t)B), "surrogate alloc");
  HIPCHK(c, tb.work.ensure(hm::sur_grad_work_bytes(B)), "surrogate alloc");
  HIPCHK(c, tb.Yh.ensure(3 * sizeof(float) * (size_t)B), "surrogate alloc");
  HIPCHK(c, tb.grad.ensure(sizeof(float) * hm::kSurParams), "surrogate alloc");
  HIPCHK(c, tb.m.ensure(sizeof(float) * hm::kSurParams), "surrogate alloc");
  HIPCHK(c, tb.v.ensure(sizeof(float) * hm::kSurParams), "surrogate alloc");
  HIPCHK(c, tb.loss.ensure(sizeof(double) * 2), "surrogate alloc");
  HIPCHK(c, hipMemsetAsync(tb.m.p, 0, sizeof(float) * hm::kSurParams, s), "surrogate");
  HIPCHK(c, hipMemsetAsync(tb.v.p, 0, sizeof(float) * hm::kSurParams, s), "surrogate");
  HIPCHK(c, hm::sur_round(s, P, Q), "surrogate");
  const long long steps_per_epoch = n / B;
  const long long total = steps_per_epoch * epochs;
  double last_loss = 0.0;
  long long t = 0;
  for (int e = 0; e < epochs; ++e) {
    HIPCHK(c, hipMemsetAsync(tb.loss.p, 0, sizeof(double), s), "surrogate");
    for (long long b = 0; b < steps_per_epoch; ++b, ++t) {
      const float* Xb = tb.X.as<float>() + 4 * b * B;
      const float* Yb = tb.Y.as<float>() + 3 * b * B;
      const double lrt = lr * (0.01 + 0.99 * 0.5 * (1.0 + cos(hm::kPi * (double)t / (double)(total ? total : 1))));
      const float* Wt = quant ? Q : P;
      HIPCHK(c, hm::sur_forward(s, B, Xb, Wt, tb.Z.as<float>(), tb.A.as<float>(), tb.Yh.as<float>(), quant != 0),
             "surrogate fwd");
      HIPCHK(c, hm::sur_loss_grad(s, B, Xb, Yb, YS, Wt, tb.Z.as<float>(), tb.A.as<float>(), tb.Yh.as<float>(),
                                  tb.D.as<float>(), tb.grad.as<float>(), tb.loss.as<double>(), tb.work.as<double>()),
             "surrogate bwd");
      HIPCHK(c, hm::sur_adam(s, P, tb.grad.as<float>(), tb.m.as<float>(), tb.v.as<float>(), (float)lrt, (int)(t + 1)),
             "surrogate adam");
      HIPCHK(c, hm::sur_round(s, P, Q), "surrogate");
    }
    double l = 0.0;
    HIPCHK(c, hipMemcpyAsync(&l, tb.loss.p, sizeof(double), hipMemcpyDeviceToHost, s), "surrogate");
    HIPCHK(c, hipStreamSynchronize(s), "surrogate");
    last_loss = l / (3.0 * (double)(steps_per_epoch * B));
  }
  // the inference image of the trained network
  HIPCHK(c, hm::sur_round(s, P, Q, hm_ctx::img_of(buf)), "surrogate");
  // validation set on other Kh values: error map and RMS error
  std::vector<double> khv, wvnv;
  kh_table(512, seed ^ 0xa11da7eull, c->g_disp, khv, wvnv);
  HIPCHK(c, hipMemcpyAsync(tb.kh.p, khv.data(), sizeof(double) * khv.size(), hipMemcpyHostToDevice, s), "surrogate");
  HIPCHK(c, hipMemcpyAsync(tb.wvn.p, wvnv.data(), sizeof(double) * wvnv.size(), hipMemcpyHostToDevice, s), "surrogate");
  HIPCHK(c, tb.M.ensure(3 * sizeof(float) * (size_t)nval), "surrogate alloc");
  HIPCHK(c, tb.cell.ensure(sizeof(int) * (size_t)nval), "surrogate alloc");
  HIPCHK(c, tb.U.ensure(3 * sizeof(float) * (size_t)nval), "surrogate alloc");
  HIPCHK(c, tb.bits.ensure(sizeof(unsigned) * hm::kSurMap), "surrogate alloc");
  if (deep)
    HIPCHK(c, hm::sur_generate_deep(s, nval, seed ^ 0x7a11dull, tb.X.as<float>(), tb.Y.as<float>(), tb.M.as<float>(),
                                    tb.cell.as<int>()), "surrogate validation");
  else
    HIPCHK(c, hm::sur_generate(s, nval, seed ^ 0x7a11dull, 512, tb.kh.as<double>(), tb.wvn.as<double>(),
                               tb.X.as<float>(), tb.Y.as<float>(), tb.M.as<float>(), tb.cell.as<int>()),
           "surrogate validation");
  (deep ? c->surd_quant : c->sur_quant) = quant ? 1 : 0;
  HIPCHK(c, hm::sur_predict(s, nval, tb.X.as<float>(), hm_ctx::dev_of(buf, quant ? 1 : 0), tb.U.as<float>()),
         "surrogate predict");
  HIPCHK(c, hipMemsetAsync(tb.bits.p, 0, sizeof(unsigned) * hm::kSurMap, s), "surrogate");
  HIPCHK(c, hipMemsetAsync(tb.loss.p, 0, sizeof(double), s), "surrogate");
  HIPCHK(c, hm::sur_error_map(s, nval, tb.Y.as<float>(), tb.M.as<float>(), tb.cell.as<int>(), tb.U.as<float>(),
                              tb.bits.as<unsigned>(), tb.loss.as<double>()), "surrogate error map");
  std::vector<unsigned> bits(hm::kSurMap);
  double sq = 0.0;
  HIPCHK(c, hipMemcpyAsync(bits.data(), tb.bits.p, sizeof(unsigned) * hm::kSurMap, hipMemcpyDeviceToHost, s), "surrogate");
  HIPCHK(c, hipMemcpyAsync(&sq, tb.loss.p, sizeof(double), hipMemcpyDeviceToHost, s), "surrogate");
  HIPCHK(c, hipStreamSynchronize(s), "surrogate");
  // bound per cell: 1.5 x the largest validation error; cells without samples are closed
  std::vector<float> map(hm::kSurMap);
  double worst = 0.0;
  long long below4 = 0, below3 = 0;
  for (int k = 0; k < hm::kSurMap; ++k) {
    float e;
    memcpy(&e, &bits[k], sizeof e);
    map[k] = bits[k] ? 1.5f * e : 1e30f;
    if (bits[k]) worst = fmax(worst, (double)e);
    below4 += map[k] <= 1e-4f;
    below3 += map[k] <= 1e-3f;
  }
  HIPCHK(c, hipMemcpyAsync(MAP, map.data(), sizeof(float) * hm::kSurMap, hipMemcpyHostToDevice, s), "surrogate");
  HIPCHK(c, hipStreamSynchronize(s), "surrogate");
  (deep ? c->surd_ready : c->sur_ready) = true;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats[0] = last_loss;
    stats[1] = sqrt(sq / (3.0 * (double)nval)) / ((yscale[0] + yscale[1] + yscale[2]) / 3.0);
    stats[2] = worst;
    stats[3] = secs;
    stats[4] = (double)below4 / hm::kSurMap;
    stats[5] = (double)below3 / hm::kSurMap;
    stats[6] = (double)total;
    stats[7] = (double)n;
  }
  (void)rc;
  return HM_OK;
}

extern "C" int hm_surrogate_train(hm_ctx* c, long long nsamples, int epochs, int batch, double lr,
                                  unsigned long long seed, int quant, double* stats) {
  if (c) ++c->cfg_gen;
  return train_surrogate(c, false, nsamples, epochs, batch, lr, seed, quant, stats);
}

extern "C" int hm_surrogate_train_deep(hm_ctx* c, long long nsamples, int epochs, int batch, double lr,
                                       unsigned long long seed, int quant, double* stats) {
  if (c) ++c->cfg_gen;
  return train_surrogate(c, true, nsamples, epochs, batch, lr, seed, quant, stats);
}

extern "C" int hm_set_surrogate_deep(hm_ctx* c, int enable, double tol) {
  if (!c) return HM_ERR_ARG - 1;
  if (enable && !c->surd_ready) return c->fail(HM_ERR_STATE, "no infinite-depth surrogate trained");
  if (!(tol >= 0.0)) return c->fail(HM_ERR_ARG - 3, "tol must be >= 0");
  HIPCHK(c, c->d_surdcnt.ensure(sizeof(unsigned long long)), "surrogate alloc");
  HIPCHK(c, hipMemsetAsync(c->d_surdcnt.p, 0, sizeof(unsigned long long), c->stream), "surrogate");
  c->surd_on = enable != 0;
  c->surd_tol = tol;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_surrogate_deep_count(hm_ctx* c, long long* taken) {
  if (!c) return HM_ERR_ARG - 1;
  if (!taken) return c->fail(HM_ERR_ARG - 2, "taken is NULL");
  unsigned long long net = 0;
  if (c->d_surdcnt.p) {
    HIPCHK(c, hipMemcpyAsync(&net, c->d_surdcnt.p, sizeof net, hipMemcpyDeviceToHost, c->stream), "surrogate counts");
    HIPCHK(c, hipStreamSynchronize(c->stream), "surrogate counts");
  }
  *taken = (long long)net;
  return HM_OK;
}

extern "C" int hm_surrogate_eval_deep(hm_ctx* c, int n, const double* kr, const double* kv, double* net,
                                      double* plain, double* bound) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 1) return c->fail(HM_ERR_ARG - 2, "n must be >= 1");
  if (!kr || !kv || !net || !plain || !bound) return c->fail(HM_ERR_ARG - 3, "arrays missing");
  if (!c->surd_ready) return c->fail(HM_ERR_STATE, "no infinite-depth surrogate trained");
  for (int i = 0; i < n; ++i)
    if (!(kr[i] >= hm::kSurKrMin && kr[i] <= hm::kSurKrMax && kv[i] >= hm::kSurKvMin && kv[i] <= 0.0))
      return c->fail(HM_ERR_ARG - 3, "point %d outside the infinite-depth surrogate's domain", i);
  return hm::sur_eval_points_deep(c->stream, n, kr, kv, c->surd_dev(), net, plain, bound) == hipSuccess
             ? HM_OK : c->fail(HM_ERR_HIP, "surrogate eval failed");
}

extern "C" int hm_surrogate_get(hm_ctx* c, float* blob, int n) {
  if (!c) return HM_ERR_ARG - 1;
  if (!blob || n < hm::kSurBlob) return c->fail(HM_ERR_ARG - 3, "blob needs %d floats", hm::kSurBlob);
  if (!c->sur_ready) return c->fail(HM_ERR_STATE, "no surrogate trained or loaded");
  const float* b = c->d_sur.as<float>();
  HIPCHK(c, hipMemcpyAsync(blob, b, sizeof(float) * hm::kSurParams, hipMemcpyDeviceToHost, c->stream), "surrogate get");
  HIPCHK(c, hipMemcpyAsync(blob + hm::kSurParams, b + 2 * hm::kSurParams, sizeof(float) * (hm::kSurOut + hm::kSurMap),
                           hipMemcpyDeviceToHost, c->stream), "surrogate get");
  HIPCHK(c, hipStreamSynchronize(c->stream), "surrogate get");
  return HM_OK;
}

extern "C" int hm_surrogate_set(hm_ctx* c, const float* blob, int n) {
  if (!c) return HM_ERR_ARG - 1;
  if (!blob || n < hm::kSurBlob) return c->fail(HM_ERR_ARG - 3, "blob needs %d floats", hm::kSurBlob);
  HIPCHK(c, c->d_sur.ensure(hm::kSurBufBytes), "surrogate alloc");
  float* b = c->d_sur.as<float>();
  HIPCHK(c, hipMemcpyAsync(b, blob, sizeof(float) * hm::kSurParams, hipMemcpyHostToDevice, c->stream), "surrogate set");
  HIPCHK(c, hipMemcpyAsync(b + 2 * hm::kSurParams, blob + hm::kSurParams, sizeof(float) * (hm::kSurOut + hm::kSurMap),
                           hipMemcpyHostToDevice, c->stream), "surrogate set");
  HIPCHK(c, hm::sur_round(c->stream, b, b + hm::kSurParams, c->sur_img()), "surrogate set");
  HIPCHK(c, hipStreamSynchronize(c->stream), "surrogate set");
  c->sur_ready = true;
  ++c->cfg_gen;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_set_surrogate(hm_ctx* c, int enable, double tol) {
  if (!c) return HM_ERR_ARG - 1;
  if (enable && !c->sur_ready) return c->fail(HM_ERR_STATE, "no surrogate trained or loaded");
  if (!(tol >= 0.0)) return c->fail(HM_ERR_ARG - 3, "tol must be >= 0");
  HIPCHK(c, c->d_fbcnt.ensure(sizeof(unsigned long long)), "surrogate alloc");
  HIPCHK(c, hipMemsetAsync(c->d_fbcnt.p, 0, sizeof(unsigned long long), c->stream), "surrogate");
  c->sur_on = enable != 0;
  c->sur_tol = tol;
  for (long long& k : c->sur_counts) k = 0;
  c->assembled = c->factored = c->solved = false;
  return HM_OK;
}

extern "C" int hm_surrogate_counts(hm_ctx* c, long long* counts) {
  if (!c) return HM_ERR_ARG - 1;
  if (!counts) return c->fail(HM_ERR_ARG - 2, "counts is NULL");
  // the network's count accumulates on the device (no host sync per assembly)
  unsigned long long net = 0;
  if (c->d_fbcnt.p) {
    HIPCHK(c, hipMemcpyAsync(&net, c->d_fbcnt.p, sizeof net, hipMemcpyDeviceToHost, c->stream), "surrogate counts");
    HIPCHK(c, hipStreamSynchronize(c->stream), "surrogate counts");
  }
  counts[0] = (long long)net;
  counts[1] = c->sur_counts[1] - (long long)net;
  counts[2] = c->sur_counts[2];
  return HM_OK;
}

extern "C" int hm_surrogate_eval(hm_ctx* c, int n, const double* rh, const double* zs, const double* zd,
                                 const double* kh, double* t_net, double* t_plain, double* bound) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 1) return c->fail(HM_ERR_ARG - 2, "n must be >= 1");
  if (!rh || !zs || !zd || !kh) return c->fail(HM_ERR_ARG - 3, "inputs missing");
  if (!c->sur_ready) return c->fail(HM_ERR_STATE, "no surrogate trained or loaded");
  for (int i = 0; i < n; ++i) {
    const double zf = 0.5 * (zs[i] + zd[i]), zp = 0.5 * (zs[i] - zd[i]);
    if (!(rh[i] >= hm::kSurRhMin && rh[i] <= hm::kSurRhMax) || !(kh[i] >= hm::kSurKhMin && kh[i] <= hm::kSurKhMax) ||
        !(zf >= -hm::kSurDepth && zf <= 0.0 && zp >= -hm::kSurDepth && zp <= 0.0))
      return c->fail(HM_ERR_ARG - 3, "point %d outside the surrogate's domain", i);
  }
  std::vector<double> wv((size_t)n * kNk);
  for (int i = 0; i < n; ++i) {
    const double w = sqrt(kh[i] * c->g_disp);
    hm_dispersion(wv.data() + (size_t)kNk * i, kNk, w, 1.0, c->g_disp);
  }
  return hm::sur_eval_points(c->stream, n, rh, zs, zd, kh, wv.data(), c->g_disp, c->sur_dev(), t_net, t_plain, bound)
             == hipSuccess ? HM_OK : c->fail(HM_ERR_HIP, "surrogate eval failed");
}

extern "C" int hm_surrogate_predict(hm_ctx* c, int n, const float* x, float* u) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 1) return c->fail(HM_ERR_ARG - 2, "n must be >= 1");
  if (!x || !u) return c->fail(HM_ERR_ARG - 3, "x or u is NULL");
  if (!c->sur_ready) return c->fail(HM_ERR_STATE, "no surrogate trained or loaded");
  DevBuf d;
  int rc = c->hip(d.ensure(sizeof(float) * 7 * (size_t)n), "surrogate predict");
  float* dx = d.as<float>();
  float* du = dx + 4 * (size_t)n;
  if (rc == HM_OK &&
      (rc = c->hip(hipMemcpyAsync(dx, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, c->stream), "surrogate predict")) == HM_OK &&
      (rc = c->hip(hm::sur_predict(c->stream, n, dx, c->sur_dev(), du), "surrogate predict")) == HM_OK &&
      (rc = c->hip(hipMemcpyAsync(u, du, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->stream), "surrogate predict")) == HM_OK)
    rc = c->hip(hipStreamSynchronize(c->stream), "surrogate predict");
  d.release();
  return rc;
}

extern "C" int hm_surrogate_grad(hm_ctx* c, int batch, const float* x, const float* u, int quant, float* grad,
                                 double* loss) {
  if (!c) return HM_ERR_ARG - 1;
  if (batch < 1 || batch > 65536) return c->fail(HM_ERR_ARG - 2, "batch must be in 1..65536");
  if (!x || !u || !grad) return c->fail(HM_ERR_ARG - 3, "arrays missing");
  if (!c->sur_ready) return c->fail(HM_ERR_STATE, "no surrogate trained or loaded");
  hipStream_t s = c->stream;
  TrainBufs tb;
  auto release = [&]() {
    DevBuf* all[] = {&tb.X, &tb.Y, &tb.Z, &tb.A, &tb.D, &tb.Yh, &tb.grad, &tb.loss, &tb.work};
    for (DevBuf* b : all) b->release();
  };
  struct Guard { std::function<void()> f; ~Guard() { f(); } } guard{release};
  const int B = batch;
  HIPCHK(c, tb.X.ensure(4 * sizeof(float) * B), "grad alloc");
  HIPCHK(c, tb.Y.ensure(3 * sizeof(float) * B), "grad alloc");
  HIPCHK(c, tb.Z.ensure(4 * sizeof(float) * hm::kSurH * (size_t)B), "grad alloc");
  HIPCHK(c, tb.A.ensure(4 * sizeof(float) * hm::kSurH * (size_t)B), "grad alloc");
  HIPCHK(c, tb.D.ensure(sizeof(float) * (4 * hm::kSurH + 3) * (size_t)B), "grad alloc");
  HIPCHK(c, tb.work.ensure(hm::sur_grad_work_bytes(B)), "grad alloc");
  HIPCHK(c, tb.Yh.ensure(3 * sizeof(float) * B), "grad alloc");
  HIPCHK(c, tb.grad.ensure(sizeof(float) * hm::kSurParams), "grad alloc");
  HIPCHK(c, tb.loss.ensure(sizeof(double)), "grad alloc");
  HIPCHK(c, hipMemcpyAsync(tb.X.p, x, 4 * sizeof(float) * B, hipMemcpyHostToDevice, s), "grad");
  HIPCHK(c, hipMemcpyAsync(tb.Y.p, u, 3 * sizeof(float) * B, hipMemcpyHostToDevice, s), "grad");
  HIPCHK(c, hipMemsetAsync(tb.loss.p, 0, sizeof(double), s), "grad");
  const hm::SurDev m = c->sur_dev();
  const float* w = quant ? m.q : m.p;
  HIPCHK(c, hm::sur_forward(s, B, tb.X.as<float>(), w, tb.Z.as<float>(), tb.A.as<float>(), tb.Yh.as<float>(), quant != 0),
         "grad fwd");
  HIPCHK(c, hm::sur_loss_grad(s, B, tb.X.as<float>(), tb.Y.as<float>(), m.yscale, w, tb.Z.as<float>(), tb.A.as<float>(),
                              tb.Yh.as<float>(), tb.D.as<float>(), tb.grad.as<float>(), tb.loss.as<double>(),
                              tb.work.as<double>()), "grad bwd");
  double l = 0.0;
  HIPCHK(c, hipMemcpyAsync(grad, tb.grad.p, sizeof(float) * hm::kSurParams, hipMemcpyDeviceToHost, s), "grad");
  HIPCHK(c, hipMemcpyAsync(&l, tb.loss.p, sizeof(double), hipMemcpyDeviceToHost, s), "grad");
  HIPCHK(c, hipStreamSynchronize(s), "grad");
  if (loss) *loss = l / (3.0 * B);
  return HM_OK;
}

extern "C" int hm_forces(hm_ctx* c, double* amas, double* bdmp, double* exfc) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->solved) return c->fail(HM_ERR_STATE, "hm_forces before hm_solve");
  const int m = c->m(), nbeta = (int)c->beta.size();
  HIPCHK(c, c->d_amas.ensure(sizeof(double) * m * m), "forces alloc");
  HIPCHK(c, c->d_bdmp.ensure(sizeof(double) * m * m), "forces alloc");
  HIPCHK(c, c->d_exfc.ensure(2 * sizeof(double) * m * (nbeta ? nbeta : 1)), "forces alloc");
  c->stage_begin(kForces);
  HIPCHK(c, hm::forces(c->stream, c->geom(), c->freq(), c->d_beta.as<double>(), nbeta, c->xw, c->yw, c->d_X.as<double>(),
                       c->n, c->d_amas.as<double>(), c->d_bdmp.as<double>(), c->d_exfc.as<double>()), "forces");
  if (amas) HIPCHK(c, hipMemcpyAsync(amas, c->d_amas.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->stream), "forces");
  if (bdmp) HIPCHK(c, hipMemcpyAsync(bdmp, c->d_bdmp.p, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->stream), "forces");
  if (exfc && nbeta) HIPCHK(c, hipMemcpyAsync(exfc, c->d_exfc.p, 2 * sizeof(double) * m * nbeta, hipMemcpyDeviceToHost, c->stream), "forces");
  HIPCHK(c, hipStreamSynchronize(c->stream), "forces");
  return c->stage_end(kForces);
}

extern "C" int hm_solve_frequency(hm_ctx* c, double w1, int precision, double* amas, double* bdmp, double* exfc) {
  int rc;
  if ((rc = hm_set_frequency(c, w1))) return rc;
  if ((rc = hm_assemble(c))) return rc;
  if ((rc = hm_factor(c, precision))) return rc;
  if ((rc = hm_solve(c, nullptr))) return rc;
  return hm_forces(c, amas, bdmp, exfc);
}

// the settings a clone must share with its context (with cfg_gen)
static std::string worker_signature(const hm_ctx* c) {
  char b[512];
  snprintf(b, sizeof b, "%d %d %d %d %d %d %d %.17g %.17g %d %d %d %.17g %d %d %d %d %.17g %d %d %d %d %.17g %d %d %d",
           c->series_variant, c->lu_coop, c->pair_sym, c->block_share, c->hm_on, c->hm_tree, c->ht_mode, c->hm_eps,
           c->hm_eps_s, c->hm_kmax, c->hm_overlap, c->gm_restart, c->gm_tol, c->gm_maxiter, c->gm_precond,
           c->sur_on, c->sur_quant, c->sur_tol, c->surd_on, c->frames, c->ht_leaf, c->nbody, c->ht_eta,
           c->sur_ready, c->pinv16_on, c->share_sym);
  return b;
}

// a clone of c's configuration: geometry (and body frames), depth, headings,
// constants, kernel and solver settings, the learned surrogate's parameters
static int clone_config(hm_ctx* w, hm_ctx* c) {
  int rc;
  if ((rc = hm_set_constants(w, c->rho, c->g, c->g_disp, c->amp, c->xw, c->yw))) return rc;
  const int n = c->n;
  if ((rc = hm_set_geometry(w, c->nbody, c->body_start.data(), c->geom_ntnd, c->geom_xyz.data(), c->ncn_host.data(),
                            c->geom_ncon.data(), c->fp_host.data(), c->ds_host.data(), c->pnsz_host.data(),
                            c->dxyz_host.data(), c->ni, c->ni ? c->fp_host.data() + 3 * (size_t)n : nullptr)))
    return rc;
  if (c->frames &&
      (rc = hm_set_body_origins(w, c->nbody, c->origin_host.data(), c->ni ? c->wp_start_host.data() : nullptr)))
    return rc;
  w->series_variant = c->series_variant;
  w->lu_coop = c->lu_coop;
  w->pair_sym = c->pair_sym;
  w->block_share = c->block_share;
  w->share_sym = c->share_sym;
  w->hm_overlap = c->hm_overlap;
  w->ht_mode = c->ht_mode;
  w->ht_leaf = c->ht_leaf;
  w->ht_eta = c->ht_eta;
  w->pinv16_on = c->pinv16_on;
  if ((rc = hm_set_depth(w, c->h))) return rc;
  if ((rc = hm_set_headings(w, (int)c->beta.size(), c->beta.data()))) return rc;
  if ((rc = hm_set_gmres(w, c->gm_restart, c->gm_tol, c->gm_maxiter, c->gm_precond))) return rc;
  if (c->sur_ready) {
    std::vector<float> blob(hm::kSurBlob);
    if ((rc = hm_surrogate_get(c, blob.data(), hm::kSurBlob)) || (rc = hm_surrogate_set(w, blob.data(), hm::kSurBlob)))
      return rc;
    w->sur_quant = c->sur_quant;
    if ((rc = hm_set_surrogate(w, c->sur_on ? 1 : 0, c->sur_tol))) return rc;
  }
  if ((c->hm_on || c->hm_tree) && (rc = hm_set_hmatrix(w, 1, c->hm_eps, c->hm_eps_s, c->hm_kmax))) return rc;
  return HM_OK;
}

extern "C" int hm_set_workers(hm_ctx* c, int nworkers) {
  if (!c) return HM_ERR_ARG - 1;
  if (nworkers < 1 || nworkers > 8) return c->fail(HM_ERR_ARG - 2, "nworkers must be in 1..8");
  c->nworkers = nworkers;
  // clones beyond the new count release their device memory now
  while ((int)c->workers.size() > nworkers - 1) {
    hm_destroy(c->workers.back());
    c->workers.pop_back();
  }
  return HM_OK;
}

extern "C" int hm_run_frequencies(hm_ctx* c, int nfreq, const double* w1, int precision, double* amas, double* bdmp,
                                  double* exfc) {
  if (!c) return HM_ERR_ARG - 1;
  if (nfreq < 0 || (nfreq && !w1)) return c->fail(HM_ERR_ARG - 2, "bad frequency list");
  const size_t mm = (size_t)c->m() * c->m(), mb = 2 * (size_t)c->m() * c->beta.size();
  auto one = [&](hm_ctx* x, int k) {
    return hm_solve_frequency(x, w1[k], precision, amas ? amas + k * mm : nullptr, bdmp ? bdmp + k * mm : nullptr,
                              exfc ? exfc + k * mb : nullptr);
  };
  // several frequencies in flight need clones of a configuration they can copy
  // (not the infinite-depth surrogate or a cluster tree built from a matrix)
  const int nw = std::min(c->nworkers, nfreq);
  const bool can = nw > 1 && c->have_depth && !c->surd_on && !c->ht_from_build && !c->hd_ready;
  if (!can) {
    for (int k = 0; k < nfreq; ++k) {
      int rc = one(c, k);
      if (rc) return rc;
    }
    return HM_OK;
  }
  const std::string sig = worker_signature(c);
  if (c->workers_gen != c->cfg_gen || c->workers_sig != sig || (int)c->workers.size() < nw - 1) {
    for (hm_ctx* w : c->workers) hm_destroy(w);
    c->workers.clear();
    for (int k = 0; k < c->nworkers - 1; ++k) {
      hm_ctx* w = nullptr;
      int rc = hm_create(&w);
      if (rc) return c->fail(rc, "worker context: hm_create failed");
      c->workers.push_back(w);
      if ((rc = clone_config(w, c))) return c->fail(rc, "worker context: %s", w->err.c_str());
    }
    c->workers_gen = c->cfg_gen;
    c->workers_sig = sig;
  }
  // each context takes the next frequency not yet taken
  std::atomic<int> next(0);
  std::vector<int> rcs(nw, HM_OK);
  auto run = [&](int t) {
    hm_ctx* x = t == 0 ? c : c->workers[t - 1];
    (void)hipSetDevice(c->device);
    for (int k = next++; k < nfreq; k = next++) {
      int rc = one(x, k);
      if (rc) { rcs[t] = rc; next = nfreq; break; }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nw; ++t) th.emplace_back(run, t);
  run(0);
  for (std::thread& t : th) t.join();
  for (int t = 0; t < nw; ++t)
    if (rcs[t]) {
      if (t > 0) return c->fail(rcs[t], "frequency worker %d: %s", t, c->workers[t - 1]->err.c_str());
      return rcs[t];
    }
  return HM_OK;
}

extern "C" int hm_set_timing(hm_ctx* c, int enable) {
  if (!c) return HM_ERR_ARG - 1;
  int rc = c->resolve_timing();
  c->timing = enable != 0;
  for (int s = 0; s < kStages; ++s) { c->stage_ms[s] = 0.0; c->stage_cnt[s] = 0; }
  return rc;
}

extern "C" int hm_get_timing(hm_ctx* c, double* ms, int* counts, int nstage) {
  if (!c) return HM_ERR_ARG - 1;
  int rc = c->resolve_timing();
  if (rc) return rc;
  for (int s = 0; s < nstage && s < kStages; ++s) {
    if (ms) ms[s] = c->stage_ms[s];
    if (counts) counts[s] = c->stage_cnt[s];
  }
  return HM_OK;
}

extern "C" int hm_get_pair_histogram(hm_ctx* c, long long* counts, int nkeys) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->have_depth) return c->fail(HM_ERR_STATE, "hm_get_pair_histogram before hm_set_depth");
  if (!counts) return c->fail(HM_ERR_ARG - 2, "counts is NULL");
  if (nkeys < hm::kNumKeys) return c->fail(HM_ERR_ARG - 3, "nkeys must be >= %d", hm::kNumKeys);
  for (int k = 0; k < nkeys; ++k) counts[k] = k < (int)c->hist_host.size() ? c->hist_host[k] : 0;
  return HM_OK;
}

extern "C" int hm_get_system(hm_ctx* c, double* amat, double* smat, double* rhs) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->assembled || c->factored) return c->fail(HM_ERR_STATE, "hm_get_system needs an assembled, unfactored system");
  if (c->hm_on) return c->fail(HM_ERR_STATE, "the H-matrix form holds no dense system");
  const size_t cplx = 2 * sizeof(double), nt = c->nt();
  if (amat) HIPCHK(c, hipMemcpyAsync(amat, c->d_A.p, cplx * nt * c->n, hipMemcpyDeviceToHost, c->stream), "get system");
  if (smat) HIPCHK(c, hipMemcpyAsync(smat, c->d_S.p, cplx * nt * c->n, hipMemcpyDeviceToHost, c->stream), "get system");
  if (rhs) HIPCHK(c, hipMemcpyAsync(rhs, c->d_R.p, cplx * nt * c->ncol(), hipMemcpyDeviceToHost, c->stream), "get system");
  HIPCHK(c, hipStreamSynchronize(c->stream), "get system");
  return HM_OK;
}

extern "C" int hm_get_normal_equations(hm_ctx* c, double* cmat, double* drhs) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->assembled || c->factored) return c->fail(HM_ERR_STATE, "hm_get_normal_equations needs an assembled, unfactored system");
  if (c->ni == 0) return c->fail(HM_ERR_STATE, "no waterplane panels: the system has no normal equations");
  const size_t cplx = 2 * sizeof(double), n = c->n;
  if (cmat) HIPCHK(c, hipMemcpyAsync(cmat, c->d_C.p, cplx * n * n, hipMemcpyDeviceToHost, c->stream), "get normal equations");
  if (drhs) HIPCHK(c, hipMemcpyAsync(drhs, c->d_D.p, cplx * n * c->ncol(), hipMemcpyDeviceToHost, c->stream), "get normal equations");
  HIPCHK(c, hipStreamSynchronize(c->stream), "get normal equations");
  return HM_OK;
}

extern "C" int hm_get_potentials(hm_ctx* c, double* mxpot) {
  if (!c) return HM_ERR_ARG - 1;
  if (!c->solved) return c->fail(HM_ERR_STATE, "hm_get_potentials before hm_solve");
  if (!mxpot) return c->fail(HM_ERR_ARG - 2, "mxpot is NULL");
  HIPCHK(c, hipMemcpyAsync(mxpot, c->d_X.p, 2 * sizeof(double) * (size_t)c->n * c->ncol(), hipMemcpyDeviceToHost, c->stream), "get potentials");
  HIPCHK(c, hipStreamSynchronize(c->stream), "get potentials");
  return HM_OK;
}

// ---------------------------------------------------------------------------
// element-wise entry points
// ---------------------------------------------------------------------------
static int green_common(hm_ctx* c, int n, bool deep, const double* const* in, double v, const double* wvn, int nk,
                        double h, const int* flag, double* grn) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 0) return c->fail(HM_ERR_ARG - 2, "n < 0");
  if (n == 0) return HM_OK;
  for (int k = 0; k < 6; ++k)
    if (!in[k]) return c->fail(HM_ERR_ARG - 3 - k, "coordinate array %d is NULL", k);
  if (!flag || !grn) return c->fail(HM_ERR_ARG - 1, "flag/grn NULL");
  if (!deep && (!wvn || nk < kNk)) return c->fail(HM_ERR_ARG - 10, "wvn must hold %d roots", kNk);
  if (!deep && !(h > 0.0)) return c->fail(HM_ERR_ARG - 12, "h must be > 0");
  DevBuf tmp;
  const size_t sz = (size_t)n * sizeof(double);
  const size_t flag_bytes = (((size_t)n * sizeof(int) + 15) / 16) * 16;
  HIPCHK(c, tmp.ensure(6 * sz + flag_bytes + 8 * sz + kNk * sizeof(double)), "green batch alloc");
  char* base = tmp.as<char>();
  double* d[6];
  for (int k = 0; k < 6; ++k) {
    d[k] = reinterpret_cast<double*>(base + k * sz);
    HIPCHK(c, hipMemcpyAsync(d[k], in[k], sz, hipMemcpyHostToDevice, c->stream), "green batch");
  }
  int* df = reinterpret_cast<int*>(base + 6 * sz);
  double* dout = reinterpret_cast<double*>(base + 6 * sz + flag_bytes);
  double* dw = dout + 8 * (size_t)n;
  HIPCHK(c, hipMemcpyAsync(df, flag, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream), "green batch");
  if (!deep) HIPCHK(c, hipMemcpyAsync(dw, wvn, kNk * sizeof(double), hipMemcpyHostToDevice, c->stream), "green batch");
  int rc = c->hip(hm::green_batch(c->stream, n, deep, d[0], d[1], d[2], d[3], d[4], d[5], v, dw, h, df, dout), "green batch");
  if (!rc) rc = c->hip(hipMemcpyAsync(grn, dout, 8 * sz, hipMemcpyDeviceToHost, c->stream), "green batch");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "green batch");
  (void)hipStreamSynchronize(c->stream);
  tmp.release();
  return rc;
}

extern "C" int hm_fingreen3d_batch(hm_ctx* c, int n, const double* xf, const double* xp, const double* yf,
                                   const double* yp, const double* zf, const double* zp, double v, const double* wvn,
                                   int nk, double h, const int* flag, double* grn) {
  const double* in[6] = {xf, xp, yf, yp, zf, zp};
  return green_common(c, n, false, in, v, wvn, nk, h, flag, grn);
}

extern "C" int hm_fingreen3d_batch_noalloc(hm_ctx* c, int n, const double* xf, const double* xp, const double* yf,
                                           const double* yp, const double* zf, const double* zp, double v,
                                           const double* wvn, int nk, double h, const int* flag, double* grn) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 0) return c->fail(HM_ERR_ARG - 2, "n < 0");
  if (n > hm::kNoAllocMaxPoints)
    return c->fail(HM_ERR_ARG - 2, "n = %d exceeds the %d pairs of the fixed device array", n, hm::kNoAllocMaxPoints);
  if (n == 0) return HM_OK;
  const double* in[6] = {xf, xp, yf, yp, zf, zp};
  for (int k = 0; k < 6; ++k)
    if (!in[k]) return c->fail(HM_ERR_ARG - 3 - k, "coordinate array %d is NULL", k);
  if (!flag || !grn) return c->fail(HM_ERR_ARG - 1, "flag/grn NULL");
  if (!wvn || nk < kNk) return c->fail(HM_ERR_ARG - 10, "wvn must hold %d roots", kNk);
  if (!(h > 0.0)) return c->fail(HM_ERR_ARG - 12, "h must be > 0");
  return c->hip(hm::noalloc_green(c->stream, n, false, in, v, wvn, h, flag, grn), "green batch (no allocation)");
}

extern "C" int hm_infgreen3d_batch(hm_ctx* c, int n, const double* xf, const double* xp, const double* yf,
                                   const double* yp, const double* zf, const double* zp, double v, const int* flag,
                                   double* grn) {
  const double* in[6] = {xf, xp, yf, yp, zf, zp};
  return green_common(c, n, true, in, v, nullptr, 0, -1.0, flag, grn);
}

// Field-point potential, pressure and elevation of every radiation mode and
// heading (CalPotentialMulti, CalPressureMulti, CalElevationMulti:
// PressureElevationMulti.f90:55-348) after hm_solve: the field points' S- and
// A-like rows against the hull, pot = (Sf W - Af X) / (2 pi or 4 pi) with W the
// right-hand sides' source terms and X the potentials, plus the incident
// potential on the diffraction columns
extern "C" int hm_field_points(hm_ctx* c, int nfp, const double* xfp, double* pot, double* prs, double* elv) {
  if (!c) return HM_ERR_ARG - 1;
  if (nfp < 0) return c->fail(HM_ERR_ARG - 2, "nfp < 0");
  if (nfp == 0) return HM_OK;
  if (!xfp) return c->fail(HM_ERR_ARG - 3, "xfp is NULL");
  if (!c->solved) return c->fail(HM_ERR_STATE, "hm_field_points before hm_solve");
  c->join_side();
  const int n = c->n, m = 6 * c->nbody, nbeta = (int)c->beta.size(), ncol = c->ncol();
  const size_t cplx = 2 * sizeof(double);
  const size_t rows = cplx * (size_t)nfp * n, outs = cplx * (size_t)nfp * ncol;
  DevBuf tmp;
  HIPCHK(c, tmp.ensure(3 * sizeof(double) * (size_t)nfp + 256 + 2 * rows + 3 * outs), "field points alloc");
  char* base = tmp.as<char>();
  double* dx = reinterpret_cast<double*>(base);
  double* Af = reinterpret_cast<double*>(base + ((3 * sizeof(double) * (size_t)nfp + 255) & ~(size_t)255));
  double* Sf = Af + 2 * (size_t)nfp * n;
  double* P = Sf + 2 * (size_t)nfp * n;
  double* Pr = P + 2 * (size_t)nfp * ncol;
  double* El = Pr + 2 * (size_t)nfp * ncol;
  hipStream_t s = c->stream;
  const hm::GeomDev gd = c->geom();
  const hm::FreqDev f = c->freq();
  int rc = c->hip(hipMemcpyAsync(dx, xfp, 3 * sizeof(double) * (size_t)nfp, hipMemcpyHostToDevice, s), "field points");
  if (!rc) rc = c->hip(hm::field_rows(s, gd, f, dx, nfp, Af, Sf), "field points");
  if (!rc) rc = c->hip(hm::cgemm<double>(s, false, nfp, ncol, n, 1.0, 0.0, Sf, nfp, 0, c->d_W.as<double>(), n, 0, 0.0,
                                         0.0, P, nfp, 0, 1), "field points");
  if (!rc) rc = c->hip(hm::cgemm<double>(s, false, nfp, ncol, n, -1.0, 0.0, Af, nfp, 0, c->d_X.as<double>(), n, 0, 1.0,
                                         0.0, P, nfp, 0, 1), "field points");
  if (!rc) rc = c->hip(hm::field_finish(s, gd, f, dx, nfp, m, nbeta, c->d_beta.as<double>(), c->xw, c->yw, P, Pr, El),
                       "field points");
  if (!rc && pot) rc = c->hip(hipMemcpyAsync(pot, P, outs, hipMemcpyDeviceToHost, s), "field points");
  if (!rc && prs) rc = c->hip(hipMemcpyAsync(prs, Pr, outs, hipMemcpyDeviceToHost, s), "field points");
  if (!rc && elv) rc = c->hip(hipMemcpyAsync(elv, El, outs, hipMemcpyDeviceToHost, s), "field points");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "field points");
  (void)hipStreamSynchronize(s);
  tmp.release();
  return rc;
}

// Motion response (SolveMotionMulti, SolveMotionMulti.f90:46-101): the body
// matrices (nbody x 6 x 6, Fortran order, element (b, i, j) at b + nbody (i +
// 6 j)) on the block diagonal (LOCALTOGLOBAL :24-40); LEFT = -w^2 (MATX + AMAS)
// + i w (BDMP + BLNR) + CRS + KSTF, RIGHT = -Im(EXFC) - i Re(EXFC); with a
// quadratic damping (largest |BQDR| >= 1e-6) the equivalent linear damping
// BQDR(i, j) w |x_j| is added and the solve repeated until the summed relative
// change of x is <= 1e-6. Each solve is the FP64 LU on the device.
extern "C" int hm_solve_motion(hm_ctx* c, int nbody, double w1, const double* amas, const double* bdmp,
                               const double* blnr, const double* bqdr, const double* matx, const double* crs,
                               const double* kstf, const double* exfc, double* dspl, int* iters) {
  if (!c) return HM_ERR_ARG - 1;
  if (nbody < 1) return c->fail(HM_ERR_ARG - 2, "nbody must be >= 1");
  if (!(w1 >= 0.0) || !isfinite(w1)) return c->fail(HM_ERR_ARG - 3, "w1 must be finite and >= 0");
  if (!amas || !bdmp || !blnr || !bqdr || !matx || !crs || !kstf || !exfc || !dspl)
    return c->fail(HM_ERR_ARG - 4, "NULL array");
  const int m = 6 * nbody;
  const size_t mm = (size_t)m * m;
  auto global = [&](const double* loc, std::vector<double>& g) {   // LOCALTOGLOBAL
    g.assign(mm, 0.0);
    for (int b = 0; b < nbody; ++b)
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          g[(size_t)(6 * b + i) + (size_t)(6 * b + j) * m] = loc[b + (size_t)nbody * (i + 6 * j)];
  };
  std::vector<double> bl, bq, mx, cr, ks;
  global(blnr, bl);
  global(bqdr, bq);
  global(matx, mx);
  global(crs, cr);
  global(kstf, ks);
  double norm = 0.0;
  for (double v : bq) norm = fmax(norm, fabs(v));
  std::vector<double> left(2 * mm), right(2 * (size_t)m), x(2 * (size_t)m, 0.0);
  DevBuf da, db, dp, di, dc;
  HIPCHK(c, da.ensure(2 * sizeof(double) * mm), "motion alloc");
  HIPCHK(c, db.ensure(2 * sizeof(double) * m), "motion alloc");
  HIPCHK(c, dp.ensure(sizeof(int) * m), "motion alloc");
  HIPCHK(c, di.ensure(sizeof(int)), "motion alloc");
  HIPCHK(c, dc.ensure(sizeof(unsigned)), "motion alloc");
  hipStream_t s = c->stream;
  // one solve with the equivalent damping of the current x (none on the first)
  auto solve = [&](bool quad) -> int {
    for (size_t e = 0; e < mm; ++e) {
      const int j = (int)(e / m);
      const double vd = quad ? bq[e] * w1 * hypot(x[2 * j], x[2 * j + 1]) : 0.0;
      left[2 * e] = -w1 * w1 * (mx[e] + amas[e]) + cr[e] + ks[e];
      left[2 * e + 1] = w1 * (bdmp[e] + bl[e] + vd);
    }
    for (int i = 0; i < m; ++i) {
      right[2 * i] = -exfc[2 * i + 1];
      right[2 * i + 1] = -exfc[2 * i];
    }
    int info = 0;
    int rc = c->hip(hipMemcpyAsync(da.p, left.data(), 2 * sizeof(double) * mm, hipMemcpyHostToDevice, s), "motion");
    if (!rc) rc = c->hip(hipMemcpyAsync(db.p, right.data(), 2 * sizeof(double) * m, hipMemcpyHostToDevice, s), "motion");
    if (!rc) rc = c->hip(hipMemsetAsync(di.p, 0, sizeof(int), s), "motion");
    if (!rc) rc = c->hip(hipMemsetAsync(dc.p, 0, sizeof(unsigned), s), "motion");
    if (!rc) rc = c->hip(hm::lu_factor<double>(s, da.as<double>(), m, 0, m, dp.as<int>(), di.as<int>(),
                                               dc.as<unsigned>(), 1), "motion");
    if (!rc) rc = c->hip(hm::lu_solve<double>(s, da.as<double>(), m, 0, m, dp.as<int>(), db.as<double>(), m, 0, 1, 1),
                         "motion");
    if (!rc) rc = c->hip(hipMemcpyAsync(right.data(), db.p, 2 * sizeof(double) * m, hipMemcpyDeviceToHost, s), "motion");
    if (!rc) rc = c->hip(hipMemcpyAsync(&info, di.p, sizeof(int), hipMemcpyDeviceToHost, s), "motion");
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "motion");
    if (rc) return rc;
    if (info > 0) return c->fail(info, "zero pivot in column %d of the motion equations", info);
    return HM_OK;
  };
  int rc = solve(false);
  if (rc) return rc;
  x = right;
  int it = 0;
  if (norm >= 1.0e-6) {
    for (double rerr = 100.0; rerr > 1.0e-6;) {
      if (++it > 500) return c->fail(HM_ERR_STATE, "quadratic damping iteration: no convergence in 500 steps");
      if ((rc = solve(true))) return rc;
      rerr = 0.0;
      for (int j = 0; j < m; ++j) {
        const double a = hypot(right[2 * j], right[2 * j + 1]);
        if (a > 0.0) rerr += hypot(right[2 * j] - x[2 * j], right[2 * j + 1] - x[2 * j + 1]) / a;
      }
      x = right;
    }
  }
  for (int j = 0; j < 2 * m; ++j) dspl[j] = x[j];
  if (iters) *iters = it;
  return HM_OK;
}

extern "C" int hm_ik01a_batch(hm_ctx* c, int n, const double* x, double* bk0, double* bk1, int mode) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 0) return c->fail(HM_ERR_ARG - 2, "n < 0");
  if (n == 0) return HM_OK;
  if (!x || !bk0 || !bk1) return c->fail(HM_ERR_ARG - 3, "NULL array");
  if (mode != 0 && mode != 1) return c->fail(HM_ERR_ARG - 6, "mode must be 0 or 1");
  DevBuf tmp;
  const size_t sz = (size_t)n * sizeof(double);
  HIPCHK(c, tmp.ensure(3 * sz), "ik01a alloc");
  double* dx = tmp.as<double>();
  double* dout = dx + n;
  std::vector<double> out(2 * (size_t)n);
  int rc = c->hip(hipMemcpyAsync(dx, x, sz, hipMemcpyHostToDevice, c->stream), "ik01a");
  if (!rc) rc = c->hip(hm::ik01a_batch(c->stream, n, dx, mode, dout), "ik01a");
  if (!rc) rc = c->hip(hipMemcpyAsync(out.data(), dout, 2 * sz, hipMemcpyDeviceToHost, c->stream), "ik01a");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "ik01a");
  (void)hipStreamSynchronize(c->stream);
  tmp.release();
  if (rc) return rc;
  for (int i = 0; i < n; ++i) {
    bk0[i] = out[2 * (size_t)i];
    bk1[i] = out[2 * (size_t)i + 1];
  }
  return HM_OK;
}

extern "C" int hm_sglintbd_batch(hm_ctx* c, int n, const int* ncn, const double* verts, const double* xf, double* sij,
                                 double* dij) {
  if (!c) return HM_ERR_ARG - 1;
  if (n < 0) return c->fail(HM_ERR_ARG - 2, "n < 0");
  if (n == 0) return HM_OK;
  if (!ncn || !verts || !xf || !sij || !dij) return c->fail(HM_ERR_ARG - 3, "NULL array");
  for (int i = 0; i < n; ++i)
    if (ncn[i] != 3 && ncn[i] != 4) return c->fail(HM_ERR_ARG - 3, "panel %d has %d vertices", i, ncn[i]);
  DevBuf tmp;
  const size_t n8 = (size_t)n * sizeof(double);
  HIPCHK(c, tmp.ensure(((n * sizeof(int) + 15) / 16) * 16 + 12 * n8 + 3 * n8 + 4 * n8), "rankine batch alloc");
  int* dn = tmp.as<int>();
  double* dv = reinterpret_cast<double*>(tmp.as<char>() + ((n * sizeof(int) + 15) / 16) * 16);
  double* dx = dv + 12 * (size_t)n;
  double* dout = dx + 3 * (size_t)n;
  HIPCHK(c, hipMemcpyAsync(dn, ncn, n * sizeof(int), hipMemcpyHostToDevice, c->stream), "rankine batch");
  HIPCHK(c, hipMemcpyAsync(dv, verts, 12 * n8, hipMemcpyHostToDevice, c->stream), "rankine batch");
  HIPCHK(c, hipMemcpyAsync(dx, xf, 3 * n8, hipMemcpyHostToDevice, c->stream), "rankine batch");
  HIPCHK(c, hm::rankine_batch(c->stream, n, dn, dv, dx, dout), "rankine batch");
  std::vector<double> out(4 * (size_t)n);
  HIPCHK(c, hipMemcpyAsync(out.data(), dout, 4 * n8, hipMemcpyDeviceToHost, c->stream), "rankine batch");
  HIPCHK(c, hipStreamSynchronize(c->stream), "rankine batch");
  tmp.release();
  for (int i = 0; i < n; ++i) {
    sij[i] = out[4 * i];
    for (int k = 0; k < 3; ++k) dij[3 * i + k] = out[4 * i + 1 + k];
  }
  return HM_OK;
}

// ---------------------------------------------------------------------------
// LAPACK drop-ins
// ---------------------------------------------------------------------------
extern "C" int hm_zgetrf(hm_ctx* c, int m, int n, double* a, int lda, int* ipiv, int* info) {
  if (!c) return HM_ERR_ARG - 1;
  if (m < 0) return c->fail(HM_ERR_ARG - 2, "m < 0");
  if (n != m) return c->fail(HM_ERR_ARG - 3, "only square matrices are supported (m=%d n=%d)", m, n);
  if (!a && n) return c->fail(HM_ERR_ARG - 4, "a is NULL");
  if (lda < (m > 1 ? m : 1)) return c->fail(HM_ERR_ARG - 5, "lda < max(1,m)");
  if (!ipiv && n) return c->fail(HM_ERR_ARG - 6, "ipiv is NULL");
  if (!info) return c->fail(HM_ERR_ARG - 7, "info is NULL");
  *info = 0;
  if (n == 0) return HM_OK;
  DevBuf da, dp, di, dc;
  const size_t cplx = 2 * sizeof(double);
  int rc = c->hip(da.ensure(cplx * (size_t)n * n), "zgetrf alloc");
  if (!rc) rc = c->hip(dp.ensure(sizeof(int) * n), "zgetrf alloc");
  if (!rc) rc = c->hip(di.ensure(sizeof(int)), "zgetrf alloc");
  if (!rc) rc = c->hip(dc.ensure(sizeof(unsigned)), "zgetrf alloc");
  if (!rc) rc = c->hip(c->d_luws.ensure(hm::lu_coop_ws_bytes(n)), "zgetrf alloc");
  if (!rc) rc = c->hip(hipMemcpy2DAsync(da.p, cplx * n, a, cplx * lda, cplx * n, n, hipMemcpyHostToDevice, c->stream), "zgetrf");
  if (!rc) rc = c->hip(hipMemsetAsync(di.p, 0, sizeof(int), c->stream), "zgetrf");
  if (!rc) rc = c->hip(hipMemsetAsync(dc.p, 0, sizeof(unsigned), c->stream), "zgetrf");
  if (!rc) rc = c->hip(hm::lu_factor<double>(c->stream, da.as<double>(), n, 0, n, dp.as<int>(), di.as<int>(), dc.as<unsigned>(), 1,
                                                c->luws()), "zgetrf");
  if (!rc) rc = c->hip(hipMemcpy2DAsync(a, cplx * lda, da.p, cplx * n, cplx * n, n, hipMemcpyDeviceToHost, c->stream), "zgetrf");
  if (!rc) rc = c->hip(hipMemcpyAsync(ipiv, dp.p, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream), "zgetrf");
  if (!rc) rc = c->hip(hipMemcpyAsync(info, di.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "zgetrf");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "zgetrf");
  (void)hipStreamSynchronize(c->stream);
  da.release(); dp.release(); di.release(); dc.release();
  if (rc) return rc;
  for (int i = 0; i < n; ++i) ipiv[i] += 1;
  return HM_OK;
}

extern "C" int hm_zgetrs(hm_ctx* c, char trans, int n, int nrhs, const double* a, int lda, const int* ipiv, double* b,
                         int ldb, int* info) {
  if (!c) return HM_ERR_ARG - 1;
  if (trans != 'N' && trans != 'n') return c->fail(HM_ERR_ARG - 2, "only TRANS='N' is supported");
  if (n < 0) return c->fail(HM_ERR_ARG - 3, "n < 0");
  if (nrhs < 0) return c->fail(HM_ERR_ARG - 4, "nrhs < 0");
  if (lda < (n > 1 ? n : 1)) return c->fail(HM_ERR_ARG - 6, "lda < max(1,n)");
  if (ldb < (n > 1 ? n : 1)) return c->fail(HM_ERR_ARG - 9, "ldb < max(1,n)");
  if (!info) return c->fail(HM_ERR_ARG - 10, "info is NULL");
  *info = 0;
  if (n == 0 || nrhs == 0) return HM_OK;
  if (!a || !ipiv || !b) return c->fail(HM_ERR_ARG - 5, "NULL array");
  std::vector<int> p0(n);
  for (int i = 0; i < n; ++i) {
    if (ipiv[i] < 1 || ipiv[i] > n) return c->fail(HM_ERR_ARG - 7, "ipiv[%d] = %d out of range", i, ipiv[i]);
    p0[i] = ipiv[i] - 1;
  }
  DevBuf da, dp, db;
  const size_t cplx = 2 * sizeof(double);
  int rc = c->hip(da.ensure(cplx * (size_t)n * n), "zgetrs alloc");
  if (!rc) rc = c->hip(dp.ensure(sizeof(int) * n), "zgetrs alloc");
  if (!rc) rc = c->hip(db.ensure(cplx * (size_t)n * nrhs), "zgetrs alloc");
  if (!rc) rc = c->hip(hipMemcpy2DAsync(da.p, cplx * n, a, cplx * lda, cplx * n, n, hipMemcpyHostToDevice, c->stream), "zgetrs");
  if (!rc) rc = c->hip(hipMemcpyAsync(dp.p, p0.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream), "zgetrs");
  if (!rc) rc = c->hip(hipMemcpy2DAsync(db.p, cplx * n, b, cplx * ldb, cplx * n, nrhs, hipMemcpyHostToDevice, c->stream), "zgetrs");
  if (!rc) rc = c->hip(hm::lu_solve<double>(c->stream, da.as<double>(), n, 0, n, dp.as<int>(), db.as<double>(), n, 0, nrhs, 1), "zgetrs");
  if (!rc) rc = c->hip(hipMemcpy2DAsync(b, cplx * ldb, db.p, cplx * n, cplx * n, nrhs, hipMemcpyDeviceToHost, c->stream), "zgetrs");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "zgetrs");
  (void)hipStreamSynchronize(c->stream);
  da.release(); dp.release(); db.release();
  return rc;
}
