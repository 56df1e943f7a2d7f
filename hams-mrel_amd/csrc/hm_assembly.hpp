// hm_assembly.hpp -- device-side problem description and assembly launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hm {

constexpr int kNk = 200;          // dispersion roots per frequency (WavDynMods.f90:59)
constexpr int kKeyStride = 128;   // classification key = regime * 128 + series length
constexpr int kNumKeys = 4 * kKeyStride;

// Panel geometry resident in HBM (all bodies combined, panel-major).
struct GeomDev {
  int n, ni, nt, nbody;
  const int* body_start;   // nbody+1 panel offsets
  const double* fp;        // nt*3 field points: hull centroids, then waterplane centroids
  const double* xyz_p;     // n*3 source centroids (= fp)
  const double* verts;     // n*12 panel vertices
  const int* ncn;          // n vertex counts
  const double* ds;        // n areas
  const double* pnsz;      // n longest edges
  const double* dxyz;      // n*6 generalised normals
  // body frames (nullptr: fp, xyz_p and verts are global coordinates). Otherwise
  // they are local to each point's body and org[2 p], org[2 p + 1] is the
  // horizontal placement of field point p's body (p < n: panel p's body)
  const double* org = nullptr;
};

// Field point i as seen from source panel j: with body frames the field point
// moves by the difference of the two bodies' placements, which is exactly zero
// within one body, so that a body's own pairs see only its local coordinates.
__device__ __forceinline__ void field_point(const GeomDev& g, int i, int j, double* xp) {
  const double* x = g.fp + 3 * (long long)i;
  xp[0] = x[0];
  xp[1] = x[1];
  xp[2] = x[2];
  if (g.org) {
    xp[0] = x[0] + (g.org[2 * (long long)i] - g.org[2 * (long long)j]);
    xp[1] = x[1] + (g.org[2 * (long long)i + 1] - g.org[2 * (long long)j + 1]);
  }
}

// a point x given in global coordinates, in the frame of source panel j
__device__ __forceinline__ void global_to_src(const GeomDev& g, int j, const double* x, double* xp) {
  xp[0] = x[0];
  xp[1] = x[1];
  xp[2] = x[2];
  if (g.org) {
    xp[0] = x[0] - g.org[2 * (long long)j];
    xp[1] = x[1] - g.org[2 * (long long)j + 1];
  }
}

// the global position of panel j's centroid (incident-wave phases)
__device__ __forceinline__ void global_centroid(const GeomDev& g, int j, double* x) {
  const double* c = g.xyz_p + 3 * (long long)j;
  x[0] = c[0];
  x[1] = c[1];
  x[2] = c[2];
  if (g.org) {
    x[0] = c[0] + g.org[2 * (long long)j];
    x[1] = c[1] + g.org[2 * (long long)j + 1];
  }
}

// Per-frequency wave state.
struct FreqDev {
  double v, w1, wk, h, rho, g, amp;
  const double* wvn;       // kNk roots on the device
  // optional (nullptr: absent): the evanescent modes' depth factors of every
  // field point, sin and cos of k_m (z_p + h) at evt[2 (m ld + p)], and each
  // mode's normalisation pm_m = 4 / (1 + sin(2 k_m h) / (2 k_m h)) / h at
  // pmt[m], m = 1 .. kNk - 1 (built by evan_table)
  const double* evt = nullptr;
  const double* pmt = nullptr;
  long long evt_ld = 0;
};
hipError_t evan_table(hipStream_t s, const GeomDev& g, const FreqDev& f, double* evt, double* pmt);

// blk > 0 restricts to the diagonal blocks of blk x blk pairs (n a multiple of blk, ni = 0);
// sym keeps PIDCOCK and EIGENE pairs only as (i, j) with i < j (for assemble_sym)
// blist (device, nblist entries): with blk > 0, only those diagonal blocks
hipError_t classify_pairs(hipStream_t s, const GeomDev& g, double h, unsigned* hist, unsigned long long* cursor,
                          uint32_t* pairs, bool count_only, int blk = 0, bool sym = false, const int* blist = nullptr,
                          int nblist = 0);
// PIDCOCK or EIGENE pairs (i, j), i < j, and their mirrors (j, i) together
// (field points = panel centroids, ni = 0): one series evaluation serves both.
// done, vreg (EIGENE, optional): pairs with done[t] != 0 take the regular parts
// of (i, j) and (j, i) from vreg[6 t .. 6 t + 5]
// The frequency-independent halves of the PIDCOCK terms (the i pi / h modes:
// K0, K1 of i pi R / h and the depth cosines) of a PIDCOCK segment sorted by
// series length nt, once per geometry and depth: 4 doubles per (pair, row),
// pair t of length nt at record tab[2 nt + 1] + (t - tab[2 nt]) nt (tab[2 nt]:
// the first pair of length nt; tab: 2 x 128 entries on the device).
struct PidCache {
  double* data;
  const long long* tab;
};
hipError_t pidcock_cache_build(hipStream_t s, const GeomDev& g, double h, const uint32_t* pairs, long long np,
                               PidCache pc);
hipError_t assemble_sym(hipStream_t s, int regime, const GeomDev& g, const FreqDev& f, const uint32_t* pairs,
                        long long np, double* A, double* S, long long ld, const uint8_t* done = nullptr,
                        const float* vreg = nullptr, PidCache pc = PidCache{nullptr, nullptr});
hipError_t assemble_grid(hipStream_t s, const GeomDev& g, const FreqDev& f, bool deep, double* A, double* S,
                         long long ld, int blk = 0, const int* blist = nullptr, int nblist = 0);
// done, vreg (optional, EIGENE only): pairs with done[t] != 0 take the regular
// part V = T - (1/r + 1/r1) of their evanescent sum T from vreg[3 t .. 3 t + 2]
// (value, R- and z-derivative) instead of evaluating the series
hipError_t assemble_regime(hipStream_t s, int regime, const GeomDev& g, const FreqDev& f, const uint32_t* pairs,
                           long long np, double* A, double* S, long long ld, int variant = 1,
                           const uint8_t* done = nullptr, const float* vreg = nullptr);
hipError_t build_w(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* beta, int nbeta, double xw,
                   double yw, double* W, int ldw);
hipError_t forces(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* beta, int nbeta, double xw,
                  double yw, const double* X, int ldx, double* amas, double* bdmp, double* exfc);
// ---- batches of independent problems (one member = one problem) ----
struct BatchProb {
  GeomDev g;                   // the member's geometry (ni waterplane field points after the n hull centroids)
  double *A, *S, *W, *R, *X;   // nt x n matrices and nt x ncol blocks (nt = n + ni), leading dimension ld = nt
  long long ld;
  int ncol, nbeta;
  const double* beta;
  double *amas, *bdmp, *exfc;  // 6nb x 6nb, 6nb x 6nb, 6nb x nbeta
};
struct BatchFreq {             // the member's wave state, rewritten per frequency
  double v, w1, wk, h, rho, g, amp, xw, yw;
  double wvn[kNk];
};
hipError_t batch_classify(hipStream_t s, const BatchProb* P, const BatchFreq* F, int p, int nt, int n, unsigned* hist,
                          unsigned long long* cursor, uint64_t* codes, bool count_only);
hipError_t batch_assemble_regime(hipStream_t s, int regime, const BatchProb* P, const BatchFreq* F,
                                 const uint64_t* codes, long long np, int variant);
hipError_t batch_assemble_all(hipStream_t s, bool deep, const BatchProb* P, const BatchFreq* F, const int* list,
                              const long long* prefix, int nlist, long long total);
hipError_t batch_build_w(hipStream_t s, const BatchProb* P, const BatchFreq* F, int count, int nmax);
hipError_t batch_forces(hipStream_t s, const BatchProb* P, const BatchFreq* F, int count, int mmax, int ncolmax);

hipError_t green_batch(hipStream_t s, int n, bool deep, const double* xf, const double* xp, const double* yf,
                       const double* yp, const double* zf, const double* zp, double v, const double* wvn, double h,
                       const int* flag, double* out);
hipError_t rankine_batch(hipStream_t s, int n, const int* ncn, const double* verts, const double* xf, double* out);
// K0, K1 at n points (out: 2 per point); mode 0 IK01A as the reference, 1 the fast form
hipError_t ik01a_batch(hipStream_t s, int n, const double* x, int mode, double* out);
// field points (CalPotentialMulti): the S- and A-like rows of nfp points xfp
// (3 per point) against the hull panels (nfp x n, column-major, no free term),
// then, on pot = Sf W - Af X (nfp x (m + nbeta)), the free term, the incident
// potential of the diffraction columns, pressure and elevation
hipError_t field_rows(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* xfp, int nfp, double* Af,
                      double* Sf);
hipError_t field_finish(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* xfp, int nfp, int m,
                        int nbeta, const double* beta, double xw, double yw, double* pot, double* prs, double* elv);

}  // namespace hm
