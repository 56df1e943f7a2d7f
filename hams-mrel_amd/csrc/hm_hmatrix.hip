// hm_hmatrix.hip -- cross approximation of the off-diagonal influence blocks
// and the H-matrix products (see hm_hmatrix.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "hm_device.hpp"
#include "hm_hmatrix.hpp"
#include "hm_linalg.hpp"
#include "hm_pairs.hpp"

namespace hm {

namespace {

constexpr int kFinishThreads = 256;

struct Proc {
  int mat, b1, b2;
  double* U;
  double* V;
};
// the process of grid row y: the y-th process that runs (h.run), or y itself
__device__ __forceinline__ int proc_of(const HmatDev& h, int y) { return h.run ? h.run[y] : y; }

__device__ __forceinline__ Proc proc(const HmatDev& h, int p) {
  const int nb = h.nb, per = nb * (nb - 1);
  Proc q;
  q.mat = p / per;
  const int r = p % per;
  q.b1 = r / (nb - 1);
  const int b2p = r % (nb - 1);
  q.b2 = b2p + (b2p >= q.b1 ? 1 : 0);
  const int b1p = q.b1 - (q.b1 > q.b2 ? 1 : 0);
  const long long blk = 2LL * h.m * h.kmax;
  q.U = h.U + ((long long)(q.mat * nb + q.b1) * (nb - 1) + b2p) * blk;
  q.V = h.V + ((long long)(q.mat * nb + q.b2) * (nb - 1) + b1p) * blk;
  return q;
}

__device__ __forceinline__ void entry(const GeomDev& g, const FreqDev& f, const double* wvn, const HmatDev& h,
                                      int mat, int i, int j, double& er, double& ei) {
  if (h.dense) {
    const long long o = (long long)i + (long long)j * h.ldd;
    er = h.dense[2 * o];
    ei = h.dense[2 * o + 1];
    return;
  }
  double ar, ai, sr, si;
  pair_entry(g, f, wvn, i, j, ar, ai, sr, si);
  er = mat ? sr : ar;
  ei = mat ? si : ai;
}

// block-wide argmax of v (first index on ties); v < 0 marks "not a candidate"
__device__ void block_argmax(double v, int idx, double* sv, int* si, double& best, int& bidx) {
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = sv[threadIdx.x + s];
      const int oi = si[threadIdx.x + s];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  best = sv[0];
  bidx = si[0];
  __syncthreads();
}

// block-wide sum of 4 doubles
template <int NT>
__device__ void block_sum4(double* v, double (*sh)[NT]) {
  for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
    __syncthreads();
  }
  for (int c = 0; c < 4; ++c) v[c] = sh[c][0];
  __syncthreads();
}

// residual of one entry: A(b1 m + i, b2 m + j) - sum_{l<k} U(i, l) V(j, l).
// Called by every lane of the workgroup (the Green function's series run in
// lane groups of 16); a lane with active = false gets nothing. LINTON entries
// come out wrong here and are redone by residual_entry_direct (k_aca_fix_*).
template <int G, bool PID, bool DIRECT>
__device__ __forceinline__ void residual_entry(const GeomDev& g, const FreqDev& f, const double* wvn, const HmatDev& h,
                                               const Proc& q, int m, int k, int i, int j, bool active, double& er,
                                               double& ei) {
  if (h.dense) {
    if (!active) return;
    entry(g, f, wvn, h, q.mat, q.b1 * m + i, q.b2 * m + j, er, ei);
  } else {
    double ar, ai, sr, si;
    pair_entry_coop<G, PID, DIRECT>(g, f, wvn, active, q.b1 * m + i, q.b2 * m + j, ar, ai, sr, si);
    if (!active) return;
    er = q.mat ? sr : ar;
    ei = q.mat ? si : ai;
  }
  for (int l = 0; l < k; ++l) {
    const double ur = q.U[2 * (i + (long long)l * m)], ui = q.U[2 * (i + (long long)l * m) + 1];
    const double vr = q.V[2 * (j + (long long)l * m)], vi = q.V[2 * (j + (long long)l * m) + 1];
    er -= ur * vr - ui * vi;
    ei -= ur * vi + ui * vr;
  }
}

// the same one lane at a time, every regime (pair_entry)
__device__ __forceinline__ void residual_entry_direct(const GeomDev& g, const FreqDev& f, const HmatDev& h,
                                                      const Proc& q, int m, int k, int i, int j, double& er,
                                                      double& ei) {
  entry(g, f, f.wvn, h, q.mat, q.b1 * m + i, q.b2 * m + j, er, ei);
  for (int l = 0; l < k; ++l) {
    const double ur = q.U[2 * (i + (long long)l * m)], ui = q.U[2 * (i + (long long)l * m) + 1];
    const double vr = q.V[2 * (j + (long long)l * m)], vi = q.V[2 * (j + (long long)l * m) + 1];
    er -= ur * vr - ui * vi;
    ei -= ur * vi + ui * vr;
  }
}

// the norm estimate's inner products from the entry kernels: with every entry of
// a row or column final after its eval launch (no fix-up launches: bodies at
// least 0.05 h apart, or an assembled matrix) and ranks below kFuseMaxRank
constexpr int kFuseMaxRank = 64;
__device__ __forceinline__ bool aca_fused(const HmatDev& h) {
  return (h.dense != nullptr || h.narrow) && h.kmax <= kFuseMaxRank && h.prow != nullptr;
}

// out[2 l], out[2 l + 1] = sum over this workgroup's lanes of conj(W(x, l)) * (er + i ei),
// l < k (W = V, x = column index of a row launch; W = U, x = row index of a
// column launch); every lane of the workgroup calls it (inactive lanes with 0)
__device__ void aca_partials(const double* W, int x, int m, int k, double er, double ei, double* out) {
  __shared__ double red[4][kFuseMaxRank][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int l = 0; l < k; ++l) {
    const double wr = W[2 * (x + (long long)l * m)], wi = W[2 * (x + (long long)l * m) + 1];
    double pr = wr * er + wi * ei, pi = wr * ei - wi * er;
    for (int o = 32; o > 0; o >>= 1) {
      pr += __shfl_xor(pr, o);
      pi += __shfl_xor(pi, o);
    }
    if (lane == 0) {
      red[wv][l][0] = pr;
      red[wv][l][1] = pi;
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < k; l += blockDim.x) {
    out[2 * l] = ((red[0][l][0] + red[1][l][0]) + red[2][l][0]) + red[3][l][0];
    out[2 * l + 1] = ((red[0][l][1] + red[1][l][1]) + red[2][l][1]) + red[3][l][1];
  }
}

// The entries of one workgroup in the order this lane evaluates: the 256
// entries (row i, columns c0 + t; or rows c0 + t, column j) sorted by near flag
// (the exact Rankine panel integral), regime and series length (bitonic sort in
// LDS), so that the lane groups and waves take entries with equal trip counts.
// Returns the entry index t this lane takes.
__device__ __forceinline__ int sorted_entry(const GeomDev& g, const FreqDev& f, const HmatDev& h, const Proc& q,
                                            bool row, int fixed, int c0, unsigned* sk) {
  const int t = threadIdx.x, m = h.m;
  unsigned key = 0xffffu;                   // past the end: last
  if (c0 + t < m) {
    key = 0;
    if (f.h > 0.0 && !h.dense) {
      const int i = q.b1 * m + (row ? fixed : c0 + t), j = q.b2 * m + (row ? c0 + t : fixed);
      double xp[3];
      field_point(g, i, j, xp);
      const double* xq = g.xyz_p + 3 * (long long)j;
      const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
      const int reg = fin_regime(rr, f.h);
      const int nt = reg == kEigene ? fin_eigene_terms(rr, f.h) : reg == kPidcock ? fin_pidcock_terms(rr, f.h) : 0;
      key = (unsigned)(near_flag(xp, xq, g.pnsz[j]) * 512 + reg * 128 + nt);
    }
  }
  sk[t] = (key << 8) | (unsigned)t;
  __syncthreads();
  for (int k = 2; k <= 256; k <<= 1)
    for (int d = k >> 1; d > 0; d >>= 1) {
      const int o = t ^ d;
      if (o > t) {
        const unsigned a = sk[t], b = sk[o];
        if ((a > b) == ((t & k) == 0)) {
          sk[t] = b;
          sk[o] = a;
        }
      }
      __syncthreads();
    }
  return (int)(sk[t] & 255u);
}

// residual pivot rows of every unconverged process, one entry per thread
// (grid: column chunks x processes)
// the entry kernels are compiled for 4 waves per SIMD (128 VGPRs, ~270 bytes a
// lane of cached spills): 21 % faster than at their natural 208 VGPRs
template <int G, bool PID, bool DIRECT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_aca_eval_row(GeomDev g, FreqDev f,
                                                                                          HmatDev h) {
  const int p = proc_of(h, blockIdx.y);
  if (h.conv[p]) return;
  __shared__ double wvn[kNk];
  if (f.h > 0.0 && !h.dense)
    for (int k = threadIdx.x; k < kNk; k += blockDim.x) wvn[k] = f.wvn[k];
  __syncthreads();
  __shared__ unsigned sk[256];
  const Proc q = proc(h, p);
  const int c0 = blockIdx.x * blockDim.x;
  const int j0 = c0 + sorted_entry(g, f, h, q, true, h.next_row[p], c0, sk);
  const bool active = j0 < h.m;
  const int j = active ? j0 : h.m - 1;
  double er = 0.0, ei = 0.0;
  residual_entry<G, PID, DIRECT>(g, f, wvn, h, q, h.m, h.rank[p], h.next_row[p], j, active, er, ei);
  if (aca_fused(h))
    aca_partials(q.V, j, h.m, h.rank[p], active ? er : 0.0, active ? ei : 0.0,
                 h.prow + 2LL * ((long long)p * gridDim.x + blockIdx.x) * h.kmax);
  if (!active) return;
  double* rb = h.rowbuf + 2LL * p * h.m;
  rb[2 * j] = er;
  rb[2 * j + 1] = ei;
}

// Cross approximation with kSplit lanes per entry (bodies at least 0.05 h
// apart in finite depth, depth-factor tables set): a workgroup takes
// kSplitEntries entries, the evanescent sums of each split over its group of
// lanes (pair_entry_split), so that a launch has kSplit times the lanes of the
// one-lane-per-entry kernels and each lane a kSplit-th of the modes.
constexpr int kSplit = 4, kSplitEntries = 256 / kSplit;

// sorted_entry over the workgroup's kSplitEntries entries (key: near flag,
// regime, number of directly summed modes); returns the entry index of slot e
__device__ __forceinline__ int sorted_entry_split(const GeomDev& g, const FreqDev& f, const HmatDev& h, const Proc& q,
                                                  bool row, int fixed, int c0, unsigned* sk, int e) {
  const int t = threadIdx.x, m = h.m;
  if (t < kSplitEntries) {
    unsigned key = 0xfffu;                   // past the end: last
    if (c0 + t < m) {
      const int i = q.b1 * m + (row ? fixed : c0 + t), j = q.b2 * m + (row ? c0 + t : fixed);
      double xp[3];
      field_point(g, i, j, xp);
      const double* xq = g.xyz_p + 3 * (long long)j;
      const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
      const int reg = fin_regime(rr, f.h);
      const int nt = reg == kEigene ? min(fin_direct_terms(rr, f.h), 255) : 0;
      key = (unsigned)(near_flag(xp, xq, g.pnsz[j]) * 1024 + reg * 256 + nt);
    }
    sk[t] = (key << 8) | (unsigned)t;
  }
  __syncthreads();
  for (int k = 2; k <= kSplitEntries; k <<= 1)
    for (int d = k >> 1; d > 0; d >>= 1) {
      const int o = t ^ d;
      if (t < kSplitEntries && o > t) {
        const unsigned a = sk[t], b = sk[o];
        if ((a > b) == ((t & k) == 0)) {
          sk[t] = b;
          sk[o] = a;
        }
      }
      __syncthreads();
    }
  return (int)(sk[e] & 255u);
}

// residual_entry with pair_entry_split; the rank sum split over the group alike
__device__ __forceinline__ void residual_entry_split(const GeomDev& g, const FreqDev& f, const double* wvn,
                                                     const Proc& q, int m, int k, int i, int j, int sub, double& er,
                                                     double& ei) {
  double ar, ai, sr, si;
  pair_entry_split<kSplit>(g, f, wvn, q.b1 * m + i, q.b2 * m + j, sub, ar, ai, sr, si);
  er = q.mat ? sr : ar;
  ei = q.mat ? si : ai;
  double pr = 0.0, pi = 0.0;
  for (int l = sub; l < k; l += kSplit) {
    const double ur = q.U[2 * (i + (long long)l * m)], ui = q.U[2 * (i + (long long)l * m) + 1];
    const double vr = q.V[2 * (j + (long long)l * m)], vi = q.V[2 * (j + (long long)l * m) + 1];
    pr += ur * vr - ui * vi;
    pi += ur * vi + ui * vr;
  }
#pragma unroll
  for (int o = 1; o < kSplit; o <<= 1) {
    pr += __shfl_xor(pr, o);
    pi += __shfl_xor(pi, o);
  }
  er -= pr;
  ei -= pi;
}

// k_aca_eval_row / k_aca_eval_col with kSplit lanes per entry
template <bool ROW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_aca_eval_split(GeomDev g,
                                                                                            FreqDev f, HmatDev h) {
  const int p = proc_of(h, blockIdx.y);
  if (h.conv[p] || (!ROW && h.skip[p])) return;
  __shared__ double wvn[kNk];
  for (int k = threadIdx.x; k < kNk; k += blockDim.x) wvn[k] = f.wvn[k];
  __syncthreads();
  __shared__ unsigned sk[kSplitEntries];
  const Proc q = proc(h, p);
  const int e = threadIdx.x / kSplit, sub = threadIdx.x % kSplit;
  const int c0 = blockIdx.x * kSplitEntries;
  const int fixed = ROW ? h.next_row[p] : h.pivot[p];
  const int x0 = c0 + sorted_entry_split(g, f, h, q, ROW, fixed, c0, sk, e);
  const bool active = x0 < h.m;
  const int x = active ? x0 : h.m - 1;
  const int k = h.rank[p];
  double er, ei;
  residual_entry_split(g, f, wvn, q, h.m, k, ROW ? fixed : x, ROW ? x : fixed, sub, er, ei);
  const bool lead = active && sub == 0;
  if (aca_fused(h))
    aca_partials(ROW ? q.V : q.U, x, h.m, k, lead ? er : 0.0, lead ? ei : 0.0,
                 (ROW ? h.prow : h.pcol) + 2LL * ((long long)p * gridDim.x + blockIdx.x) * h.kmax);
  if (!lead) return;
  if (ROW) {
    double* rb = h.rowbuf + 2LL * p * h.m;
    rb[2 * x] = er;
    rb[2 * x + 1] = ei;
  } else {
    double* u = q.U + 2LL * k * h.m;
    u[2 * x] = er;
    u[2 * x + 1] = ei;
  }
}

// pivot column j* = argmax |r| over unused columns, v_k = r / r(j*); a zero
// residual row is retired and the pivot row moves to the next unused row
__global__ __launch_bounds__(256) void k_aca_pivot_row(HmatDev h) {
  const int p = proc_of(h, blockIdx.x);
  if (h.conv[p]) return;
  __shared__ double sv[256];
  __shared__ int si[256];
  const Proc q = proc(h, p);
  const int m = h.m, k = h.rank[p], i = h.next_row[p];
  const double* rb = h.rowbuf + 2LL * p * m;
  const int* cu = h.col_used + (long long)p * m;
  double best = -1.0;
  int bj = m;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double a = rb[2 * j] * rb[2 * j] + rb[2 * j + 1] * rb[2 * j + 1];
    if (!cu[j] && a > best) { best = a; bj = j; }
  }
  double emax;
  int js;
  block_argmax(best, bj, sv, si, emax, js);
  if (!(emax > 0.0) || js >= m) {
    if (threadIdx.x == 0) {
      int* ru = h.row_used + (long long)p * m;
      ru[i] = 1;
      int nr = -1;
      for (int t = 1; t <= m; ++t)
        if (!ru[(i + t) % m]) { nr = (i + t) % m; break; }
      if (nr < 0) h.conv[p] = 1;
      else h.next_row[p] = nr;
      h.skip[p] = 1;
    }
    return;
  }
  const double pr = rb[2 * js], pi = rb[2 * js + 1], den = pr * pr + pi * pi;
  const double ir = pr / den, ii = -pi / den;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double er = rb[2 * j], ei = rb[2 * j + 1];
    q.V[2 * (j + (long long)k * m)] = er * ir - ei * ii;
    q.V[2 * (j + (long long)k * m) + 1] = er * ii + ei * ir;
  }
  if (threadIdx.x == 0) {
    h.pivot[p] = js;
    h.col_used[(long long)p * m + js] = 1;
    h.skip[p] = 0;
  }
}

// residual pivot columns u_k = A(:, j*) - U V(j*, :)^T, one entry per thread
template <int G, bool PID, bool DIRECT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_aca_eval_col(GeomDev g, FreqDev f,
                                                                                          HmatDev h) {
  const int p = proc_of(h, blockIdx.y);
  if (h.conv[p] || h.skip[p]) return;
  __shared__ double wvn[kNk];
  if (f.h > 0.0 && !h.dense)
    for (int k = threadIdx.x; k < kNk; k += blockDim.x) wvn[k] = f.wvn[k];
  __syncthreads();
  __shared__ unsigned sk[256];
  const Proc q = proc(h, p);
  const int c0 = blockIdx.x * blockDim.x;
  const int i0 = c0 + sorted_entry(g, f, h, q, false, h.pivot[p], c0, sk);
  const bool active = i0 < h.m;
  const int i = active ? i0 : h.m - 1;
  const int k = h.rank[p];
  double er = 0.0, ei = 0.0;
  residual_entry<G, PID, DIRECT>(g, f, wvn, h, q, h.m, k, i, h.pivot[p], active, er, ei);
  if (aca_fused(h))
    aca_partials(q.U, i, h.m, k, active ? er : 0.0, active ? ei : 0.0,
                 h.pcol + 2LL * ((long long)p * gridDim.x + blockIdx.x) * h.kmax);
  if (!active) return;
  q.U[2 * (i + (long long)k * h.m)] = er;
  q.U[2 * (i + (long long)k * h.m) + 1] = ei;
}

// LINTON entries of the rows / columns just evaluated (R/h < 5e-4: vertically
// aligned panels of two bodies; rare), and PIDCOCK entries when the
// EIGENE-only entry kernels ran, one lane each
__global__ __launch_bounds__(256) void k_aca_fix_row(GeomDev g, FreqDev f, HmatDev h) {
  const int p = proc_of(h, blockIdx.y);
  if (h.dense || h.conv[p]) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= h.m) return;
  const Proc q = proc(h, p);
  const int i = h.next_row[p];
  if (!deferred_entry(g, f, q.b1 * h.m + i, q.b2 * h.m + j, !h.narrow)) return;
  double er, ei;
  residual_entry_direct(g, f, h, q, h.m, h.rank[p], i, j, er, ei);
  double* rb = h.rowbuf + 2LL * p * h.m;
  rb[2 * j] = er;
  rb[2 * j + 1] = ei;
}
__global__ __launch_bounds__(256) void k_aca_fix_col(GeomDev g, FreqDev f, HmatDev h) {
  const int p = proc_of(h, blockIdx.y);
  if (h.dense || h.conv[p] || h.skip[p]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= h.m) return;
  const Proc q = proc(h, p);
  const int k = h.rank[p], j = h.pivot[p];
  if (!deferred_entry(g, f, q.b1 * h.m + i, q.b2 * h.m + j, !h.narrow)) return;
  double er, ei;
  residual_entry_direct(g, f, h, q, h.m, k, i, j, er, ei);
  q.U[2 * (i + (long long)k * h.m)] = er;
  q.U[2 * (i + (long long)k * h.m) + 1] = ei;
}

// norm estimate, stopping test and next pivot row = argmax |u_k| over unused rows
__global__ __launch_bounds__(kFinishThreads) void k_aca_finish_col(HmatDev h) {
  const int p = proc_of(h, blockIdx.x);
  if (h.conv[p] || h.skip[p]) return;
  __shared__ double sh[4][kFinishThreads];
  __shared__ double sv[kFinishThreads];
  __shared__ int si[kFinishThreads];
  const Proc q = proc(h, p);
  const int m = h.m, k = h.rank[p], ir0 = h.next_row[p];
  const double* uk = q.U + 2LL * k * m;
  const double* vk = q.V + 2LL * k * m;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < m; t += blockDim.x) {
    acc[0] += uk[2 * t] * uk[2 * t] + uk[2 * t + 1] * uk[2 * t + 1];
    acc[1] += vk[2 * t] * vk[2 * t] + vk[2 * t + 1] * vk[2 * t + 1];
  }
  block_sum4<kFinishThreads>(acc, sh);
  const double nu2 = acc[0], nv2 = acc[1];
  // |sum_l u_l v_l^T|_F^2 = sum_{l,l'} (u_l^H u_l')(v_l^H v_l'): add the new
  // diagonal term and twice the real part of the cross terms with l < k
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double cross = 0.0;
  if (aca_fused(h)) {
    // from the partial sums of the row and column launches, in chunk order:
    // u_l^H u_k, and v_l^H v_k = (v_l^H r) / r(j*)
    const int nch = (m + h.ech - 1) / h.ech;
    const double* rb = h.rowbuf + 2LL * p * m;
    const int js = h.pivot[p];
    const double pr = rb[2 * js], pi = rb[2 * js + 1], den = pr * pr + pi * pi;
    const double ir = pr / den, ii = -pi / den;
    const double* prow = h.prow + 2LL * (long long)p * nch * h.kmax;
    const double* pcol = h.pcol + 2LL * (long long)p * nch * h.kmax;
    double c = 0.0;
    for (int l = threadIdx.x; l < k; l += blockDim.x) {
      double ur = 0.0, ui = 0.0, rr = 0.0, ri = 0.0;
      for (int ch = 0; ch < nch; ++ch) {
        ur += pcol[2 * ((long long)ch * h.kmax + l)];
        ui += pcol[2 * ((long long)ch * h.kmax + l) + 1];
        rr += prow[2 * ((long long)ch * h.kmax + l)];
        ri += prow[2 * ((long long)ch * h.kmax + l) + 1];
      }
      const double vr = rr * ir - ri * ii, vi = rr * ii + ri * ir;
      c += ur * vr - ui * vi;
    }
    double cc[4] = {c, 0.0, 0.0, 0.0};
    block_sum4<kFinishThreads>(cc, sh);
    cross = cc[0];
  } else {
  for (int l = wv; l < k; l += nw) {
    const double* ul = q.U + 2LL * l * m;
    const double* vl = q.V + 2LL * l * m;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    for (int t = lane; t < m; t += 64) {
      c0 += ul[2 * t] * uk[2 * t] + ul[2 * t + 1] * uk[2 * t + 1];    // conj(u_l) u_k
      c1 += ul[2 * t] * uk[2 * t + 1] - ul[2 * t + 1] * uk[2 * t];
      c2 += vl[2 * t] * vk[2 * t] + vl[2 * t + 1] * vk[2 * t + 1];    // conj(v_l) v_k
      c3 += vl[2 * t] * vk[2 * t + 1] - vl[2 * t + 1] * vk[2 * t];
    }
    for (int o = 32; o > 0; o >>= 1) {
      c0 += __shfl_xor(c0, o);
      c1 += __shfl_xor(c1, o);
      c2 += __shfl_xor(c2, o);
      c3 += __shfl_xor(c3, o);
    }
    cross += c0 * c2 - c1 * c3;
  }
  {
    double c[4] = {lane == 0 ? cross : 0.0, 0.0, 0.0, 0.0};
    block_sum4<kFinishThreads>(c, sh);
    cross = c[0];
  }
  }
  const double n2 = h.norm2[p] + nu2 * nv2 + 2.0 * cross;
  // next pivot row
  const int* ru = h.row_used + (long long)p * m;
  double best = -1.0;
  int bi = m;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double a = uk[2 * i] * uk[2 * i] + uk[2 * i + 1] * uk[2 * i + 1];
    if (!ru[i] && i != ir0 && a > best) { best = a; bi = i; }
  }
  double bmax;
  int bidx;
  block_argmax(best, bi, sv, si, bmax, bidx);
  if (threadIdx.x == 0) {
    h.row_used[(long long)p * m + ir0] = 1;
    h.norm2[p] = n2;
    h.rank[p] = k + 1;
    const bool small = sqrt(nu2 * nv2) < (q.mat ? h.eps_s : h.eps) * sqrt(fmax(n2, 1e-30));
    if (small) h.conv[p] = 1;
    else if (bidx >= m) h.conv[p] = 1;        // every row used: the approximation is exact
    else if (k + 1 >= h.kmax) h.conv[p] = 2;  // rank cap reached before the tolerance
    else h.next_row[p] = bidx;
  }
}

// processes whose block is another process's: marked done (conv = 3) before
// the first step
__global__ void k_aca_share_mark(HmatDev h, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P && h.share[p] != p) h.conv[p] = 3;
}

// after the last step: a shared process takes its representative's factors
// (the first rank columns of U and V; the rest stay zero), rank, norm and state.
// Related by a body symmetry s, block(p)[pi(i), pi(j)] = block(r)[i, j] for
// the symmetry's panel map pi, so row k of p's factors is row pi^-1(k) of r's.
__global__ __launch_bounds__(256) void k_aca_share_copy(HmatDev h) {
  const int p = blockIdx.y;
  const int r = h.share[p];
  if (r == p) return;
  const Proc q = proc(h, p), qr = proc(h, r);
  const long long len = 2LL * h.m * h.rank[r];
  const int sy = h.psym ? h.psym[p] : 0;
  const int* ip = sy ? h.sperm + (long long)sy * h.m : nullptr;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long long)gridDim.x * blockDim.x) {
    long long src = e;
    if (ip) {
      const long long col = e / (2LL * h.m), w = e - col * 2LL * h.m;
      src = col * 2LL * h.m + 2LL * ip[w >> 1] + (w & 1);
    }
    q.U[e] = qr.U[src];
    q.V[e] = qr.V[src];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    h.rank[p] = h.rank[r];
    h.norm2[p] = h.norm2[r];
    h.conv[p] = h.conv[r];
  }
}

__global__ void k_count_open(const int* conv, int n, int* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) mine += conv[p] != 1;
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) *out = cnt;
}

__global__ void k_count_active(const int* conv, int n, int* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) mine += conv[p] == 0;
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) *out = cnt;
}

// Tstack[b1][t, c] = Tbuf[b2][row, c] for tmap[b1 ru + t] = b2 rv + row (0 for -1)
template <typename T>
__global__ void k_permute_map(int nb, int ru, int rv, int ncol, const int* __restrict__ tmap,
                              const T* __restrict__ tb, T* __restrict__ ts) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)ru * ncol;
  if (e >= (long long)nb * per) return;
  const int b1 = (int)(e / per);
  const long long rem = e - b1 * per;
  const int c = (int)(rem / ru), t = (int)(rem - (long long)c * ru);
  const int v = tmap[(long long)b1 * ru + t];
  T re = T(0), im = T(0);
  if (v >= 0) {
    const int b2 = v / rv, row = v - b2 * rv;
    const long long src = ((long long)b2 * ncol + c) * rv + row;
    re = tb[2 * src];
    im = tb[2 * src + 1];
  }
  ts[2 * e] = re;
  ts[2 * e + 1] = im;
}

// block q's first rank columns (r m contiguous complex entries) to their place
// in the packed factors: U blocks by block row (blockIdx.y = b1 (nb-1) + b2'),
// V blocks by block column (blockIdx.y = b2 (nb-1) + b1')
__global__ __launch_bounds__(256) void k_pack_lowrank(HmatDev h, int mat, const int* __restrict__ offu,
                                                      const int* __restrict__ offv, int ru, int rv, double* Uc,
                                                      double* Vc) {
  const int nb = h.nb, per = nb * (nb - 1), q = blockIdx.y;
  const long long blk = 2LL * h.m * h.kmax;
  // U: block row b1, partner b2'
  {
    const int b1 = q / (nb - 1), b2p = q % (nb - 1);
    const int p = mat * per + q;
    const long long len = 2LL * h.rank[p] * h.m;
    const double* src = h.U + ((long long)(mat * nb + b1) * (nb - 1) + b2p) * blk;
    double* dst = Uc + 2LL * ((long long)b1 * h.m * ru + (long long)offu[q] * h.m);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long long)gridDim.x * blockDim.x)
      dst[e] = src[e];
  }
  // V: block column b2, partner b1'
  {
    const int b2 = q / (nb - 1), b1p = q % (nb - 1);
    const int b1 = b1p + (b1p >= b2 ? 1 : 0), b2p = b2 - (b2 > b1 ? 1 : 0);
    const int p = mat * per + b1 * (nb - 1) + b2p;
    const long long len = 2LL * h.rank[p] * h.m;
    const double* src = h.V + ((long long)(mat * nb + b2) * (nb - 1) + b1p) * blk;
    double* dst = Vc + 2LL * ((long long)b2 * h.m * rv + (long long)offv[q] * h.m);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long long)gridDim.x * blockDim.x)
      dst[e] = src[e];
  }
}

}  // namespace

size_t hmat_bytes(int nb, int m, int kmax, int nmat) {
  const size_t P = (size_t)hmat_processes(nb, nmat);
  size_t b = 0;
  b += 2 * 16 * P * m * kmax;      // U, V (complex FP64)
  b += 16 * P * m;                 // rowbuf
  b += 8 * P;                      // norm2
  b += 4 * 5 * P;                  // rank, conv, next_row, pivot, skip
  b += 4 * 2 * P * m;              // row_used, col_used
  b += 4;                          // counter
  b += 2 * 16 * P * (size_t)((m + 63) / 64) * kmax;   // prow, pcol
  return b + 15 * 256;             // alignment of the 15 arrays
}

HmatDev hmat_carve(void* base, int nb, int m, int kmax, double eps, double eps_s, int nmat, const double* dense,
                   long long ldd) {
  const size_t P = (size_t)hmat_processes(nb, nmat);
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~(size_t)255; return q; };
  HmatDev h;
  h.nb = nb;
  h.m = m;
  h.kmax = kmax;
  h.eps = eps;
  h.eps_s = eps_s;
  h.nmat = nmat;
  h.dense = dense;
  h.ldd = ldd;
  h.U = (double*)take(16 * P * m * kmax);
  h.V = (double*)take(16 * P * m * kmax);
  h.rowbuf = (double*)take(16 * P * m);
  h.norm2 = (double*)take(8 * P);
  h.rank = (int*)take(4 * P);
  h.conv = (int*)take(4 * P);
  h.next_row = (int*)take(4 * P);
  h.pivot = (int*)take(4 * P);
  h.skip = (int*)take(4 * P);
  h.row_used = (int*)take(4 * P * m);
  h.col_used = (int*)take(4 * P * m);
  h.counter = (int*)take(4);
  const size_t part = 16 * P * (size_t)((m + 63) / 64) * kmax;
  h.prow = (double*)take(part);
  h.pcol = (double*)take(part);
  h.ech = 256;
  h.narrow = 0;
  h.share = nullptr;
  h.run = nullptr;
  h.nrun = 0;
  h.psym = nullptr;
  h.sperm = nullptr;
  return h;
}

// one step of every open process: residual pivot row, pivot column, residual
// pivot column, norm estimate and next pivot row
static hipError_t aca_step(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h0, int P,
                           const AcaTimer* tm) {
  auto mark = [tm](int b) { if (tm) tm->mark(tm->user, b); };
  // the processes that run: all, or the list h.run of h.nrun
  if (h0.run) P = h0.nrun;
  // bodies at least 0.05 h apart in finite depth with the depth-factor
  // tables: kSplit lanes per entry
  const bool split = h0.narrow && !h0.dense && f.h > 0.0 && f.evt != nullptr;
  HmatDev h = h0;
  h.ech = split ? kSplitEntries : 256;
  if (split) {
    const dim3 ev((h.m + kSplitEntries - 1) / kSplitEntries, P);
    mark(1);
    hipLaunchKernelGGL(k_aca_eval_split<true>, ev, dim3(256), 0, s, g, f, h);
    mark(0);
    hipLaunchKernelGGL(k_aca_pivot_row, dim3(P), dim3(256), 0, s, h);
    mark(1);
    hipLaunchKernelGGL(k_aca_eval_split<false>, ev, dim3(256), 0, s, g, f, h);
    mark(0);
    hipLaunchKernelGGL(k_aca_finish_col, dim3(P), dim3(kFinishThreads), 0, s, h);
    return hipGetLastError();
  }
  const dim3 eval((h.m + 255) / 256, P);
  mark(1);
  if (h.narrow) hipLaunchKernelGGL((k_aca_eval_row<8, false, true>), eval, dim3(256), 0, s, g, f, h);
  else hipLaunchKernelGGL((k_aca_eval_row<16, true, false>), eval, dim3(256), 0, s, g, f, h);
  mark(0);
  // narrow: every pair of bodies is more than 0.05 h apart (bounding circles of
  // the panel centroids), so no entry is LINTON or PIDCOCK and the fix-up
  // kernels would only test and return; skipping them keeps the stream from
  // waiting for CUs behind the other stream's launches for nothing
  if (!h.narrow) hipLaunchKernelGGL(k_aca_fix_row, eval, dim3(256), 0, s, g, f, h);
  hipLaunchKernelGGL(k_aca_pivot_row, dim3(P), dim3(256), 0, s, h);
  mark(1);
  if (h.narrow) hipLaunchKernelGGL((k_aca_eval_col<8, false, true>), eval, dim3(256), 0, s, g, f, h);
  else hipLaunchKernelGGL((k_aca_eval_col<16, true, false>), eval, dim3(256), 0, s, g, f, h);
  mark(0);
  if (!h.narrow) hipLaunchKernelGGL(k_aca_fix_col, eval, dim3(256), 0, s, g, f, h);
  hipLaunchKernelGGL(k_aca_finish_col, dim3(P), dim3(kFinishThreads), 0, s, h);
  return hipGetLastError();
}

hipError_t aca_begin(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int blind, int* steps,
                     const AcaTimer* tm) {
  const int P = hmat_processes(h.nb, h.nmat);
  const size_t uv = 16 * (size_t)P * h.m * h.kmax;
  hipError_t e;
  if ((e = hipMemsetAsync(h.U, 0, uv, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.V, 0, uv, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.norm2, 0, 8 * (size_t)P, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.rank, 0, 4 * (size_t)P, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.conv, 0, 4 * (size_t)P, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.next_row, 0, 4 * (size_t)P, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.skip, 0, 4 * (size_t)P, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.row_used, 0, 4 * (size_t)P * h.m, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(h.col_used, 0, 4 * (size_t)P * h.m, s)) != hipSuccess) return e;
  if (h.share) {
    hipLaunchKernelGGL(k_aca_share_mark, dim3((P + 255) / 256), dim3(256), 0, s, h, P);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // each row step either adds a rank (then a column step) or retires a zero row:
  // at most kmax + m row steps per process
  const int n = blind < h.kmax + h.m ? blind : h.kmax + h.m;
  for (int st = 0; st < n; ++st)
    if ((e = aca_step(s, g, f, h, P, tm)) != hipSuccess) return e;
  *steps = n;
  return hipSuccess;
}

hipError_t aca_end(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int* steps, int* unconverged,
                   const AcaTimer* tm) {
  const int P = hmat_processes(h.nb, h.nmat);
  int* cnt = h.counter;
  hipError_t e;
  int st = *steps;
  const int max_steps = h.kmax + h.m;
  // further steps while a process is open (a process stops at its tolerance,
  // at rank kmax, or when every row is used): polled every 4 steps
  for (;;) {
    int open = 0;
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, s, h.conv, P, cnt);
    if ((e = hipMemcpyAsync(&open, cnt, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (open == 0 || st >= max_steps) break;
    for (int q = 0; q < 4 && st < max_steps; ++q, ++st)
      if ((e = aca_step(s, g, f, h, P, tm)) != hipSuccess) return e;
  }
  if (h.share) {
    hipLaunchKernelGGL(k_aca_share_copy, dim3(64, P), dim3(256), 0, s, h);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_count_open, dim3(1), dim3(256), 0, s, h.conv, P, cnt);
  int bad = 0;
  if ((e = hipMemcpyAsync(&bad, cnt, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (unconverged) *unconverged = bad;
  *steps = st;
  return hipSuccess;
}

hipError_t aca_build(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int* unconverged,
                     int* steps, const AcaTimer* tm) {
  int st = 0;
  hipError_t e = aca_begin(s, g, f, h, 0, &st, tm);
  if (e != hipSuccess) return e;
  e = aca_end(s, g, f, h, &st, unconverged, tm);
  if (steps) *steps = st;
  return e;
}

void lowrank_plan(int nb, const int* rank, LowRankPlan& plan) {
  const int nq = nb * (nb - 1);
  plan.offu.assign(nq, 0);
  plan.offv.assign(nq, 0);
  plan.ru = plan.rv = 0;
  std::vector<int> rowsum(nb, 0), colsum(nb, 0);
  for (int b1 = 0; b1 < nb; ++b1)
    for (int b2p = 0; b2p < nb - 1; ++b2p) {
      const int q = b1 * (nb - 1) + b2p;
      plan.offu[q] = rowsum[b1];
      rowsum[b1] += rank[q];
    }
  for (int b2 = 0; b2 < nb; ++b2)
    for (int b1p = 0; b1p < nb - 1; ++b1p) {
      const int b1 = b1p + (b1p >= b2 ? 1 : 0), b2p = b2 - (b2 > b1 ? 1 : 0);
      const int q = b2 * (nb - 1) + b1p;
      plan.offv[q] = colsum[b2];
      colsum[b2] += rank[b1 * (nb - 1) + b2p];
    }
  for (int b = 0; b < nb; ++b) {
    plan.ru = std::max(plan.ru, rowsum[b]);
    plan.rv = std::max(plan.rv, colsum[b]);
  }
  plan.tmap.assign((size_t)nb * std::max(plan.ru, 1), -1);
  for (int b1 = 0; b1 < nb; ++b1)
    for (int b2p = 0; b2p < nb - 1; ++b2p) {
      const int b2 = b2p + (b2p >= b1 ? 1 : 0), b1p = b1 - (b1 > b2 ? 1 : 0);
      const int q = b1 * (nb - 1) + b2p, r = rank[q];
      const int vrow = plan.offv[b2 * (nb - 1) + b1p];
      for (int k = 0; k < r; ++k) plan.tmap[(size_t)b1 * plan.ru + plan.offu[q] + k] = b2 * plan.rv + vrow + k;
    }
}

hipError_t hmat_pack(hipStream_t s, const HmatDev& h, int mat, const LowRankPlan& plan, int* meta, double* Uc,
                     double* Vc, LowRank* out) {
  const int nb = h.nb, nq = nb * (nb - 1);
  out->ru = plan.ru;
  out->rv = plan.rv;
  out->tmap = meta + 2 * nq;
  hipError_t e;
  if ((e = hipMemcpyAsync(meta, plan.offu.data(), sizeof(int) * nq, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(meta + nq, plan.offv.data(), sizeof(int) * nq, hipMemcpyHostToDevice, s)) != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(meta + 2 * nq, plan.tmap.data(), sizeof(int) * plan.tmap.size(), hipMemcpyHostToDevice, s)) !=
      hipSuccess)
    return e;
  if (plan.ru > 0 && plan.rv > 0) {
    if ((e = hipMemsetAsync(Uc, 0, 2 * sizeof(double) * (size_t)nb * h.m * plan.ru, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(Vc, 0, 2 * sizeof(double) * (size_t)nb * h.m * plan.rv, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pack_lowrank, dim3(16, nq), dim3(256), 0, s, h, mat, meta, meta + nq, plan.ru, plan.rv, Uc,
                       Vc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // the plan's host vectors may change after the return
  return hipStreamSynchronize(s);
}

// Wc[i, b q + t] = X[b m + i, column t of body b]: its own 6 radiation
// columns (t < 6), then the nbeta diffraction ones (q = 6 + nbeta)
__global__ void k_rhs_gather(int nb, int m, int nbeta, const double* __restrict__ X, long long ldx,
                             double* __restrict__ xc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, bt = blockIdx.y;
  if (i >= m) return;
  const int q = 6 + nbeta, b = bt / q, t = bt % q;
  const int c = t < 6 ? 6 * b + t : 6 * nb + (t - 6);
  const long long src = (long long)b * m + i + (long long)c * ldx, dst = (long long)i + (long long)bt * m;
  xc[2 * dst] = X[2 * src];
  xc[2 * dst + 1] = X[2 * src + 1];
}

// Y = the products back in place: radiation column c of body b's rows from
// body b's own columns, zero for the other bodies' radiation columns
__global__ void k_rhs_scatter(int nb, int m, int nbeta, const double* __restrict__ yc, double* __restrict__ Y,
                              long long ldy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= nb * m) return;
  const int q = 6 + nbeta, b = r / m, i = r % m;
  double yr = 0.0, yi = 0.0;
  const int t = c < 6 * nb ? (c / 6 == b ? c % 6 : -1) : 6 + (c - 6 * nb);
  if (t >= 0) {
    const long long src = (long long)i + (long long)(b * q + t) * m;
    yr = yc[2 * src];
    yi = yc[2 * src + 1];
  }
  const long long dst = (long long)r + (long long)c * ldy;
  Y[2 * dst] = yr;
  Y[2 * dst + 1] = yi;
}

hipError_t hmat_diag_rhs(hipStream_t s, int nb, int m, int nbeta, const double* D, long long ldd, const double* X,
                         long long ldx, double* Y, long long ldy, double* xc, double* yc) {
  const int q = 6 + nbeta;
  hipLaunchKernelGGL(k_rhs_gather, dim3((m + 255) / 256, nb * q), dim3(256), 0, s, nb, m, nbeta, X, ldx, xc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = cgemm_3m<double>(s, false, m, nb * q, m, 1.0, 0.0, D, (int)ldd, 0, xc, m, 0, 0.0, 0.0, yc, m, 0, 1)) !=
      hipSuccess)
    return e;
  hipLaunchKernelGGL(k_rhs_scatter, dim3((nb * m + 255) / 256, 6 * nb + nbeta), dim3(256), 0, s, nb, m, nbeta, yc, Y,
                     ldy);
  return hipGetLastError();
}

template <typename T>
hipError_t hmat_apply(hipStream_t s, int nb, int m, const LowRank& lr, const T* D, long long ldd, long long sD,
                      const T* U, const T* V, const T* X, long long ldx, T* Y, long long ldy, int ncol, T* tbuf,
                      T* tstack) {
  hipError_t e;
  // diagonal blocks (D null: Y holds the caller's values and only the
  // off-diagonal part is added). sD = 0, one block for all bodies: with X and Y
  // of leading dimension nb m, block b's rows of column c are column b + c nb
  // of X and Y read as m x (nb ncol) matrices of leading dimension m, so the
  // product is one GEMM over all bodies' columns (the same sums per entry).
  // Every product takes three real products per complex one: in FP64 (the
  // residuals) the error stays at FP64's, far below the GMRES tolerance; in
  // FP32 (the Arnoldi operator) far below the FP16 preconditioner's rounding
  auto gemm = [&](bool ta, int gm, int gn, int gk, const T* A, int lda, long long sA, const T* B, int ldb,
                  long long sB, T beta, T* C, int ldc, long long sC, int cnt) {
    return cgemm_3m<T>(s, ta, gm, gn, gk, T(1), T(0), A, lda, sA, B, ldb, sB, beta, T(0), C, ldc, sC, cnt);
  };
  if (D && sD == 0 && ldx == (long long)nb * m && ldy == ldx)
    e = gemm(false, m, nb * ncol, m, D, (int)ldd, 0, X, m, 0, T(0), Y, m, 0, 1);
  else
    e = D ? gemm(false, m, ncol, m, D, (int)ldd, sD, X, (int)ldx, m, T(0), Y, (int)ldy, m, nb) : hipSuccess;
  if (e != hipSuccess || lr.ru == 0 || lr.rv == 0) return e;
  // T_b2 = V_b2^T x_b2 for every block column b2
  e = gemm(true, lr.rv, ncol, m, V, m, (long long)lr.rv * m, X, (int)ldx, m, T(0), tbuf, lr.rv,
           (long long)lr.rv * ncol, nb);
  if (e != hipSuccess) return e;
  const long long tot = (long long)nb * lr.ru * ncol;
  hipLaunchKernelGGL((k_permute_map<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nb, lr.ru, lr.rv, ncol,
                     lr.tmap, tbuf, tstack);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // y_b1 += U_b1 [T_{b1,b2}]_b2
  return gemm(false, m, ncol, lr.ru, U, m, (long long)lr.ru * m, tstack, lr.ru, (long long)lr.ru * ncol, T(1), Y,
              (int)ldy, m, nb);
}

template hipError_t hmat_apply<double>(hipStream_t, int, int, const LowRank&, const double*, long long, long long,
                                       const double*, const double*, const double*, long long, double*, long long, int,
                                       double*, double*);
template hipError_t hmat_apply<float>(hipStream_t, int, int, const LowRank&, const float*, long long, long long,
                                      const float*, const float*, const float*, long long, float*, long long, int,
                                      float*, float*);

}  // namespace hm
