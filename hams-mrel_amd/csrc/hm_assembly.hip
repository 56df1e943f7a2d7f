// hm_assembly.hip -- influence-matrix assembly and force reductions on gfx950.
//
// Replaces the per-frequency table fill CALGREEN_MULTI (CalGreenFuncMulti.f90:112)
// followed by the element integrations BODINT/RBC/DBC (BodyIntgrMulti.f90:48-173)
// and the fills of ASSB_LEFT/RBC/DBC (AssbMatxMulti.f90:53-251; IRR rows:
// AssbMatx_irr.f90:53-466). The 4-plane RKBN/CGRN tables are never stored: one
// thread evaluates the Rankine panel integral (near pairs only) and the wave
// Green function of its (field point i, source panel j) pair and writes
//   A[i,j] = (i==j ? 2*pi : 0) + sum_k (dS_k + dG_k*ds_j) n_jk       (flag = 1)
//                                 or (sum_k dG_k n_jk) * ds_j          (flag = 0)
//   S[i,j] = S_ij + G*ds_j  (flag = 1)   or   G*ds_j  (flag = 0)
// straight into the column-major system matrix A and single-layer matrix S.
// The radiation/diffraction right-hand sides are then S*W (a GEMM) with W the
// per-panel normal components and incident normal velocities.
//
// Finite depth: pairs are classified once per geometry by regime and series
// length (both depend on R/h only) and each regime runs its own kernel over a
// contiguous pair list sorted by series length, so that lanes of a wave follow
// the same branch and trip count.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hm_assembly.hpp"
#include "hm_coop.hpp"
#include "hm_device.hpp"
#include "hm_pairs.hpp"

namespace hm {

__device__ __forceinline__ void load_wvn(double* s, const double* g, int nk) {
  for (int k = threadIdx.x; k < nk; k += blockDim.x) s[k] = g[k];
  __syncthreads();
}

// one pair: field point i (hull centroid, or waterplane centroid when i >= n),
// source panel j; REG = regime (or -1 for infinite depth, -2 for generic dispatch)
template <int REG>
__device__ __forceinline__ void assemble_pair(const GeomDev& g, const FreqDev& f, const double* wvn, int i, int j,
                                              double* A, double* S, long long ld) {
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  if (REG == -1) {
    infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gr);
  } else if (REG == -2) {
    fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, wvn, f.h, flag, gr);
  } else {
    const double xpq = xq[0] - xp[0], ypq = xq[1] - xp[1];
    const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
    FinOut fo{0, 0, 0, 0, 0, 0};
    if (REG == kLinton) fin_linton(rr, xq[2], xp[2], f.v, wvn, f.h, fo);
    else if (REG == kPidcock) fin_pidcock(rr, xq[2], xp[2], f.v, wvn, f.h, fo);
    else if (REG == kEigene) fin_eigene(rr, xq[2], xp[2], wvn, f.h, fo);
    else fin_eigen(rr, xq[2], xp[2], wvn, f.h, fo);
    if (rr == 0.0) { fo.gr_re = 0.0; fo.gr_im = 0.0; }
    fin_to_grn(fo, xpq, ypq, xq[2], xp[2], rr, true, flag, gr);
  }
  write_pair_g(g, i, j, xp, flag, gr, A, S, ld);
}

// 2-D grid over all pairs (infinite depth, or finite depth with dispatch per pair)
template <int REG>
__global__ __launch_bounds__(256) void k_assemble_grid(GeomDev g, FreqDev f, double* A, double* S, long long ld,
                                                       int blk, const int* __restrict__ blist) {
  __shared__ double wvn[kNk];
  if (REG != -1) load_wvn(wvn, f.wvn, kNk);
  // blk > 0: only the diagonal blocks of blk x blk pairs (blockIdx.z = block,
  // or blist[blockIdx.z])
  const int b0 = blk > 0 ? (blist ? blist[blockIdx.z] : (int)blockIdx.z) * blk : 0;
  const int i = b0 + blockIdx.y * 16 + (threadIdx.x >> 4);
  const int j = b0 + blockIdx.x * 16 + (threadIdx.x & 15);
  if (i >= (blk > 0 ? b0 + blk : g.nt) || j >= (blk > 0 ? b0 + blk : g.n)) return;
  assemble_pair<REG>(g, f, wvn, i, j, A, S, ld);
}

// the rest of one EIGENE / PIDCOCK pair once its series limits L are known
// (regular = true: L is the regular part V = T - (1/r + 1/r1) of an EIGENE
// pair, the 1/r + 1/r1 strip already applied)
template <int REG>
__device__ __forceinline__ void coop_finish(const GeomDev& g, const double* wvn, double h, int i, int j, double rr,
                                            double L1, double L2, double L3, double* A, double* S, long long ld,
                                            bool regular = false) {
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double zf = xq[2], zp = xp[2];
  FinOut fo{0, 0, 0, 0, 0, 0};
  if (REG == kEigene) {
    fin_propagating(rr, zf + h, zp + h, wvn[0], h, fo);
    fo.g_re += L1;
    fo.gr_re += L2;
    fo.gz_re += L3;
    if (!regular) fin_strip(rr, zf, zp, fo);
  } else {
    fin_pidcock_head(rr, zf, zp, wvn, h, fo);
    fin_pidcock_tail(L1, L2, L3, rr, zf, zp, h, fo);
  }
  if (rr == 0.0) { fo.gr_re = 0.0; fo.gr_im = 0.0; }
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  fin_to_grn(fo, xq[0] - xp[0], xq[1] - xp[1], zf, zp, rr, true, flag, gr);
  write_pair_g(g, i, j, xp, flag, gr, A, S, ld);
}

// one regime over its classified pair list (i in the high 16 bits, j in the
// low). EIGENE with a mask: pairs with done[t] != 0 take the regular part
// vreg[3 t .. 3 t + 2] instead of the series.
template <int REG>
__global__ __launch_bounds__(256) void k_assemble_list(GeomDev g, FreqDev f, const uint32_t* __restrict__ pairs,
                                                       long long np, double* A, double* S, long long ld,
                                                       const uint8_t* __restrict__ done,
                                                       const float* __restrict__ vreg) {
  __shared__ double wvn[kNk];
  load_wvn(wvn, f.wvn, kNk);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const uint32_t p = pairs[t];
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  if (REG == kEigene && done && done[t]) {
    double xp[3];
    field_point(g, i, j, xp);
    const double* xq = g.xyz_p + 3 * (long long)j;
    coop_finish<kEigene>(g, wvn, f.h, i, j, horiz_dist(xq[0], xp[0], xq[1], xp[1]), vreg[3 * t], vreg[3 * t + 1],
                         vreg[3 * t + 2], A, S, ld, true);
    return;
  }
  assemble_pair<REG>(g, f, wvn, i, j, A, S, ld);
}

// The epsilon-accelerated regimes with G lanes per pair (see hm_coop.hpp): each
// lane owns one pair of the list; the group evaluates the series and their
// epsilon limits for its G pairs in turn, then every lane finishes its own pair.

// The FP64 series are latency bound: the default shapes (K <= 7 register
// slots) are compiled for 5 waves per SIMD (96 VGPRs, a few hundred bytes a lane
// of spills that stay in cache), measured 29 % faster than at the 2 waves their
// natural 200 VGPRs allow; the K = 13 shapes keep the default (they spill
// heavily below 256 VGPRs).
template <int K>
struct CoopWaves { static constexpr int value = K <= 7 ? 5 : 1; };

template <int REG, int G, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CoopWaves<K>::value, 8))) void k_assemble_coop(GeomDev g, FreqDev f, const uint32_t* __restrict__ pairs,
                                                       long long np, double* A, double* S, long long ld,
                                                       const uint8_t* __restrict__ done,
                                                       const float* __restrict__ vreg) {
  __shared__ double wvn[kNk];
  load_wvn(wvn, f.wvn, kNk);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < np;
  // a pair whose regular part is given (masked EIGENE list) runs no series
  const bool given = REG == kEigene && valid && done && done[t];
  const uint32_t p = pairs[valid ? t : np - 1];
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2];
  // no series for lanes without a pair to evaluate (end of list, or done)
  const int nt = !valid || given ? 0 : REG == kEigene ? fin_eigene_terms(rr, h) : fin_pidcock_terms(rr, h);
  const int l = threadIdx.x & (G - 1);
  double L1 = 0.0, L2 = 0.0, L3 = 0.0;
  for (int q = 0; q < G; ++q) {
    const double r = group_bcast<G>(rr, q);
    const double zfh = group_bcast<G>(zf, q) + h, zph = group_bcast<G>(zp, q) + h;
    const int ntq = group_bcast<G>(nt, q);
    if (ntq == 0) continue;                  // uniform over the group
    double l1, l2, l3;
    coop_series<REG, G, K>(ntq, r, zfh, zph, h, f.v, wvn, l, l1, l2, l3);
    if (l == q) { L1 = l1; L2 = l2; L3 = l3; }
  }
  if (!valid) return;
  if (given) {
    L1 = vreg[3 * t];
    L2 = vreg[3 * t + 1];
    L3 = vreg[3 * t + 2];
  }
  coop_finish<REG>(g, wvn, h, i, j, rr, L1, L2, L3, A, S, ld, given);
}

// PIDCOCK or EIGENE pairs (i, j), i < j, of blocks whose field points are
// their panels' centroids (the H-matrix diagonal blocks): the group evaluates
// the series of the pair and of its mirror together (four series, four limits
// instead of six and six), then every lane finishes (i, j) and (j, i). EIGENE
// with a mask: pairs with done[t] != 0 take the regular parts of both
// directions from vreg[6 t .. 6 t + 5] instead.
template <int REG, int G, int K, bool CACHED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CoopWaves<K>::value, 8))) void
k_assemble_sym(GeomDev g, FreqDev f, const uint32_t* __restrict__ pairs, long long np, double* A, double* S,
               long long ld, const uint8_t* __restrict__ done, const float* __restrict__ vreg, PidCache pc) {
  __shared__ double wvn[kNk];
  load_wvn(wvn, f.wvn, kNk);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < np;
  const bool given = REG == kEigene && valid && done && done[t];
  const uint32_t p = pairs[valid ? t : np - 1];
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2];
  const int nt = !valid || given ? 0 : REG == kEigene ? fin_eigene_terms(rr, h) : fin_pidcock_terms(rr, h);
  // the pair's record in the PIDCOCK cache (rows of nt x 4 doubles; pairs sorted by nt)
  long long rec = 0;
  if (CACHED && valid) rec = pc.tab[2 * nt + 1] + (t - pc.tab[2 * nt]) * nt;
  const int l = threadIdx.x & (G - 1);
  double L1 = 0.0, L2 = 0.0, L3 = 0.0, L4 = 0.0;
  for (int q = 0; q < G; ++q) {
    const int ntq = group_bcast<G>(nt, q);
    if (ntq == 0) continue;                  // uniform over the group
    const double r = group_bcast<G>(rr, q);
    const double zfh = group_bcast<G>(zf, q) + h, zph = group_bcast<G>(zp, q) + h;
    const double* tc = nullptr;
    if (CACHED) {
      const unsigned lo = __shfl((unsigned)(rec & 0xffffffffll), q, G), hi = __shfl((unsigned)(rec >> 32), q, G);
      tc = pc.data + 4 * (long long)(((unsigned long long)hi << 32) | lo);
    }
    double l1, l2, l3, l4;
    coop_series_sym<REG, G, K, CACHED>(ntq, r, zfh, zph, h, f.v, wvn, l, l1, l2, l3, l4, tc);
    if (l == q) { L1 = l1; L2 = l2; L3 = l3; L4 = l4; }
  }
  if (!valid) return;
  if (given) {
    const float* v = vreg + 6 * t;
    coop_finish<REG>(g, wvn, h, i, j, rr, v[0], v[1], v[2], A, S, ld, true);
    coop_finish<REG>(g, wvn, h, j, i, rr, v[3], v[4], v[5], A, S, ld, true);
    return;
  }
  coop_finish<REG>(g, wvn, h, i, j, rr, L1, L2, L3, A, S, ld);
  coop_finish<REG>(g, wvn, h, j, i, rr, L1, L2, L4, A, S, ld);
}

// the i pi / h halves of every term of every pair of a PIDCOCK segment sorted
// by series length (thread per pair and row)
__global__ __launch_bounds__(256) void k_pidcock_cache(GeomDev g, double h, const uint32_t* __restrict__ pairs,
                                                       long long np, PidCache pc) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = e / kPidcockMaxTerms;
  const int row = (int)(e % kPidcockMaxTerms);
  if (t >= np) return;
  const uint32_t p = pairs[t];
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const int nt = fin_pidcock_terms(rr, h);
  if (row < 1 || row >= nt) return;
  const long long rec = pc.tab[2 * nt + 1] + (t - pc.tab[2 * nt]) * nt;
  fin_pidcock_tilde(row, rr, xq[2] + h, xp[2] + h, h, pc.data + 4 * (rec + row));
}

hipError_t pidcock_cache_build(hipStream_t s, const GeomDev& g, double h, const uint32_t* pairs, long long np,
                               PidCache pc) {
  if (np <= 0 || !pc.data) return hipSuccess;
  const long long work = np * kPidcockMaxTerms;
  hipLaunchKernelGGL(k_pidcock_cache, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, g, h, pairs, np, pc);
  return hipGetLastError();
}

hipError_t assemble_sym(hipStream_t s, int regime, const GeomDev& g, const FreqDev& f, const uint32_t* pairs,
                        long long np, double* A, double* S, long long ld, const uint8_t* done, const float* vreg,
                        PidCache pc) {
  if (np <= 0) return hipSuccess;
  const dim3 gd((unsigned)((np + 255) / 256)), bd(256);
  if (regime == kPidcock && pc.data)
    hipLaunchKernelGGL((k_assemble_sym<kPidcock, 16, 7, true>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, nullptr,
                       nullptr, pc);
  else if (regime == kPidcock)
    hipLaunchKernelGGL((k_assemble_sym<kPidcock, 16, 7, false>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, nullptr,
                       nullptr, pc);
  else if (regime == kEigene)
    hipLaunchKernelGGL((k_assemble_sym<kEigene, 8, 7, false>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg,
                       PidCache{nullptr, nullptr});
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---- classification --------------------------------------------------------
__device__ __forceinline__ int pair_key(const GeomDev& g, double h, int i, int j) {
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const int reg = fin_regime(rr, h);
  int len = 0;
  if (reg == kEigene) len = fin_eigene_terms(rr, h);
  else if (reg == kPidcock) len = fin_pidcock_terms(rr, h);
  return reg * kKeyStride + len;
}

// sym: PIDCOCK and EIGENE pairs are kept only as (i, j) with i < j (their
// mirrors are assembled with them, k_assemble_sym)
__device__ __forceinline__ bool mirrored_out(bool sym, int key, int i, int j) {
  const int reg = key / kKeyStride;
  return sym && (reg == kPidcock || reg == kEigene) && i > j;
}

__global__ __launch_bounds__(256) void k_classify_count(GeomDev g, double h, unsigned* hist, int blk, int sym,
                                                        const int* __restrict__ blist) {
  __shared__ unsigned sh[kNumKeys];
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  const int b0 = blk > 0 ? (blist ? blist[blockIdx.z] : (int)blockIdx.z) * blk : 0;
  const int i = b0 + blockIdx.y * 16 + (threadIdx.x >> 4);
  const int j = b0 + blockIdx.x * 16 + (threadIdx.x & 15);
  if (i < (blk > 0 ? b0 + blk : g.nt) && j < (blk > 0 ? b0 + blk : g.n)) {
    const int key = pair_key(g, h, i, j);
    if (!mirrored_out(sym != 0, key, i, j)) atomicAdd(&sh[key], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x)
    if (sh[k]) atomicAdd(&hist[k], sh[k]);
}

// one global atomic per (workgroup, key): ranks within the workgroup come from LDS
__global__ __launch_bounds__(256) void k_classify_scatter(GeomDev g, double h, unsigned long long* cursor,
                                                          uint32_t* pairs, int blk, int sym,
                                                          const int* __restrict__ blist) {
  __shared__ unsigned cnt[kNumKeys];
  __shared__ unsigned long long base[kNumKeys];
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) cnt[k] = 0;
  __syncthreads();
  const int b0 = blk > 0 ? (blist ? blist[blockIdx.z] : (int)blockIdx.z) * blk : 0;
  const int i = b0 + blockIdx.y * 16 + (threadIdx.x >> 4);
  const int j = b0 + blockIdx.x * 16 + (threadIdx.x & 15);
  const bool inb = i < (blk > 0 ? b0 + blk : g.nt) && j < (blk > 0 ? b0 + blk : g.n);
  const int key = inb ? pair_key(g, h, i, j) : 0;
  const bool in = inb && !mirrored_out(sym != 0, key, i, j);
  const unsigned rank = in ? atomicAdd(&cnt[key], 1u) : 0u;
  __syncthreads();
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x)
    if (cnt[k]) base[k] = atomicAdd(&cursor[k], (unsigned long long)cnt[k]);
  __syncthreads();
  if (in) pairs[base[key] + rank] = ((uint32_t)i << 16) | (uint32_t)j;
}

static dim3 pair_grid(const GeomDev& g, int blk, int nblk = 0) {
  if (blk > 0) return dim3((blk + 15) / 16, (blk + 15) / 16, nblk > 0 ? nblk : g.n / blk);
  return dim3((g.n + 15) / 16, (g.nt + 15) / 16);
}

hipError_t classify_pairs(hipStream_t s, const GeomDev& g, double h, unsigned* hist,
                          unsigned long long* cursor, uint32_t* pairs, bool count_only, int blk, bool sym,
                          const int* blist, int nblist) {
  if (blist && (blk <= 0 || nblist <= 0)) return hipErrorInvalidValue;
  const dim3 grid = pair_grid(g, blk, blist ? nblist : 0);
  if (count_only) hipLaunchKernelGGL(k_classify_count, grid, dim3(256), 0, s, g, h, hist, blk, sym ? 1 : 0, blist);
  else hipLaunchKernelGGL(k_classify_scatter, grid, dim3(256), 0, s, g, h, cursor, pairs, blk, sym ? 1 : 0, blist);
  return hipGetLastError();
}

hipError_t assemble_grid(hipStream_t s, const GeomDev& g, const FreqDev& f, bool deep, double* A, double* S,
                         long long ld, int blk, const int* blist, int nblist) {
  if (blist && (blk <= 0 || nblist <= 0)) return hipErrorInvalidValue;
  const dim3 grid = pair_grid(g, blk, blist ? nblist : 0);
  if (deep) hipLaunchKernelGGL((k_assemble_grid<-1>), grid, dim3(256), 0, s, g, f, A, S, ld, blk, blist);
  else hipLaunchKernelGGL((k_assemble_grid<-2>), grid, dim3(256), 0, s, g, f, A, S, ld, blk, blist);
  return hipGetLastError();
}

hipError_t assemble_regime(hipStream_t s, int regime, const GeomDev& g, const FreqDev& f, const uint32_t* pairs,
                           long long np, double* A, double* S, long long ld, int variant, const uint8_t* done,
                           const float* vreg) {
  if (np <= 0) return hipSuccess;
  const dim3 gd((unsigned)((np + 255) / 256)), bd(256);
  switch (regime) {
    case kLinton: hipLaunchKernelGGL((k_assemble_list<kLinton>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg); break;
    case kPidcock:
      if (variant == 1) hipLaunchKernelGGL((k_assemble_coop<kPidcock, 16, 7>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      else if (variant == 2) hipLaunchKernelGGL((k_assemble_coop<kPidcock, 8, 13>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      else hipLaunchKernelGGL((k_assemble_list<kPidcock>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      break;
    case kEigene:
      if (variant == 1) hipLaunchKernelGGL((k_assemble_coop<kEigene, 8, 7>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      else if (variant == 2) hipLaunchKernelGGL((k_assemble_coop<kEigene, 4, 13>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      else hipLaunchKernelGGL((k_assemble_list<kEigene>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg);
      break;
    default: hipLaunchKernelGGL((k_assemble_list<kEigen>), gd, bd, 0, s, g, f, pairs, np, A, S, ld, done, vreg); break;
  }
  return hipGetLastError();
}

// ---- incident wave (PatclVelct.f90) -----------------------------------------
struct Incident { double w0, wk0; };
__device__ __forceinline__ Incident incident_params(const FreqDev& f) {
  if (fabs(f.v) < F32(1.e-8)) return {F32(1.e-20), F32(1.e-20)};
  if (fabs(f.v + 1.0) < F32(1.e-8)) return {F32(1.e+20), F32(1.e+20)};
  return {f.w1, f.wk};
}

// DINP . n at source panel j for heading beta (DINP :83-123); VINP-style deep
// test (H < 0) as in the reference
__device__ void dinp_dot_n(const FreqDev& f, const double* x, const double* n, double beta, double xw, double yw,
                           double& qr, double& qi) {
  const Incident p = incident_params(f);
  double sb, cb;
  sincos(beta, &sb, &cb);
  const double wkx = p.wk0 * ((x[0] - xw) * cb + (x[1] - yw) * sb);
  double se, ce;
  sincos(wkx, &se, &ce);
  double hx, vz;   // horizontal and vertical depth factors
  if (f.h < 0.0) { hx = exp(p.wk0 * x[2]); vz = hx; }
  else {
    hx = cosh(p.wk0 * (x[2] + f.h)) / sinh(p.wk0 * f.h);
    vz = sinh(p.wk0 * (x[2] + f.h)) / sinh(p.wk0 * f.h);
  }
  const double a = p.w0 * f.amp;
  // dx = a*hx*e*cos/i, dy = a*hx*e*sin/i, dz = -i*a*vz*e/i = -a*vz*e; e = ce + i se; (u)/i = -i u
  const double dxr = a * hx * cb * se, dxi = -a * hx * cb * ce;
  const double dyr = a * hx * sb * se, dyi = -a * hx * sb * ce;
  const double dzr = -a * vz * ce, dzi = -a * vz * se;
  qr = dxr * n[0] + dyr * n[1] + dzr * n[2];
  qi = dxi * n[0] + dyi * n[1] + dzi * n[2];
}

// VINP at field panel i (VINP :42-76; deep formula when H < 0 or H > 500)
__device__ void vinp(const FreqDev& f, const double* x, double beta, double xw, double yw, double& vr, double& vi) {
  const Incident p = incident_params(f);
  double sb, cb;
  sincos(beta, &sb, &cb);
  const double wkx = p.wk0 * ((x[0] - xw) * cb + (x[1] - yw) * sb);
  double se, ce;
  sincos(wkx, &se, &ce);
  double amp;   // VINP = (i * c * e)/i for finite depth, (-i * c * e)/i for deep water
  if (f.h < 0.0 || f.h > 500.0) {
    amp = -f.amp * f.g / p.w0 * exp(p.wk0 * x[2]);
  } else {
    const double dum = -f.amp * p.w0 / (p.wk0 * sinh(p.wk0 * f.h));
    amp = dum * cosh(p.wk0 * (x[2] + f.h));
    // finite depth: VINP = CI*DUM*cosh*e / CI = DUM*cosh*e
    vr = amp * ce;
    vi = amp * se;
    return;
  }
  // deep: VINP = -CI*c*e / CI = -c*e with c = AMP*G/W0*exp(kz)
  vr = amp * ce;
  vi = amp * se;
}

// W[j, 6b+d] = n_jd for j in body b; W[j, m+h] = -(DINP.n)_j,h
__device__ __forceinline__ void build_w_body(const GeomDev& g, const FreqDev& f, const double* beta, int nbeta,
                                             double xw, double yw, double* W, long long ldw, int j) {
  const int m = 6 * g.nbody;
  int b = 0;
  while (b + 1 < g.nbody && j >= g.body_start[b + 1]) ++b;
  for (int c = 0; c < m; ++c) {
    const long long off = (long long)j + (long long)c * ldw;
    W[2 * off] = (c / 6 == b) ? g.dxyz[6 * (long long)j + c % 6] : 0.0;
    W[2 * off + 1] = 0.0;
  }
  for (int h = 0; h < nbeta; ++h) {
    double qr, qi, xg[3];
    global_centroid(g, j, xg);
    dinp_dot_n(f, xg, g.dxyz + 6 * (long long)j, beta[h], xw, yw, qr, qi);
    const long long off = (long long)j + (long long)(m + h) * ldw;
    W[2 * off] = -qr;
    W[2 * off + 1] = -qi;
  }
}

__global__ void k_build_w(GeomDev g, FreqDev f, const double* beta, int nbeta, double xw, double yw, double* W,
                          int ldw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.n) return;
  build_w_body(g, f, beta, nbeta, xw, yw, W, ldw, j);
}

// the evanescent modes' depth factors per field point and their normalisations
// (FreqDev::evt, pmt), formed as fin_evanescent forms them
__global__ __launch_bounds__(256) void k_evan_table(GeomDev g, FreqDev f, double* __restrict__ evt,
                                                    double* __restrict__ pmt) {
#pragma clang fp contract(off)
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y + 1;
  const double km = f.wvn[m];
  if (e == 0) {
    const double twokmh = 2.0 * km * f.h;
    pmt[m] = 4.0 / (1.0 + sin(twokmh) / twokmh) / f.h;
  }
  if (e >= g.nt) return;
  const double zh = g.fp[3 * e + 2] + f.h;
  double sk, ck;
  sincos(km * zh, &sk, &ck);
  evt[2 * ((long long)m * g.nt + e)] = sk;
  evt[2 * ((long long)m * g.nt + e) + 1] = ck;
}

hipError_t evan_table(hipStream_t s, const GeomDev& g, const FreqDev& f, double* evt, double* pmt) {
  hipLaunchKernelGGL(k_evan_table, dim3((g.nt + 255) / 256, kNk - 1), dim3(256), 0, s, g, f, evt, pmt);
  return hipGetLastError();
}

hipError_t build_w(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* beta, int nbeta, double xw,
                   double yw, double* W, int ldw) {
  hipLaunchKernelGGL(k_build_w, dim3((g.n + 255) / 256), dim3(256), 0, s, g, f, beta, nbeta, xw, yw, W, ldw);
  return hipGetLastError();
}

// Force reductions (RFORCE_MULTI/EFORCE_MULTI, PotentWavForceMulti.f90:45-226):
// one workgroup per (force row 6*b+d, potential column c).
__device__ __forceinline__ void forces_body(const GeomDev& g, const FreqDev& f, const double* beta, int nbeta,
                                            double xw, double yw, const double* X, long long ldx, double* amas,
                                            double* bdmp, double* exfc, int row, int c) {
  const int b = row / 6, d = row % 6, m = 6 * g.nbody;
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  for (int i = g.body_start[b] + threadIdx.x; i < g.body_start[b + 1]; i += blockDim.x) {
    const long long off = (long long)i + (long long)c * ldx;
    double pr = X[2 * off], pi = X[2 * off + 1];
    if (c >= m) {
      double vr, vi, xg[3];
      global_centroid(g, i, xg);
      vinp(f, xg, beta[c - m], xw, yw, vr, vi);
      pr += vr;
      pi += vi;
    }
    const double w = g.ds[i] * g.dxyz[6 * (long long)i + d];
    ar += pr * w;
    ai += pi * w;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) { sr[threadIdx.x] += sr[threadIdx.x + st]; si[threadIdx.x] += si[threadIdx.x + st]; }
    __syncthreads();
  }
  if (threadIdx.x) return;
  if (c < m) {
    amas[row + (long long)c * m] = f.rho * sr[0];            // Im(i rho z) = rho Re z
    bdmp[row + (long long)c * m] = f.rho * f.w1 * si[0];     // -Re(i rho w z) = rho w Im z
  } else {
    const long long off = row + (long long)(c - m) * m;      // i w rho z
    exfc[2 * off] = -f.w1 * f.rho * si[0];
    exfc[2 * off + 1] = f.w1 * f.rho * sr[0];
  }
}

__global__ __launch_bounds__(256) void k_forces(GeomDev g, FreqDev f, const double* beta, int nbeta, double xw,
                                                double yw, const double* X, int ldx, double* amas, double* bdmp,
                                                double* exfc) {
  forces_body(g, f, beta, nbeta, xw, yw, X, ldx, amas, bdmp, exfc, blockIdx.x, blockIdx.y);
}

hipError_t forces(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* beta, int nbeta, double xw,
                  double yw, const double* X, int ldx, double* amas, double* bdmp, double* exfc) {
  const int m = 6 * g.nbody;
  hipLaunchKernelGGL(k_forces, dim3(m, m + nbeta), dim3(256), 0, s, g, f, beta, nbeta, xw, yw, X, ldx, amas, bdmp, exfc);
  return hipGetLastError();
}

// ---- batches of independent problems ------------------------------------------
// Pair codes: member << 40 | i << 20 | j. The list is laid out so that every
// 64-lane wave holds pairs of one member only: each (key, member) run is padded
// to a multiple of 64 with kNoPair entries, and the deep members' pair ranges
// are padded likewise. The member index is then wave-uniform
// (readfirstlane), so its descriptors load into scalar registers as the
// single-problem kernels' arguments do; a wave stages its member's dispersion
// roots in LDS. One workgroup per 256 entries: a grid-stride (persistent) loop
// over the list was measured to double the register allocation of these
// kernels (occupancy 2 -> 1 wave per SIMD) and is not used.
constexpr uint64_t kNoPair = ~0ull;

__device__ __forceinline__ void decode_pair(uint64_t c, int& p, int& i, int& j) {
  p = (int)(c >> 40);
  i = (int)((c >> 20) & 0xfffffu);
  j = (int)(c & 0xfffffu);
}
__device__ __forceinline__ FreqDev batch_freq(const BatchFreq& b) {
  FreqDev f;
  f.v = b.v; f.w1 = b.w1; f.wk = b.wk; f.h = b.h; f.rho = b.rho; f.g = b.g; f.amp = b.amp;
  f.wvn = b.wvn;
  return f;
}
// the wave's member, from its first entry (always a real pair)
__device__ __forceinline__ int wave_member(const uint64_t* codes, long long wave_base) {
  return __builtin_amdgcn_readfirstlane((int)(codes[wave_base] >> 40));
}
// every wave copies its member's roots into its LDS slice (all waves of the
// workgroup call it)
__device__ __forceinline__ const double* stage_wave_wvn(const BatchFreq* F, int p, double (*swvn)[kNk]) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = lane; k < kNk; k += 64) swvn[w][k] = F[p].wvn[k];
  __syncthreads();
  return swvn[w];
}

__global__ __launch_bounds__(256) void k_bclassify_count(const BatchProb* __restrict__ P,
                                                         const BatchFreq* __restrict__ F, int p, unsigned* hist) {
  __shared__ unsigned sh[kNumKeys];
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  const GeomDev g = P[p].g;
  const int i = blockIdx.y * 16 + (threadIdx.x >> 4), j = blockIdx.x * 16 + (threadIdx.x & 15);
  if (i < g.nt && j < g.n) atomicAdd(&sh[pair_key(g, F[p].h, i, j)], 1u);
  __syncthreads();
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x)
    if (sh[k]) atomicAdd(&hist[k], sh[k]);
}

// scatter with one atomic per (workgroup, key): ranks within the workgroup from LDS
__global__ __launch_bounds__(256) void k_bclassify_scatter(const BatchProb* __restrict__ P,
                                                           const BatchFreq* __restrict__ F, int p,
                                                           unsigned long long* cursor, uint64_t* codes) {
  __shared__ unsigned cnt[kNumKeys];
  __shared__ unsigned long long base[kNumKeys];
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x) cnt[k] = 0;
  __syncthreads();
  const GeomDev g = P[p].g;
  const int i = blockIdx.y * 16 + (threadIdx.x >> 4), j = blockIdx.x * 16 + (threadIdx.x & 15);
  const bool in = i < g.nt && j < g.n;
  const int key = in ? pair_key(g, F[p].h, i, j) : 0;
  const unsigned rank = in ? atomicAdd(&cnt[key], 1u) : 0u;
  __syncthreads();
  for (int k = threadIdx.x; k < kNumKeys; k += blockDim.x)
    if (cnt[k]) base[k] = atomicAdd(&cursor[k], (unsigned long long)cnt[k]);
  __syncthreads();
  if (in) codes[base[key] + rank] = ((uint64_t)p << 40) | ((uint64_t)i << 20) | (uint64_t)j;
}

template <int REG>
__global__ __launch_bounds__(256) void k_bassemble_list(const BatchProb* __restrict__ P,
                                                        const BatchFreq* __restrict__ F,
                                                        const uint64_t* __restrict__ codes, long long np) {
  __shared__ double swvn[4][kNk];
  const long long base = (long long)blockIdx.x * blockDim.x;
  const long long wb = base + (threadIdx.x & ~63);
  const bool wave_in = wb < np;
  const int p = wave_in ? wave_member(codes, wb) : 0;
  const double* wvn = stage_wave_wvn(F, p, swvn);
  const long long t = base + threadIdx.x;
  const uint64_t c = t < np ? codes[t] : kNoPair;
  if (c == kNoPair) return;
  int pp, i, j;
  decode_pair(c, pp, i, j);
  const FreqDev f = batch_freq(F[p]);
  assemble_pair<REG>(P[p].g, f, wvn, i, j, P[p].A, P[p].S, P[p].ld);
}

template <int REG, int G, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CoopWaves<K>::value, 8))) void k_bassemble_coop(const BatchProb* __restrict__ P,
                                                        const BatchFreq* __restrict__ F,
                                                        const uint64_t* __restrict__ codes, long long np) {
  const int l = threadIdx.x & (G - 1);
  __shared__ double swvn[4][kNk];
  {
    const long long base = (long long)blockIdx.x * blockDim.x;
    const long long wb = base + (threadIdx.x & ~63);
    const bool wave_in = wb < np;
    const int p = wave_in ? wave_member(codes, wb) : 0;
    const double* wvn = stage_wave_wvn(F, p, swvn);
    if (!wave_in) return;                    // the whole wave (no group operations pending)
    const long long t = base + threadIdx.x;
    uint64_t c = t < np ? codes[t] : kNoPair;
    const bool valid = c != kNoPair;
    if (!valid) c = codes[wb];               // padding: evaluate the wave's first pair, write nothing
    int pp, i, j;
    decode_pair(c, pp, i, j);
    const GeomDev& g = P[p].g;
    double xp[3];
    field_point(g, i, j, xp);
    const double* xq = g.xyz_p + 3 * (long long)j;
    const double h = F[p].h, v = F[p].v;
    const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
    const double zf = xq[2], zp = xp[2];
    const int nt = REG == kEigene ? fin_eigene_terms(rr, h) : fin_pidcock_terms(rr, h);
    double L1 = 0.0, L2 = 0.0, L3 = 0.0;
    for (int q = 0; q < G; ++q) {
      const double r = group_bcast<G>(rr, q);
      const double zfh = group_bcast<G>(zf, q) + h, zph = group_bcast<G>(zp, q) + h;
      const int ntq = group_bcast<G>(nt, q);
      double l1, l2, l3;
      coop_series<REG, G, K>(ntq, r, zfh, zph, h, v, wvn, l, l1, l2, l3);
      if (l == q) { L1 = l1; L2 = l2; L3 = l3; }
    }
    if (valid) coop_finish<REG>(g, wvn, h, i, j, rr, L1, L2, L3, P[p].A, P[p].S, P[p].ld);
  }
}

// all pairs of the members in `list` (infinite depth): member ranges of nt n
// pairs (nt = hull + waterplane field points) padded to multiples of 64 in `prefix`
template <bool DEEP>
__global__ __launch_bounds__(256) void k_bassemble_all(const BatchProb* __restrict__ P,
                                                       const BatchFreq* __restrict__ F, const int* __restrict__ list,
                                                       const long long* __restrict__ prefix, int nlist) {
  __shared__ double swvn[4][kNk];
  const long long total = prefix[nlist];
  {
    const long long base = (long long)blockIdx.x * blockDim.x;
    const long long wb = base + (threadIdx.x & ~63);
    int lo = 0;
    if (wb < total) {
      int hi = nlist - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= wb) lo = mid;
        else hi = mid - 1;
      }
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const int p = list[lo];
    const double* wvn = DEEP ? nullptr : stage_wave_wvn(F, p, swvn);
    const long long e = base + threadIdx.x - prefix[lo];
    const int n = P[p].g.n, nt = P[p].g.nt;
    if (wb >= total || e >= (long long)nt * n) return;
    const int i = (int)(e % nt), j = (int)(e / nt);
    const FreqDev f = batch_freq(F[p]);
    assemble_pair<DEEP ? -1 : -2>(P[p].g, f, DEEP ? f.wvn : wvn, i, j, P[p].A, P[p].S, P[p].ld);
  }
}

__global__ void k_bbuild_w(const BatchProb* __restrict__ P, const BatchFreq* __restrict__ F) {
  const BatchProb& bp = P[blockIdx.y];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= bp.g.n) return;
  const BatchFreq& fr = F[blockIdx.y];
  build_w_body(bp.g, batch_freq(fr), bp.beta, bp.nbeta, fr.xw, fr.yw, bp.W, bp.ld, j);
}

__global__ __launch_bounds__(256) void k_bforces(const BatchProb* __restrict__ P, const BatchFreq* __restrict__ F) {
  const BatchProb& bp = P[blockIdx.z];
  const int m = 6 * bp.g.nbody;
  if ((int)blockIdx.x >= m || (int)blockIdx.y >= m + bp.nbeta) return;
  const BatchFreq& fr = F[blockIdx.z];
  forces_body(bp.g, batch_freq(fr), bp.beta, bp.nbeta, fr.xw, fr.yw, bp.X, bp.ld, bp.amas, bp.bdmp, bp.exfc,
              blockIdx.x, blockIdx.y);
}

hipError_t batch_classify(hipStream_t s, const BatchProb* P, const BatchFreq* F, int p, int nt, int n, unsigned* hist,
                          unsigned long long* cursor, uint64_t* codes, bool count_only) {
  const dim3 grid((n + 15) / 16, (nt + 15) / 16);
  if (count_only) hipLaunchKernelGGL(k_bclassify_count, grid, dim3(256), 0, s, P, F, p, hist);
  else hipLaunchKernelGGL(k_bclassify_scatter, grid, dim3(256), 0, s, P, F, p, cursor, codes);
  return hipGetLastError();
}

hipError_t batch_assemble_regime(hipStream_t s, int regime, const BatchProb* P, const BatchFreq* F,
                                 const uint64_t* codes, long long np, int variant) {
  if (np <= 0) return hipSuccess;
  const dim3 gd((unsigned)((np + 255) / 256));
  switch (regime) {
    case kLinton: hipLaunchKernelGGL((k_bassemble_list<kLinton>), gd, dim3(256), 0, s, P, F, codes, np); break;
    case kPidcock:
      if (variant == 1) hipLaunchKernelGGL((k_bassemble_coop<kPidcock, 16, 7>), gd, dim3(256), 0, s, P, F, codes, np);
      else if (variant == 2) hipLaunchKernelGGL((k_bassemble_coop<kPidcock, 8, 13>), gd, dim3(256), 0, s, P, F, codes, np);
      else hipLaunchKernelGGL((k_bassemble_list<kPidcock>), gd, dim3(256), 0, s, P, F, codes, np);
      break;
    case kEigene:
      if (variant == 1) hipLaunchKernelGGL((k_bassemble_coop<kEigene, 8, 7>), gd, dim3(256), 0, s, P, F, codes, np);
      else if (variant == 2) hipLaunchKernelGGL((k_bassemble_coop<kEigene, 4, 13>), gd, dim3(256), 0, s, P, F, codes, np);
      else hipLaunchKernelGGL((k_bassemble_list<kEigene>), gd, dim3(256), 0, s, P, F, codes, np);
      break;
    default: hipLaunchKernelGGL((k_bassemble_list<kEigen>), gd, dim3(256), 0, s, P, F, codes, np); break;
  }
  return hipGetLastError();
}

hipError_t batch_assemble_all(hipStream_t s, bool deep, const BatchProb* P, const BatchFreq* F, const int* list,
                              const long long* prefix, int nlist, long long total) {
  if (nlist <= 0 || total <= 0) return hipSuccess;
  const dim3 gd((unsigned)((total + 255) / 256));
  if (deep) hipLaunchKernelGGL((k_bassemble_all<true>), gd, dim3(256), 0, s, P, F, list, prefix, nlist);
  else hipLaunchKernelGGL((k_bassemble_all<false>), gd, dim3(256), 0, s, P, F, list, prefix, nlist);
  return hipGetLastError();
}

hipError_t batch_build_w(hipStream_t s, const BatchProb* P, const BatchFreq* F, int count, int nmax) {
  hipLaunchKernelGGL(k_bbuild_w, dim3((nmax + 255) / 256, count), dim3(256), 0, s, P, F);
  return hipGetLastError();
}

hipError_t batch_forces(hipStream_t s, const BatchProb* P, const BatchFreq* F, int count, int mmax, int ncolmax) {
  hipLaunchKernelGGL(k_bforces, dim3(mmax, ncolmax, count), dim3(256), 0, s, P, F);
  return hipGetLastError();
}

// ---- element-wise parity kernels ---------------------------------------------
__global__ void k_green_batch(int n, int deep, const double* xf, const double* xp, const double* yf, const double* yp,
                              const double* zf, const double* zp, double v, const double* wvn_g, double h,
                              const int* flag, double* out) {
  __shared__ double wvn[kNk];
  if (!deep) load_wvn(wvn, wvn_g, kNk);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Grn4 g;
  if (deep) infgreen3d(xf[t], xp[t], yf[t], yp[t], zf[t], zp[t], v, flag[t], g);
  else fingreen3d(xf[t], xp[t], yf[t], yp[t], zf[t], zp[t], v, wvn, h, flag[t], g);
  for (int k = 0; k < 4; ++k) {
    out[8 * (long long)t + 2 * k] = g.re[k];
    out[8 * (long long)t + 2 * k + 1] = g.im[k];
  }
}

__global__ void k_rankine_batch(int n, const int* ncn, const double* verts, const double* xf, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Rk r;
  rankine(ncn[t], verts + 12 * (long long)t, xf + 3 * (long long)t, r);
  out[4 * (long long)t] = r.s;
  out[4 * (long long)t + 1] = r.dx;
  out[4 * (long long)t + 2] = r.dy;
  out[4 * (long long)t + 3] = r.dz;
}

// K0, K1 at n points: mode 0 the reference's IK01A (sf_k01), 1 the Horner-form
// evaluation of the mixed-precision kernels (sf_k01_fast)
__global__ void k_ik01a(int n, const double* __restrict__ x, int mode, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double k0, k1;
  if (mode) sf_k01_fast(x[t], k0, k1);
  else sf_k01(x[t], k0, k1);
  out[2 * (long long)t] = k0;
  out[2 * (long long)t + 1] = k1;
}

hipError_t ik01a_batch(hipStream_t s, int n, const double* x, int mode, double* out) {
  hipLaunchKernelGGL(k_ik01a, dim3((n + 255) / 256), dim3(256), 0, s, n, x, mode, out);
  return hipGetLastError();
}

hipError_t green_batch(hipStream_t s, int n, bool deep, const double* xf, const double* xp, const double* yf,
                       const double* yp, const double* zf, const double* zp, double v, const double* wvn, double h,
                       const int* flag, double* out) {
  hipLaunchKernelGGL(k_green_batch, dim3((n + 255) / 256), dim3(256), 0, s, n, deep ? 1 : 0, xf, xp, yf, yp, zf, zp, v,
                     wvn, h, flag, out);
  return hipGetLastError();
}
hipError_t rankine_batch(hipStream_t s, int n, const int* ncn, const double* verts, const double* xf, double* out) {
  hipLaunchKernelGGL(k_rankine_batch, dim3((n + 255) / 256), dim3(256), 0, s, n, ncn, verts, xf, out);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Field points (CalPotentialMulti, PressureElevationMulti.f90:55-272)
// ---------------------------------------------------------------------------

// S-like (DUM(1)) and A-like (DUM(2)) entries of nfp field points against every
// hull panel, as the assembly's rows but without the free term: Sf, Af are
// nfp x n, column-major (entry (p, j) at p + j nfp), one thread per entry
__global__ __launch_bounds__(256) void k_field_rows(GeomDev g, FreqDev f, const double* __restrict__ xfp, int nfp,
                                                    double* __restrict__ Af, double* __restrict__ Sf) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nfp * g.n) return;
  const int p = (int)(t % nfp), j = (int)(t / nfp);
  double xp[3];
  global_to_src(g, j, xfp + 3 * (long long)p, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  if (f.h < 0.0) infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gr);
  else fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, f.wvn, f.h, flag, gr);
  double ar, ai, sr, si;
  pair_values(g, -1, j, xp, flag, gr, ar, ai, sr, si);   // -1: no panel is the field point's own
  Af[2 * t] = ar;
  Af[2 * t + 1] = ai;
  Sf[2 * t] = sr;
  Sf[2 * t + 1] = si;
}

// the field point lies on hull panel j: within max(1e-6, 1e-3 PNSZ) of its plane
// and inside its polygon (the cross products of the edges with the point keep
// one sign, to 1e-3 PNSZ^2) -- CalPotentialMulti's free-term test
__device__ bool on_panel(const GeomDev& g, int j, const double* xg) {
  double x[3];
  global_to_src(g, j, xg, x);
  const double* v = g.verts + 12 * (long long)j;
  const double* en = g.dxyz + 6 * (long long)j;
  const int nv = g.ncn[j];
  const double tolp = fmax(1.0e-6, 1.0e-3 * g.pnsz[j]), tola = 1.0e-3 * g.pnsz[j] * g.pnsz[j];
  const double dpln = (x[0] - v[0]) * en[0] + (x[1] - v[1]) * en[1] + (x[2] - v[2]) * en[2];
  if (fabs(dpln) > tolp) return false;
  double smin = 1.0e30, smax = -1.0e30;
  for (int k = 0; k < nv; ++k) {
    const double* a = v + 3 * k;
    const double* b = v + 3 * (k + 1 < nv ? k + 1 : 0);
    const double ev0 = b[0] - a[0], ev1 = b[1] - a[1], ev2 = b[2] - a[2];
    const double wv0 = x[0] - a[0], wv1 = x[1] - a[1], wv2 = x[2] - a[2];
    const double sv = (ev1 * wv2 - ev2 * wv1) * en[0] + (ev2 * wv0 - ev0 * wv2) * en[1] + (ev0 * wv1 - ev1 * wv0) * en[2];
    smin = fmin(smin, sv);
    smax = fmax(smax, sv);
  }
  return smin >= -tola || smax <= tola;
}

// one workgroup per field point: the free term (2 pi on the hull, 4 pi in the
// fluid), the incident potential on the diffraction columns (ISOL = 1), and
// pressure i w rho phi and elevation i w / g phi (rho phi and phi at zero or
// infinite frequency); pot holds the integral sums on entry
__global__ __launch_bounds__(256) void k_field_finish(GeomDev g, FreqDev f, const double* __restrict__ xfp, int nfp,
                                                      int m, int nbeta, const double* __restrict__ beta, double xw,
                                                      double yw, double* __restrict__ pot, double* __restrict__ prs,
                                                      double* __restrict__ elv) {
  const int p = blockIdx.x;
  const double* x = xfp + 3 * (long long)p;
  int on = 0;
  for (int j = threadIdx.x; j < g.n && !on; j += blockDim.x) on = on_panel(g, j, x) ? 1 : 0;
  on = __syncthreads_or(on);
  const double sld = on ? 2.0 * kPi : 4.0 * kPi;
  const bool limit = fabs(f.v) < F32(1.e-8) || fabs(f.v + 1.0) < F32(1.e-8);
  for (int c = threadIdx.x; c < m + nbeta; c += blockDim.x) {
    const long long o = 2 * ((long long)p + (long long)c * nfp);
    double pr = pot[o] / sld, pi = pot[o + 1] / sld;
    if (c >= m && !limit) {
      double vr, vi;
      vinp(f, x, beta[c - m], xw, yw, vr, vi);
      pr += vr;
      pi += vi;
    }
    pot[o] = pr;
    pot[o + 1] = pi;
    if (limit) {
      prs[o] = f.rho * pr;
      prs[o + 1] = f.rho * pi;
      elv[o] = pr;
      elv[o + 1] = pi;
    } else {
      prs[o] = -f.w1 * f.rho * pi;
      prs[o + 1] = f.w1 * f.rho * pr;
      elv[o] = -f.w1 / f.g * pi;
      elv[o + 1] = f.w1 / f.g * pr;
    }
  }
}

hipError_t field_rows(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* xfp, int nfp, double* Af,
                      double* Sf) {
  const long long tot = (long long)nfp * g.n;
  hipLaunchKernelGGL(k_field_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, f, xfp, nfp, Af, Sf);
  return hipGetLastError();
}

hipError_t field_finish(hipStream_t s, const GeomDev& g, const FreqDev& f, const double* xfp, int nfp, int m,
                        int nbeta, const double* beta, double xw, double yw, double* pot, double* prs, double* elv) {
  hipLaunchKernelGGL(k_field_finish, dim3(nfp), dim3(256), 0, s, g, f, xfp, nfp, m, nbeta, beta, xw, yw, pot, prs,
                     elv);
  return hipGetLastError();
}

}  // namespace hm
