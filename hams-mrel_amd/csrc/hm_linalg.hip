// hm_linalg.hip -- dense complex linear algebra of the solve stage on gfx950:
// the work of LAPACK ZGETRF/ZGETRS at the reference's call sites
// (AssbMatxMulti.f90:112,:279,:346; AssbMatx_irr.f90:175,:491) plus the complex
// GEMMs of the irregular-frequency normal equations (AssbMatx_irr.f90:153-169,
// :300-318) and of the right-hand-side assembly.
//
// Storage: column-major complex matrices, interleaved (re, im), leading
// dimension ld; a batch of matrices is addressed by a per-matrix element stride.
// Everything is templated on the real type T (double for the plain path, float
// for the mixed-precision factorisation that FP64 iterative refinement repairs).
//
// * cgemm: C = alpha*op(A)*B + beta*C with op = identity or plain transpose.
//   64x64 output tile per 256-thread workgroup (2x2 waves of 32x32), K staged
//   through LDS in chunks of 16, inner products on MFMA 16x16x4 (f64 or f32);
//   a complex product is four real MFMAs into separate Re/Im accumulators.
// * LU with partial pivoting, right-looking, panel width kNb: per panel column
//   one launch scales the column, applies the rank-1 update to the rest of the
//   panel and -- in the last workgroup to finish -- searches the next column's
//   pivot (first max of |re|+|im|, as izamax) and swaps it into place. Then the
//   panel's row swaps are applied to the other columns, the block row is solved
//   with the unit-lower panel block, and the trailing matrix gets one GEMM.
// * Triangular solves with many right-hand sides: per 64-row block a
//   diagonal-block solve (64 lanes per right-hand side) and a GEMM update.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <utility>
#include <stdint.h>

#include "hm_linalg.hpp"

namespace hm {

template <typename T> struct vec4;
template <> struct vec4<double> { typedef double __attribute__((ext_vector_type(4))) type; };
template <> struct vec4<float> { typedef float __attribute__((ext_vector_type(4))) type; };

// Row of accumulator element v held by `lane` in a 16x16 MFMA result (column is
// lane % 16). Measured on gfx950: the f32 16x16x4 result holds 4 consecutive rows
// per lane group, the f64 16x16x4 result interleaves them with stride 4.
template <typename T> __device__ __forceinline__ int mfma_row(int lane, int v);
template <> __device__ __forceinline__ int mfma_row<float>(int lane, int v) { return 4 * (lane >> 4) + v; }
template <> __device__ __forceinline__ int mfma_row<double>(int lane, int v) { return (lane >> 4) + 4 * v; }

__device__ __forceinline__ vec4<double>::type mfma16(double a, double b, vec4<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ vec4<float>::type mfma16(float a, float b, vec4<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kBM = 64, kBN = 64, kBK = 16;

// One 64 x 64 tile (rows row0.., columns col0..) of
// C[m x n] = alpha * op(A)[m x k] * B[k x n] + beta * C, K staged through LDS
// in chunks of KB (8, 16 or 32; 16 by default, the grouped factorisation's
// autotuner picks among them). M3: the complex product from three real ones
// (P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2,
// Im = P3 - P1 - P2), three MFMAs per complex step instead of four; the
// imaginary part's error is then bounded by (|Ar| + |Ai|)(|Br| + |Bi|) rather
// than by its own terms.
template <typename T, bool TA, int KB = kBK, bool M3 = false>
__device__ __forceinline__ void cgemm_tile(int m, int n, int k, T ar, T ai, const T* __restrict__ A, int lda,
                                           const T* __restrict__ B, int ldb, T br, T bi, T* __restrict__ C,
                                           int ldc, int row0, int col0) {
  static_assert(KB == 8 || KB == 16 || KB == 32, "K chunk");
  constexpr int Q = KB / 4;            // elements of each operand chunk per thread
  __shared__ T as_re[KB][kBM + 1], as_im[KB][kBM + 1], bs_re[KB][kBN + 1], bs_im[KB][kBN + 1];
  typedef typename vec4<T>::type v4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  v4 cre[2][2], cim[2][2], cx[2][2];   // M3: P1, P3, P2
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int v = 0; v < 4; ++v) { cre[i][j][v] = 0; cim[i][j][v] = 0; cx[i][j][v] = 0; }

  // op(A)[row0:+64, k0:+KB] and B[k0:+KB, col0:+64] go through registers into
  // LDS; the next chunk's global loads are issued before the current chunk's
  // MFMAs, so their latency hides behind the arithmetic
  T pa_r[Q], pa_i[Q], pb_r[Q], pb_i[Q];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int r, kk;
      if (!TA) { r = tid & 63; kk = (tid >> 6) + 4 * q; }
      else { kk = tid % KB; r = tid / KB + (256 / KB) * q; }
      const int gr = row0 + r, gk = k0 + kk;
      pa_r[q] = 0;
      pa_i[q] = 0;
      if (gr < m && gk < k) {
        const long long off = TA ? (long long)gk + (long long)gr * lda : (long long)gr + (long long)gk * lda;
        pa_r[q] = A[2 * off];
        pa_i[q] = A[2 * off + 1];
      }
      const int bk = tid % KB, bc = tid / KB + (256 / KB) * q;
      const int gbk = k0 + bk, gbc = col0 + bc;
      pb_r[q] = 0;
      pb_i[q] = 0;
      if (gbk < k && gbc < n) {
        const long long off = (long long)gbk + (long long)gbc * ldb;
        pb_r[q] = B[2 * off];
        pb_i[q] = B[2 * off + 1];
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int r, kk;
      if (!TA) { r = tid & 63; kk = (tid >> 6) + 4 * q; }
      else { kk = tid % KB; r = tid / KB + (256 / KB) * q; }
      as_re[kk][r] = pa_r[q];
      as_im[kk][r] = pa_i[q];
      const int bk = tid % KB, bc = tid / KB + (256 / KB) * q;
      bs_re[bk][bc] = pb_r[q];
      bs_im[bk][bc] = pb_i[q];
    }
  };
  if (k > 0) {
    fetch(0);
    stash();
  }
  __syncthreads();
  for (int k0 = 0; k0 < k; k0 += KB) {
    const bool more = k0 + KB < k;
    if (more) fetch(k0 + KB);
    for (int k4 = 0; k4 < KB; k4 += 4) {
      const int kr = k4 + (lane >> 4), li = lane & 15;
      T a_re[2], a_im[2], b_re[2], b_im[2];
      for (int s = 0; s < 2; ++s) {
        a_re[s] = as_re[kr][wr + 16 * s + li];
        a_im[s] = as_im[kr][wr + 16 * s + li];
        b_re[s] = bs_re[kr][wc + 16 * s + li];
        b_im[s] = bs_im[kr][wc + 16 * s + li];
      }
      if (M3) {
        T a_s[2], b_s[2];
        for (int s = 0; s < 2; ++s) {
          a_s[s] = a_re[s] + a_im[s];
          b_s[s] = b_re[s] + b_im[s];
        }
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            cre[i][j] = mfma16(a_re[i], b_re[j], cre[i][j]);
            cx[i][j] = mfma16(a_im[i], b_im[j], cx[i][j]);
            cim[i][j] = mfma16(a_s[i], b_s[j], cim[i][j]);
          }
      } else {
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            cre[i][j] = mfma16(a_re[i], b_re[j], cre[i][j]);
            cre[i][j] = mfma16(-a_im[i], b_im[j], cre[i][j]);
            cim[i][j] = mfma16(a_re[i], b_im[j], cim[i][j]);
            cim[i][j] = mfma16(a_im[i], b_re[j], cim[i][j]);
          }
      }
    }
    __syncthreads();
    if (more) {
      stash();
      __syncthreads();
    }
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int v = 0; v < 4; ++v) {
        const int gr = row0 + wr + 16 * i + mfma_row<T>(lane, v);
        const int gc = col0 + wc + 16 * j + (lane & 15);
        if (gr < m && gc < n) {
          const long long off = (long long)gr + (long long)gc * ldc;
          const T pr = M3 ? cre[i][j][v] - cx[i][j][v] : cre[i][j][v];
          const T pi = M3 ? cim[i][j][v] - cre[i][j][v] - cx[i][j][v] : cim[i][j][v];
          T outr = ar * pr - ai * pi, outi = ar * pi + ai * pr;
          if (br != T(0) || bi != T(0)) {
            const T c_r = C[2 * off], c_i = C[2 * off + 1];
            outr += br * c_r - bi * c_i;
            outi += br * c_i + bi * c_r;
          }
          C[2 * off] = outr;
          C[2 * off + 1] = outi;
        }
      }
}

// blockIdx.z = batch index
template <typename T, bool TA>
__global__ __launch_bounds__(256) void k_cgemm(int m, int n, int k, T ar, T ai, const T* __restrict__ A,
                                               int lda, long long sA, const T* __restrict__ B, int ldb,
                                               long long sB, T br, T bi, T* __restrict__ C, int ldc,
                                               long long sC) {
  cgemm_tile<T, TA>(m, n, k, ar, ai, A + 2 * sA * blockIdx.z, lda, B + 2 * sB * blockIdx.z, ldb, br, bi,
                    C + 2 * sC * blockIdx.z, ldc, blockIdx.x * kBM, blockIdx.y * kBN);
}

// blockIdx.z = batch index; three real products a complex step
template <typename T, bool TA>
__global__ __launch_bounds__(256, 2) void k_cgemm_3m(int m, int n, int k, T ar, T ai, const T* __restrict__ A, int lda,
                                                  long long sA, const T* __restrict__ B, int ldb, long long sB, T br,
                                                  T bi, T* __restrict__ C, int ldc, long long sC) {
  cgemm_tile<T, TA, kBK, true>(m, n, k, ar, ai, A + 2 * sA * blockIdx.z, lda, B + 2 * sB * blockIdx.z, ldb, br, bi,
                               C + 2 * sC * blockIdx.z, ldc, blockIdx.x * kBM, blockIdx.y * kBN);
}

// blockIdx.z = group member; tiles outside the member's m x n exit
template <typename T, bool TA, int KB>
__global__ __launch_bounds__(256) void k_cgemm_grouped(const GemmDesc<T>* __restrict__ d) {
  const GemmDesc<T> g = d[blockIdx.z];
  const int row0 = blockIdx.x * kBM, col0 = blockIdx.y * kBN;
  if (row0 >= g.m || col0 >= g.n || g.k <= 0) return;
  cgemm_tile<T, TA, KB>(g.m, g.n, g.k, g.ar, g.ai, g.A, (int)g.lda, g.B, (int)g.ldb, g.br, g.bi, g.C, (int)g.ldc,
                        row0, col0);
}

template <typename T>
hipError_t cgemm(hipStream_t s, bool transA, int m, int n, int k, T ar, T ai, const T* A, int lda,
                 long long sA, const T* B, int ldb, long long sB, T br, T bi, T* C, int ldc,
                 long long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return hipSuccess;
  dim3 grid((m + kBM - 1) / kBM, (n + kBN - 1) / kBN, batch);
  if (transA)
    hipLaunchKernelGGL((k_cgemm<T, true>), grid, dim3(256), 0, s, m, n, k, ar, ai, A, lda, sA, B, ldb, sB, br, bi, C, ldc, sC);
  else
    hipLaunchKernelGGL((k_cgemm<T, false>), grid, dim3(256), 0, s, m, n, k, ar, ai, A, lda, sA, B, ldb, sB, br, bi, C, ldc, sC);
  return hipGetLastError();
}

template <typename T>
hipError_t cgemm_3m(hipStream_t s, bool transA, int m, int n, int k, T ar, T ai, const T* A, int lda, long long sA,
                    const T* B, int ldb, long long sB, T br, T bi, T* C, int ldc, long long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return hipSuccess;
  dim3 grid((m + kBM - 1) / kBM, (n + kBN - 1) / kBN, batch);
  if (transA)
    hipLaunchKernelGGL((k_cgemm_3m<T, true>), grid, dim3(256), 0, s, m, n, k, ar, ai, A, lda, sA, B, ldb, sB, br, bi,
                       C, ldc, sC);
  else
    hipLaunchKernelGGL((k_cgemm_3m<T, false>), grid, dim3(256), 0, s, m, n, k, ar, ai, A, lda, sA, B, ldb, sB, br, bi,
                       C, ldc, sC);
  return hipGetLastError();
}
template hipError_t cgemm_3m<double>(hipStream_t, bool, int, int, int, double, double, const double*, int, long long,
                                     const double*, int, long long, double, double, double*, int, long long, int);
template hipError_t cgemm_3m<float>(hipStream_t, bool, int, int, int, float, float, const float*, int, long long,
                                    const float*, int, long long, float, float, float*, int, long long, int);

template <typename T>
hipError_t cgemm_grouped(hipStream_t s, bool transA, const GemmDesc<T>* d, int count, int mmax, int nmax, int kb) {
  if (mmax <= 0 || nmax <= 0 || count <= 0) return hipSuccess;
  dim3 grid((mmax + kBM - 1) / kBM, (nmax + kBN - 1) / kBN, count);
#define HM_GG(TA, KB) hipLaunchKernelGGL((k_cgemm_grouped<T, TA, KB>), grid, dim3(256), 0, s, d)
  if (transA) {
    if (kb == 8) HM_GG(true, 8); else if (kb == 32) HM_GG(true, 32); else HM_GG(true, 16);
  } else {
    if (kb == 8) HM_GG(false, 8); else if (kb == 32) HM_GG(false, 32); else HM_GG(false, 16);
  }
#undef HM_GG
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LU
// ---------------------------------------------------------------------------

// block-wide first-max search of |re|+|im| over rows [r0, n) of column j;
// returns the winning row in *prow (shared). Tie -> smallest row.
template <typename T>
__device__ void block_pivot_search(const T* a, int lda, int n, int r0, int j, int* prow, T* pval) {
  __shared__ T smax[256];
  __shared__ int sidx[256];
  T best = T(-1);
  int bi = n;
  for (int r = r0 + threadIdx.x; r < n; r += blockDim.x) {
    const long long off = (long long)r + (long long)j * lda;
    const T v = fabs(a[2 * off]) + fabs(a[2 * off + 1]);
    if (v > best) { best = v; bi = r; }
  }
  smax[threadIdx.x] = best;
  sidx[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const T o = smax[threadIdx.x + s];
      const int oi = sidx[threadIdx.x + s];
      if (o > smax[threadIdx.x] || (o == smax[threadIdx.x] && oi < sidx[threadIdx.x])) {
        smax[threadIdx.x] = o;
        sidx[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { *prow = sidx[0]; *pval = smax[0]; }
  __syncthreads();
}

// swap rows j and p over columns [c0, c1), record the pivot, flag a zero pivot
template <typename T>
__device__ void block_apply_pivot(T* a, int lda, int j, int p, T pmax, int c0, int c1, int* ipiv, int* info) {
  if (p != j)
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
      const long long oj = (long long)j + (long long)c * lda, op = (long long)p + (long long)c * lda;
      const T xr = a[2 * oj], xi = a[2 * oj + 1];
      a[2 * oj] = a[2 * op];
      a[2 * oj + 1] = a[2 * op + 1];
      a[2 * op] = xr;
      a[2 * op + 1] = xi;
    }
  if (threadIdx.x == 0) {
    ipiv[j] = p;
    if (!(pmax > T(0)) && *info == 0) *info = j + 1;
  }
}

// pivot of the first column of a panel: one workgroup per matrix
template <typename T>
__global__ __launch_bounds__(256) void k_lu_pivot(T* A, int lda, long long sA, int n, int j, int c0, int c1,
                                                  int* ipiv, int* info) {
  T* a = A + 2 * sA * blockIdx.z;
  __shared__ int prow;
  __shared__ T pval;
  block_pivot_search(a, lda, n, j, j, &prow, &pval);
  block_apply_pivot(a, lda, j, prow, pval, c0, c1, ipiv + (long long)n * blockIdx.z, info + blockIdx.z);
}

// one panel column j (pivot already in place): scale L(:,j), rank-1 update of
// panel columns (j, c1); the last of the nblk workgroups to finish pivots
// column j+1. bx = this workgroup's index among the nblk.
template <typename T>
__device__ __forceinline__ void lu_col_body(T* a, int lda, int n, int j, int c0, int c1, int* ipiv, int* info,
                                            unsigned* counter, int nblk, int bx) {
  __shared__ T u_re[kLuNb], u_im[kLuNb];
  __shared__ int last;
  const int ncu = c1 - j - 1;
  for (int c = threadIdx.x; c < ncu; c += blockDim.x) {
    const long long off = (long long)j + (long long)(j + 1 + c) * lda;
    u_re[c] = a[2 * off];
    u_im[c] = a[2 * off + 1];
  }
  const long long od = (long long)j + (long long)j * lda;
  const T dr = a[2 * od], di = a[2 * od + 1];
  const T den = dr * dr + di * di;
  T rr = 0, ri = 0;
  if (den > T(0)) { rr = dr / den; ri = -di / den; }
  __syncthreads();
  const int r = j + 1 + bx * blockDim.x + threadIdx.x;
  if (r < n) {
    const long long ol = (long long)r + (long long)j * lda;
    const T xr = a[2 * ol], xi = a[2 * ol + 1];
    const T lr = xr * rr - xi * ri, li = xr * ri + xi * rr;
    a[2 * ol] = lr;
    a[2 * ol + 1] = li;
    for (int c = 0; c < ncu; ++c) {
      const long long oc = (long long)r + (long long)(j + 1 + c) * lda;
      a[2 * oc] -= lr * u_re[c] - li * u_im[c];
      a[2 * oc + 1] -= lr * u_im[c] + li * u_re[c];
    }
  }
  if (j + 1 >= c1) return;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = (atomicAdd(counter, 1u) == (unsigned)nblk - 1u);
  __syncthreads();
  if (!last) return;
  __threadfence();
  __shared__ int prow;
  __shared__ T pval;
  block_pivot_search(a, lda, n, j + 1, j + 1, &prow, &pval);
  block_apply_pivot(a, lda, j + 1, prow, pval, c0, c1, ipiv, info);
  if (threadIdx.x == 0) *counter = 0u;
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_col(T* A, int lda, long long sA, int n, int j, int c0, int c1,
                                                int* ipiv, int* info, unsigned* counter) {
  lu_col_body<T>(A + 2 * sA * blockIdx.z, lda, n, j, c0, c1, ipiv + (long long)n * blockIdx.z, info + blockIdx.z,
                 counter + blockIdx.z, gridDim.x, blockIdx.x);
}

// A whole panel [k0, k1) of one large matrix in one cooperative launch: the
// panel's rows [k0, n) are split into slabs of kCpRows rows, one per workgroup,
// held in LDS for the whole panel. Per column j: each workgroup finds its
// candidate pivot (largest |re| + |im| at rows >= j, smallest row on ties, as
// block_pivot_search) and publishes its value, row and panel row; workgroup 0,
// which holds rows k0 .. k0 + kCpRows - 1, publishes row j; one grid-wide
// barrier; then every workgroup picks the same winner from the candidates, the
// owners of rows j and p swap them in their slabs, and each workgroup scales
// L(:, j) and applies the rank-1 update to its own rows, in LDS. One barrier a
// column instead of one launch a column and a separate pivot search, and the
// panel goes through HBM once. The exchange area (two column parities, so that
// the next column's candidates never overwrite ones still being read):
// cval[2][W], crow[2][W], cdat[2][W][64] complex, jdat[2][64] complex.
constexpr int kCpRows = 128;

template <typename T>
__global__ __launch_bounds__(256) void k_lu_panel_coop(T* A, int lda, int n, int k0, int k1, int* ipiv, int* info,
                                                       T* cval, int* crow, T* cdat, T* jdat) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int W = gridDim.x, w = blockIdx.x, t = threadIdx.x;
  const int kb = k1 - k0;
  const int rbeg = k0 + w * kCpRows;
  const int nr = n - rbeg < kCpRows ? n - rbeg : kCpRows;
  __shared__ T S[kCpRows * kLuNb * 2];   // the slab, column-major: row r of column c at 2 (c kCpRows + r)
  __shared__ T U[kLuNb * 2];             // the pivot row (row j after the swap)
  __shared__ T rv[256];
  __shared__ int rx[256];
  for (int e = t; e < nr * kb; e += 256) {
    const int r = e % nr, c = e / nr;
    const long long o = 2 * ((long long)(rbeg + r) + (long long)(k0 + c) * lda);
    S[2 * (c * kCpRows + r)] = A[o];
    S[2 * (c * kCpRows + r) + 1] = A[o + 1];
  }
  __syncthreads();
  for (int j = k0; j < k1; ++j) {
    const int jj = j - k0, par = jj & 1;
    // this slab's candidate
    T best = T(-1);
    int bi = n;
    for (int r = t; r < nr; r += 256) {
      if (rbeg + r < j) continue;
      const T v = fabs(S[2 * (jj * kCpRows + r)]) + fabs(S[2 * (jj * kCpRows + r) + 1]);
      if (v > best) { best = v; bi = rbeg + r; }
    }
    rv[t] = best;
    rx[t] = bi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) {
        const T o = rv[t + h];
        const int oi = rx[t + h];
        if (o > rv[t] || (o == rv[t] && oi < rx[t])) { rv[t] = o; rx[t] = oi; }
      }
      __syncthreads();
    }
    const int mine = rx[0];
    if (t == 0) {
      cval[par * W + w] = rv[0];
      crow[par * W + w] = mine;
    }
    if (mine < n)
      for (int c = t; c < kb; c += 256) {
        const int r = mine - rbeg;
        cdat[2 * ((par * W + w) * kLuNb + c)] = S[2 * (c * kCpRows + r)];
        cdat[2 * ((par * W + w) * kLuNb + c) + 1] = S[2 * (c * kCpRows + r) + 1];
      }
    if (w == 0)
      for (int c = t; c < kb; c += 256) {
        jdat[2 * (par * kLuNb + c)] = S[2 * (c * kCpRows + jj)];
        jdat[2 * (par * kLuNb + c) + 1] = S[2 * (c * kCpRows + jj) + 1];
      }
    __syncthreads();
    grid.sync();
    // the winner, the same in every workgroup
    T b2 = T(-1);
    int i2 = n;
    for (int q = t; q < W; q += 256) {
      const T v = cval[par * W + q];
      const int r = crow[par * W + q];
      if (v > b2 || (v == b2 && r < i2)) { b2 = v; i2 = r; }
    }
    rv[t] = b2;
    rx[t] = i2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) {
        const T o = rv[t + h];
        const int oi = rx[t + h];
        if (o > rv[t] || (o == rv[t] && oi < rx[t])) { rv[t] = o; rx[t] = oi; }
      }
      __syncthreads();
    }
    const T pmax = rv[0];
    const int p = rx[0] < n ? rx[0] : j;   // a column without a comparable entry keeps its row
    const int wp = (p - k0) / kCpRows;
    for (int c = t; c < kb; c += 256) {
      const T* src = p == j ? jdat + 2 * (par * kLuNb + c) : cdat + 2 * ((par * W + wp) * kLuNb + c);
      U[2 * c] = src[0];
      U[2 * c + 1] = src[1];
    }
    if (w == 0 && t == 0) {
      ipiv[j] = p;
      if (!(pmax > T(0)) && *info == 0) *info = j + 1;
    }
    __syncthreads();
    if (p != j) {
      if (w == 0)
        for (int c = t; c < kb; c += 256) {
          S[2 * (c * kCpRows + jj)] = U[2 * c];
          S[2 * (c * kCpRows + jj) + 1] = U[2 * c + 1];
        }
      if (w == wp)
        for (int c = t; c < kb; c += 256) {
          S[2 * (c * kCpRows + p - rbeg)] = jdat[2 * (par * kLuNb + c)];
          S[2 * (c * kCpRows + p - rbeg) + 1] = jdat[2 * (par * kLuNb + c) + 1];
        }
      __syncthreads();
    }
    // L(:, j) and the rank-1 update of the panel's columns (j, k1), own rows > j
    const T dr = U[2 * jj], di = U[2 * jj + 1];
    const T den = dr * dr + di * di;
    T qr = 0, qi = 0;
    if (den > T(0)) { qr = dr / den; qi = -di / den; }
    for (int r = t; r < nr; r += 256) {
      if (rbeg + r <= j) continue;
      T* col = S + 2 * r;
      const T xr = col[2 * jj * kCpRows], xi = col[2 * jj * kCpRows + 1];
      const T lr = xr * qr - xi * qi, li = xr * qi + xi * qr;
      col[2 * jj * kCpRows] = lr;
      col[2 * jj * kCpRows + 1] = li;
      for (int c = jj + 1; c < kb; ++c) {
        col[2 * c * kCpRows] -= lr * U[2 * c] - li * U[2 * c + 1];
        col[2 * c * kCpRows + 1] -= lr * U[2 * c + 1] + li * U[2 * c];
      }
    }
    __syncthreads();
  }
  for (int e = t; e < nr * kb; e += 256) {
    const int r = e % nr, c = e / nr;
    const long long o = 2 * ((long long)(rbeg + r) + (long long)(k0 + c) * lda);
    A[o] = S[2 * (c * kCpRows + r)];
    A[o + 1] = S[2 * (c * kCpRows + r) + 1];
  }
}

// A whole panel [k0, k1) of a small matrix by one workgroup per matrix: for
// each column the pivot search, the row swap, the scaling of L(:, j) and the
// rank-1 update of the panel's columns (j, k1), with the panel in L2. Same
// arithmetic per row as lu_col_body; one launch per panel instead of one per
// column (the per-column launches of n <= 4096 matrices are latency bound).
template <typename T>
__global__ __launch_bounds__(256) void k_lu_panel(T* A, int lda, long long sA, int n, int k0, int k1, int* ipiv,
                                                  int* info) {
  T* a = A + 2 * sA * blockIdx.z;
  int* pv = ipiv + (long long)n * blockIdx.z;
  int* inf = info + blockIdx.z;
  __shared__ int prow;
  __shared__ T pval;
  __shared__ T u_re[kLuNb], u_im[kLuNb];
  for (int j = k0; j < k1; ++j) {
    block_pivot_search(a, lda, n, j, j, &prow, &pval);
    block_apply_pivot(a, lda, j, prow, pval, k0, k1, pv, inf);
    __syncthreads();
    const int ncu = k1 - j - 1;
    for (int c = threadIdx.x; c < ncu; c += blockDim.x) {
      const long long off = (long long)j + (long long)(j + 1 + c) * lda;
      u_re[c] = a[2 * off];
      u_im[c] = a[2 * off + 1];
    }
    const long long od = (long long)j + (long long)j * lda;
    const T dr = a[2 * od], di = a[2 * od + 1];
    const T den = dr * dr + di * di;
    T rr = 0, ri = 0;
    if (den > T(0)) { rr = dr / den; ri = -di / den; }
    __syncthreads();
    for (int r = j + 1 + threadIdx.x; r < n; r += blockDim.x) {
      const long long ol = (long long)r + (long long)j * lda;
      const T xr = a[2 * ol], xi = a[2 * ol + 1];
      const T lr = xr * rr - xi * ri, li = xr * ri + xi * rr;
      a[2 * ol] = lr;
      a[2 * ol + 1] = li;
      for (int c = 0; c < ncu; ++c) {
        const long long oc = (long long)r + (long long)(j + 1 + c) * lda;
        a[2 * oc] -= lr * u_re[c] - li * u_im[c];
        a[2 * oc + 1] -= lr * u_im[c] + li * u_re[c];
      }
    }
    __syncthreads();
  }
}

// A whole panel [k0, k1) of an FP32 matrix of n <= 1024 R rows, one 1024-thread
// workgroup per matrix, in sub-panels of 8 columns held in registers (thread t
// owns rows t, t + 1024, ...): per column the pivot search (largest |re| + |im|,
// smallest row on ties, as block_pivot_search), the swap and the rank-1 update
// in registers; then the sub-panel's swaps on the panel's other columns, U12 =
// L11^-1 A12 and one rank-8 update of the panel's remaining columns. The panel
// goes through memory once per sub-panel instead of once per column.
template <int R>
__global__ __launch_bounds__(1024) void k_lu_panel_reg(float* A, int lda, long long sA, int n, int k0, int k1,
                                                       int* ipiv, int* info) {
  constexpr int SP = 8, NT = 1024;
  float* a = A + 2 * sA * blockIdx.z;
  int* pv = ipiv + (long long)n * blockIdx.z;
  int* inf = info + blockIdx.z;
  __shared__ float wv[NT / 64];
  __shared__ int wi[NT / 64];
  __shared__ int piv_s[SP];
  __shared__ float pr_re[SP], pr_im[SP], xr_re[SP], xr_im[SP];
  __shared__ float l11_re[SP][SP], l11_im[SP][SP];
  __shared__ float u12_re[SP][kLuNb], u12_im[SP][kLuNb];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j0 = k0; j0 < k1; j0 += SP) {
    const int w = k1 - j0 < SP ? k1 - j0 : SP;
    float x_re[R][SP], x_im[R][SP];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = tid + NT * q;
#pragma unroll
      for (int c = 0; c < SP; ++c) {
        const bool in = r < n && c < w;
        const long long o = (long long)r + (long long)(j0 + c) * lda;
        x_re[q][c] = in ? a[2 * o] : 0.0f;
        x_im[q][c] = in ? a[2 * o + 1] : 0.0f;
      }
    }
#pragma unroll
    for (int c = 0; c < SP; ++c) {
      if (c >= w) break;
      const int j = j0 + c;
      // pivot: the largest |re| + |im| over rows >= j, the smallest row on ties
      float best = -1.0f;
      int bi = n;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        if (r >= j && r < n) {
          const float v = fabsf(x_re[q][c]) + fabsf(x_im[q][c]);
          if (v > best) { best = v; bi = r; }
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
      __syncthreads();
      if (tid == 0) {
        float b = wv[0];
        int i = wi[0];
        for (int k = 1; k < NT / 64; ++k)
          if (wv[k] > b || (wv[k] == b && wi[k] < i)) { b = wv[k]; i = wi[k]; }
        piv_s[c] = i;
        pv[j] = i;
        if (!(b > 0.0f) && *inf == 0) *inf = j + 1;
      }
      __syncthreads();
      const int p = piv_s[c];
      // swap rows j and p of the sub-panel through LDS
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        if (r == p) {
#pragma unroll
          for (int cc = 0; cc < SP; ++cc) { pr_re[cc] = x_re[q][cc]; pr_im[cc] = x_im[q][cc]; }
        }
        if (r == j) {
#pragma unroll
          for (int cc = 0; cc < SP; ++cc) { xr_re[cc] = x_re[q][cc]; xr_im[cc] = x_im[q][cc]; }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        if (p != j && r == j) {
#pragma unroll
          for (int cc = 0; cc < SP; ++cc) { x_re[q][cc] = pr_re[cc]; x_im[q][cc] = pr_im[cc]; }
        }
        if (p != j && r == p) {
#pragma unroll
          for (int cc = 0; cc < SP; ++cc) { x_re[q][cc] = xr_re[cc]; x_im[q][cc] = xr_im[cc]; }
        }
      }
      // the pivot row (row j after the swap) is pr; scale L(:, j), rank-1 update of the sub-panel
      const float dr = pr_re[c], di = pr_im[c];
      const float den = dr * dr + di * di;
      float rr = 0.0f, ri = 0.0f;
      if (den > 0.0f) { rr = dr / den; ri = -di / den; }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        if (r > j && r < n) {
          const float xr = x_re[q][c], xi = x_im[q][c];
          const float lr = xr * rr - xi * ri, li = xr * ri + xi * rr;
          x_re[q][c] = lr;
          x_im[q][c] = li;
#pragma unroll
          for (int cc = c + 1; cc < SP; ++cc) {
            x_re[q][cc] -= lr * pr_re[cc] - li * pr_im[cc];
            x_im[q][cc] -= lr * pr_im[cc] + li * pr_re[cc];
          }
        }
      }
      __syncthreads();                        // pr / xr / wv reused by the next column
    }
    // the factored sub-panel back to memory; L11 (rows j0 .. j0 + w - 1) to LDS
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = tid + NT * q;
      if (r >= j0 && r < n)
#pragma unroll
        for (int c = 0; c < SP; ++c)
          if (c < w) {
            const long long o = (long long)r + (long long)(j0 + c) * lda;
            a[2 * o] = x_re[q][c];
            a[2 * o + 1] = x_im[q][c];
          }
      if (r >= j0 && r < j0 + w)
#pragma unroll
        for (int c = 0; c < SP; ++c) { l11_re[r - j0][c] = x_re[q][c]; l11_im[r - j0][c] = x_im[q][c]; }
    }
    __syncthreads();
    // the sub-panel's swaps on the panel's other columns, in order
    const int nother = (k1 - k0) - w;
    if (tid < nother) {
      const int col = tid < j0 - k0 ? k0 + tid : j0 + w + (tid - (j0 - k0));
      float* cp = a + 2 * (long long)col * lda;
      for (int c = 0; c < w; ++c) {
        const int j = j0 + c, p = piv_s[c];
        if (p != j) {
          const float tr = cp[2 * j], ti = cp[2 * j + 1];
          cp[2 * j] = cp[2 * p];
          cp[2 * j + 1] = cp[2 * p + 1];
          cp[2 * p] = tr;
          cp[2 * p + 1] = ti;
        }
      }
    }
    __syncthreads();
    // U12 = L11^-1 A12 (unit lower L11), one thread per remaining panel column
    const int nrest = k1 - j0 - w;
    if (tid < nrest) {
      const int col = j0 + w + tid;
      float ur[SP], ui[SP];
#pragma unroll
      for (int c = 0; c < SP; ++c) {
        if (c < w) {
          const long long o = (long long)(j0 + c) + (long long)col * lda;
          ur[c] = a[2 * o];
          ui[c] = a[2 * o + 1];
        } else {
          ur[c] = ui[c] = 0.0f;
        }
      }
#pragma unroll
      for (int c = 1; c < SP; ++c)
#pragma unroll
        for (int k = 0; k < c; ++k) {
          ur[c] -= l11_re[c][k] * ur[k] - l11_im[c][k] * ui[k];
          ui[c] -= l11_re[c][k] * ui[k] + l11_im[c][k] * ur[k];
        }
#pragma unroll
      for (int c = 0; c < SP; ++c) {
        if (c < w) {
          const long long o = (long long)(j0 + c) + (long long)col * lda;
          a[2 * o] = ur[c];
          a[2 * o + 1] = ui[c];
        }
        u12_re[c][tid] = ur[c];
        u12_im[c][tid] = ui[c];
      }
    }
    __syncthreads();
    // rank-w update of rows >= j0 + w, columns [j0 + w, k1): A22 -= L21 U12
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = tid + NT * q;
      if (r >= j0 + w && r < n)
        for (int t = 0; t < nrest; ++t) {
          const long long o = (long long)r + (long long)(j0 + w + t) * lda;
          float vr = a[2 * o], vi = a[2 * o + 1];
#pragma unroll
          for (int c = 0; c < SP; ++c) {
            vr -= x_re[q][c] * u12_re[c][t] - x_im[q][c] * u12_im[c][t];
            vi -= x_re[q][c] * u12_im[c][t] + x_im[q][c] * u12_re[c][t];
          }
          a[2 * o] = vr;
          a[2 * o + 1] = vi;
        }
    }
    __syncthreads();
  }
}

// apply the panel's swaps (rows [k0,k1)) to columns [0,k0) and [k1,n)
template <typename T>
__device__ __forceinline__ void laswp_body(T* a, int lda, int n, int k0, int k1, const int* pv, int c) {
  if (c >= n - (k1 - k0)) return;
  if (c >= k0) c += k1 - k0;
  T* col = a + 2 * (long long)c * lda;
  for (int j = k0; j < k1; ++j) {
    const int p = pv[j];
    if (p != j) {
      const T xr = col[2 * j], xi = col[2 * j + 1];
      col[2 * j] = col[2 * p];
      col[2 * j + 1] = col[2 * p + 1];
      col[2 * p] = xr;
      col[2 * p + 1] = xi;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_laswp(T* A, int lda, long long sA, int n, int k0, int k1,
                                                  const int* ipiv) {
  laswp_body<T>(A + 2 * sA * blockIdx.z, lda, n, k0, k1, ipiv + (long long)n * blockIdx.z,
                blockIdx.x * blockDim.x + threadIdx.x);
}

// X[kb x ncols] <- L^{-1} X (unit lower, L = T[0:kb,0:kb]) or U^{-1} X (upper).
// 64 lanes per right-hand side, 4 right-hand sides per workgroup.
template <typename T, bool UPPER>
__device__ __forceinline__ void trsm_body(const T* t, int ldt, int kb, T* x, int ldx, int ncols, int bx) {
  __shared__ T l_re[kLuNb][kLuNb + 1], l_im[kLuNb][kLuNb + 1];
  __shared__ T x_re[4][kLuNb], x_im[4][kLuNb];
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
    const int r = e % kb, c = e / kb;
    const long long off = (long long)r + (long long)c * ldt;
    l_re[r][c] = t[2 * off];
    l_im[r][c] = t[2 * off + 1];
  }
  const int q = threadIdx.x >> 6, r = threadIdx.x & 63;
  const int col = bx * 4 + q;
  const bool act = col < ncols && r < kb;
  if (act) {
    const long long off = (long long)r + (long long)col * ldx;
    x_re[q][r] = x[2 * off];
    x_im[q][r] = x[2 * off + 1];
  }
  __syncthreads();
  if (!UPPER) {
    for (int k = 0; k < kb; ++k) {
      if (act && r > k) {
        const T lr = l_re[r][k], li = l_im[r][k], br = x_re[q][k], bi = x_im[q][k];
        x_re[q][r] -= lr * br - li * bi;
        x_im[q][r] -= lr * bi + li * br;
      }
      __syncthreads();
    }
  } else {
    for (int k = kb - 1; k >= 0; --k) {
      if (act && r == k) {
        const T dr = l_re[k][k], di = l_im[k][k], den = dr * dr + di * di;
        const T xr = x_re[q][k], xi = x_im[q][k];
        x_re[q][k] = (xr * dr + xi * di) / den;
        x_im[q][k] = (xi * dr - xr * di) / den;
      }
      __syncthreads();
      if (act && r < k) {
        const T lr = l_re[r][k], li = l_im[r][k], br = x_re[q][k], bi = x_im[q][k];
        x_re[q][r] -= lr * br - li * bi;
        x_im[q][r] -= lr * bi + li * br;
      }
      __syncthreads();
    }
  }
  if (act) {
    const long long off = (long long)r + (long long)col * ldx;
    x[2 * off] = x_re[q][r];
    x[2 * off + 1] = x_im[q][r];
  }
}

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_trsm_block(const T* Tm, int ldt, long long sT, int kb, T* X, int ldx,
                                                    long long sX, int ncols) {
  trsm_body<T, UPPER>(Tm + 2 * sT * blockIdx.z, ldt, kb, X + 2 * sX * blockIdx.z, ldx, ncols, blockIdx.x);
}

size_t lu_coop_ws_bytes(int n) {
  const size_t W = (size_t)((n + kCpRows - 1) / kCpRows);
  return 2 * W * sizeof(double) + 2 * W * sizeof(int) + 2 * W * kLuNb * 2 * sizeof(double) +
         2 * kLuNb * 2 * sizeof(double) + 1024;
}

// workgroups of k_lu_panel_coop<T> the device holds at once (0: no cooperative launch)
template <typename T>
static int lu_coop_capacity() {
  static int cap = -1;
  if (cap < 0) {
    int dev = 0, cus = 0, per = 0, coop = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_lu_panel_coop<T>, 256, 0) != hipSuccess)
      coop = 0;
    cap = coop ? cus * per : 0;
  }
  return cap;
}

template <typename T>
hipError_t lu_factor(hipStream_t s, T* A, int lda, long long sA, int n, int* ipiv, int* info,
                     unsigned* counter, int batch, void* coop_ws) {
  const bool reg = sizeof(T) == sizeof(float) && n <= 2048;
  bool coop = coop_ws && batch == 1 && !reg && (n + kCpRows - 1) / kCpRows <= lu_coop_capacity<T>();
  for (int k0 = 0; k0 < n; k0 += kLuNb) {
    const int k1 = k0 + kLuNb < n ? k0 + kLuNb : n;
    if (coop) {
      const int W = (n - k0 + kCpRows - 1) / kCpRows, Wmax = (n + kCpRows - 1) / kCpRows;
      char* b = static_cast<char*>(coop_ws);
      T* cval = reinterpret_cast<T*>(b);
      int* crow = reinterpret_cast<int*>(b + 2 * (size_t)Wmax * sizeof(double));
      T* cdat = reinterpret_cast<T*>(b + ((2 * (size_t)Wmax * (sizeof(double) + sizeof(int)) + 255) & ~(size_t)255));
      T* jdat = cdat + 2 * (size_t)Wmax * kLuNb * 2;
      int kk0 = k0, kk1 = k1, nn = n, ld = lda;
      void* args[] = {&A, &ld, &nn, &kk0, &kk1, &ipiv, &info, &cval, &crow, &cdat, &jdat};
      hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_lu_panel_coop<T>), dim3(W), dim3(256),
                                                args, 0, s);
      if (e != hipSuccess) {
        if (k0 > 0) return e;
        (void)hipGetLastError();   // refused before anything ran: per-column panels instead
        coop = false;
      }
    }
    if (coop) {
    } else if (reg) {
      float* Af = reinterpret_cast<float*>(A);
      if (n <= 1024) hipLaunchKernelGGL((k_lu_panel_reg<1>), dim3(1, 1, batch), dim3(1024), 0, s, Af, lda, sA, n, k0, k1, ipiv, info);
      else hipLaunchKernelGGL((k_lu_panel_reg<2>), dim3(1, 1, batch), dim3(1024), 0, s, Af, lda, sA, n, k0, k1, ipiv, info);
    } else if (sizeof(T) == sizeof(float) && n <= 4096) {
      hipLaunchKernelGGL((k_lu_panel<T>), dim3(1, 1, batch), dim3(256), 0, s, A, lda, sA, n, k0, k1, ipiv, info);
    } else {
      hipLaunchKernelGGL((k_lu_pivot<T>), dim3(1, 1, batch), dim3(256), 0, s, A, lda, sA, n, k0, k0, k1, ipiv, info);
      for (int j = k0; j < k1; ++j) {
        const int rows = n - j - 1;
        const int gx = rows > 0 ? (rows + 255) / 256 : 1;
        hipLaunchKernelGGL((k_lu_col<T>), dim3(gx, 1, batch), dim3(256), 0, s, A, lda, sA, n, j, k0, k1, ipiv, info,
                           counter);
      }
    }
    const int other = n - (k1 - k0);
    if (other > 0)
      hipLaunchKernelGGL((k_lu_laswp<T>), dim3((other + 255) / 256, 1, batch), dim3(256), 0, s, A, lda, sA, n, k0, k1, ipiv);
    if (k1 < n) {
      const int nc = n - k1;
      const long long d = (long long)k0 + (long long)k0 * lda;
      hipLaunchKernelGGL((k_trsm_block<T, false>), dim3((nc + 3) / 4, 1, batch), dim3(256), 0, s,
                         A + 2 * d, lda, sA, k1 - k0, A + 2 * ((long long)k0 + (long long)k1 * lda), lda, sA, nc);
      hipError_t e = cgemm_3m<T>(s, false, n - k1, nc, k1 - k0, T(-1), T(0),
                              A + 2 * ((long long)k1 + (long long)k0 * lda), lda, sA,
                              A + 2 * ((long long)k0 + (long long)k1 * lda), lda, sA, T(1), T(0),
                              A + 2 * ((long long)k1 + (long long)k1 * lda), lda, sA, batch);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

// B[n x nrhs] rows permuted by the pivot sequence (one thread per column)
template <typename T>
__device__ __forceinline__ void apply_pivots_body(T* B, int ldb, int n, int nrhs, const int* pv, int c) {
  if (c >= nrhs) return;
  T* col = B + 2 * (long long)c * ldb;
  for (int j = 0; j < n; ++j) {
    const int p = pv[j];
    if (p != j) {
      const T xr = col[2 * j], xi = col[2 * j + 1];
      col[2 * j] = col[2 * p];
      col[2 * j + 1] = col[2 * p + 1];
      col[2 * p] = xr;
      col[2 * p + 1] = xi;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_apply_pivots(T* B, int ldb, long long sB, int n, int nrhs,
                                                      const int* ipiv) {
  apply_pivots_body<T>(B + 2 * sB * blockIdx.z, ldb, n, nrhs, ipiv + (long long)n * blockIdx.z,
                       blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
hipError_t lu_solve(hipStream_t s, const T* A, int lda, long long sA, int n, const int* ipiv, T* B, int ldb,
                    long long sB, int nrhs, int batch) {
  hipLaunchKernelGGL((k_apply_pivots<T>), dim3((nrhs + 255) / 256, 1, batch), dim3(256), 0, s, B, ldb, sB, n, nrhs, ipiv);
  // forward: unit lower
  for (int k0 = 0; k0 < n; k0 += kLuNb) {
    const int kb = (n - k0) < kLuNb ? n - k0 : kLuNb;
    hipLaunchKernelGGL((k_trsm_block<T, false>), dim3((nrhs + 3) / 4, 1, batch), dim3(256), 0, s,
                       A + 2 * ((long long)k0 + (long long)k0 * lda), lda, sA, kb, B + 2 * (long long)k0, ldb, sB, nrhs);
    if (k0 + kb < n) {
      hipError_t e = cgemm<T>(s, false, n - k0 - kb, nrhs, kb, T(-1), T(0),
                              A + 2 * ((long long)(k0 + kb) + (long long)k0 * lda), lda, sA,
                              B + 2 * (long long)k0, ldb, sB, T(1), T(0), B + 2 * (long long)(k0 + kb), ldb, sB, batch);
      if (e != hipSuccess) return e;
    }
  }
  // backward: upper
  const int nblk = (n + kLuNb - 1) / kLuNb;
  for (int b = nblk - 1; b >= 0; --b) {
    const int k0 = b * kLuNb, kb = (n - k0) < kLuNb ? n - k0 : kLuNb;
    hipLaunchKernelGGL((k_trsm_block<T, true>), dim3((nrhs + 3) / 4, 1, batch), dim3(256), 0, s,
                       A + 2 * ((long long)k0 + (long long)k0 * lda), lda, sA, kb, B + 2 * (long long)k0, ldb, sB, nrhs);
    if (k0 > 0) {
      hipError_t e = cgemm<T>(s, false, k0, nrhs, kb, T(-1), T(0), A + 2 * ((long long)k0 * lda), lda, sA,
                              B + 2 * (long long)k0, ldb, sB, T(1), T(0), B, ldb, sB, batch);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}


template <typename T>
hipError_t lu_inverse_nopiv(hipStream_t s, const T* A, int lda, long long sA, int n, T* B, int ldb, long long sB,
                            int batch) {
  // forward: L Y = I; Y is unit lower triangular, so rows [k0, k0 + kb) hold
  // nonzeros in columns < k0 + kb only
  for (int k0 = 0; k0 < n; k0 += kLuNb) {
    const int kb = (n - k0) < kLuNb ? n - k0 : kLuNb;
    const int nc = k0 + kb;
    hipLaunchKernelGGL((k_trsm_block<T, false>), dim3((nc + 3) / 4, 1, batch), dim3(256), 0, s,
                       A + 2 * ((long long)k0 + (long long)k0 * lda), lda, sA, kb, B + 2 * (long long)k0, ldb, sB, nc);
    if (k0 + kb < n) {
      hipError_t e = cgemm<T>(s, false, n - k0 - kb, nc, kb, T(-1), T(0),
                              A + 2 * ((long long)(k0 + kb) + (long long)k0 * lda), lda, sA,
                              B + 2 * (long long)k0, ldb, sB, T(1), T(0), B + 2 * (long long)(k0 + kb), ldb, sB, batch);
      if (e != hipSuccess) return e;
    }
  }
  // backward: U X = Y, all columns
  const int nblk = (n + kLuNb - 1) / kLuNb;
  for (int b = nblk - 1; b >= 0; --b) {
    const int k0 = b * kLuNb, kb = (n - k0) < kLuNb ? n - k0 : kLuNb;
    hipLaunchKernelGGL((k_trsm_block<T, true>), dim3((n + 3) / 4, 1, batch), dim3(256), 0, s,
                       A + 2 * ((long long)k0 + (long long)k0 * lda), lda, sA, kb, B + 2 * (long long)k0, ldb, sB, n);
    if (k0 > 0) {
      hipError_t e = cgemm<T>(s, false, k0, n, kb, T(-1), T(0), A + 2 * ((long long)k0 * lda), lda, sA,
                              B + 2 * (long long)k0, ldb, sB, T(1), T(0), B, ldb, sB, batch);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grouped factorisation and solve: matrices of different orders in one launch
// per step (blockIdx.z = member). The host walks the panels of the largest
// member; a member whose order is exhausted exits, and every grid is sized for
// the largest member (smaller members' surplus workgroups exit at once).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

template <typename T>
__global__ __launch_bounds__(256) void k_lu_pivot_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  if (k0 >= g.n) return;
  __shared__ int prow;
  __shared__ T pval;
  block_pivot_search(g.A, (int)g.lda, g.n, k0, k0, &prow, &pval);
  block_apply_pivot(g.A, (int)g.lda, k0, prow, pval, k0, imin(k0 + kLuNb, g.n), g.ipiv, g.info);
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_col_g(const LuDesc<T>* __restrict__ d, int j, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  if (j >= g.n) return;
  const int rows = g.n - j - 1, nblk = rows > 0 ? (rows + 255) / 256 : 1;
  if ((int)blockIdx.x >= nblk) return;
  lu_col_body<T>(g.A, (int)g.lda, g.n, j, k0, imin(k0 + kLuNb, g.n), g.ipiv, g.info, g.counter, nblk, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_laswp_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  if (k0 >= g.n) return;
  laswp_body<T>(g.A, (int)g.lda, g.n, k0, imin(k0 + kLuNb, g.n), g.ipiv, blockIdx.x * blockDim.x + threadIdx.x);
}

// block row of U: L11^{-1} A12
template <typename T>
__global__ __launch_bounds__(256) void k_lu_trsm_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  const int k1 = imin(k0 + kLuNb, g.n);
  if (k1 >= g.n || (int)blockIdx.x * 4 >= g.n - k1) return;
  trsm_body<T, false>(g.A + 2 * ((long long)k0 + (long long)k0 * g.lda), (int)g.lda, k1 - k0,
                      g.A + 2 * ((long long)k0 + (long long)k1 * g.lda), (int)g.lda, g.n - k1, blockIdx.x);
}

// trailing update A22 -= A21 A12
template <typename T, int KB>
__global__ __launch_bounds__(256, 2) void k_lu_update_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  const int k1 = imin(k0 + kLuNb, g.n), nr = g.n - k1;
  const int row0 = blockIdx.x * kBM, col0 = blockIdx.y * kBN;
  if (nr <= 0 || row0 >= nr || col0 >= nr) return;
  const long long lda = g.lda;
  cgemm_tile<T, false, KB, true>(nr, nr, k1 - k0, T(-1), T(0), g.A + 2 * ((long long)k1 + (long long)k0 * lda),
                                 (int)lda, g.A + 2 * ((long long)k0 + (long long)k1 * lda), (int)lda, T(1), T(0),
                                 g.A + 2 * ((long long)k1 + (long long)k1 * lda), (int)lda, row0, col0);
}

template <typename T>
hipError_t lu_factor_grouped(hipStream_t s, const LuDesc<T>* d, int count, int nmax, int kb) {
  if (count <= 0 || nmax <= 0) return hipSuccess;
  for (int k0 = 0; k0 < nmax; k0 += kLuNb) {
    const int k1 = k0 + kLuNb < nmax ? k0 + kLuNb : nmax;
    hipLaunchKernelGGL((k_lu_pivot_g<T>), dim3(1, 1, count), dim3(256), 0, s, d, k0);
    for (int j = k0; j < k1; ++j) {
      const int rows = nmax - j - 1;
      const int gx = rows > 0 ? (rows + 255) / 256 : 1;
      hipLaunchKernelGGL((k_lu_col_g<T>), dim3(gx, 1, count), dim3(256), 0, s, d, j, k0);
    }
    hipLaunchKernelGGL((k_lu_laswp_g<T>), dim3((nmax + 255) / 256, 1, count), dim3(256), 0, s, d, k0);
    if (k1 < nmax) {
      const int nc = nmax - k1;
      hipLaunchKernelGGL((k_lu_trsm_g<T>), dim3((nc + 3) / 4, 1, count), dim3(256), 0, s, d, k0);
      const dim3 ug((nc + kBM - 1) / kBM, (nc + kBN - 1) / kBN, count);
      if (kb == 8) hipLaunchKernelGGL((k_lu_update_g<T, 8>), ug, dim3(256), 0, s, d, k0);
      else if (kb == 32) hipLaunchKernelGGL((k_lu_update_g<T, 32>), ug, dim3(256), 0, s, d, k0);
      else hipLaunchKernelGGL((k_lu_update_g<T, 16>), ug, dim3(256), 0, s, d, k0);
    }
  }
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_apply_pivots_g(const LuDesc<T>* __restrict__ d) {
  const LuDesc<T> g = d[blockIdx.z];
  apply_pivots_body<T>(g.B, (int)g.ldb, g.n, g.nrhs, g.ipiv, blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_solve_trsm_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  if (k0 >= g.n || (int)blockIdx.x * 4 >= g.nrhs) return;
  const int kb = imin(kLuNb, g.n - k0);
  trsm_body<T, UPPER>(g.A + 2 * ((long long)k0 + (long long)k0 * g.lda), (int)g.lda, kb, g.B + 2 * (long long)k0,
                      (int)g.ldb, g.nrhs, blockIdx.x);
}

// forward: B[k0+kb:, :] -= A[k0+kb:, k0:k0+kb] B[k0:k0+kb, :]
// backward: B[0:k0, :] -= A[0:k0, k0:k0+kb] B[k0:k0+kb, :]
template <typename T, bool BACK>
__global__ __launch_bounds__(256) void k_solve_update_g(const LuDesc<T>* __restrict__ d, int k0) {
  const LuDesc<T> g = d[blockIdx.z];
  if (k0 >= g.n) return;
  const int kb = imin(kLuNb, g.n - k0);
  const int r0 = BACK ? 0 : k0 + kb, nr = BACK ? k0 : g.n - k0 - kb;
  const int row0 = blockIdx.x * kBM, col0 = blockIdx.y * kBN;
  if (nr <= 0 || row0 >= nr || col0 >= g.nrhs) return;
  cgemm_tile<T, false>(nr, g.nrhs, kb, T(-1), T(0), g.A + 2 * ((long long)r0 + (long long)k0 * g.lda), (int)g.lda,
                       g.B + 2 * (long long)k0, (int)g.ldb, T(1), T(0), g.B + 2 * (long long)r0, (int)g.ldb, row0,
                       col0);
}

template <typename T>
hipError_t lu_solve_grouped(hipStream_t s, const LuDesc<T>* d, int count, int nmax, int nrhs_max) {
  if (count <= 0 || nmax <= 0 || nrhs_max <= 0) return hipSuccess;
  const dim3 tg((nrhs_max + 3) / 4, 1, count);
  hipLaunchKernelGGL((k_apply_pivots_g<T>), dim3((nrhs_max + 255) / 256, 1, count), dim3(256), 0, s, d);
  for (int k0 = 0; k0 < nmax; k0 += kLuNb) {
    hipLaunchKernelGGL((k_solve_trsm_g<T, false>), tg, dim3(256), 0, s, d, k0);
    const int nr = nmax - k0 - kLuNb;
    if (nr > 0)
      hipLaunchKernelGGL((k_solve_update_g<T, false>), dim3((nr + kBM - 1) / kBM, (nrhs_max + kBN - 1) / kBN, count),
                         dim3(256), 0, s, d, k0);
  }
  const int nblk = (nmax + kLuNb - 1) / kLuNb;
  for (int b = nblk - 1; b >= 0; --b) {
    const int k0 = b * kLuNb;
    hipLaunchKernelGGL((k_solve_trsm_g<T, true>), tg, dim3(256), 0, s, d, k0);
    if (k0 > 0)
      hipLaunchKernelGGL((k_solve_update_g<T, true>), dim3((k0 + kBM - 1) / kBM, (nrhs_max + kBN - 1) / kBN, count),
                         dim3(256), 0, s, d, k0);
  }
  return hipGetLastError();
}

// per member: dst (FP32) <- src (FP64), or dst (FP64) (+)= src (FP32)
__global__ void k_convert_g(const ConvDesc* __restrict__ d, int to_single, int overwrite) {
  const ConvDesc g = d[blockIdx.z];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.m * g.n) return;
  const int r = (int)(e % g.m), c = (int)(e / g.m);
  const long long o64 = (long long)r + (long long)c * g.ld64, o32 = (long long)r + (long long)c * g.ld32;
  if (to_single) {
    g.f32[2 * o32] = (float)g.f64[2 * o64];
    g.f32[2 * o32 + 1] = (float)g.f64[2 * o64 + 1];
  } else if (overwrite) {
    g.f64[2 * o64] = g.f32[2 * o32];
    g.f64[2 * o64 + 1] = g.f32[2 * o32 + 1];
  } else {
    g.f64[2 * o64] += g.f32[2 * o32];
    g.f64[2 * o64 + 1] += g.f32[2 * o32 + 1];
  }
}

hipError_t convert_grouped(hipStream_t s, const ConvDesc* d, int count, long long max_elems, bool to_single,
                           bool overwrite) {
  if (count <= 0 || max_elems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_convert_g, dim3((unsigned)((max_elems + 255) / 256), 1, count), dim3(256), 0, s, d,
                     to_single ? 1 : 0, overwrite ? 1 : 0);
  return hipGetLastError();
}

// per member: dst <- src (FP64 m x n blocks)
__global__ void k_copy_g(const CopyDesc* __restrict__ d) {
  const CopyDesc g = d[blockIdx.z];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.m * g.n) return;
  const int r = (int)(e % g.m), c = (int)(e / g.m);
  const long long os = (long long)r + (long long)c * g.lds, od = (long long)r + (long long)c * g.ldd;
  g.dst[2 * od] = g.src[2 * os];
  g.dst[2 * od + 1] = g.src[2 * os + 1];
}
hipError_t copy_grouped(hipStream_t s, const CopyDesc* d, int count, long long max_elems) {
  if (count <= 0 || max_elems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_g, dim3((unsigned)((max_elems + 255) / 256), 1, count), dim3(256), 0, s, d);
  return hipGetLastError();
}

// per member: largest |re|,|im| of the FP32 correction over that of the FP64
// solution (one workgroup per member)
__global__ __launch_bounds__(256) void k_correction_ratio_g(const ConvDesc* __restrict__ d, double* out) {
  const ConvDesc g = d[blockIdx.x];
  __shared__ double s32[256], s64[256];
  double m32 = 0.0, m64 = 0.0;
  for (long long e = threadIdx.x; e < (long long)g.m * g.n; e += blockDim.x) {
    const int r = (int)(e % g.m), c = (int)(e / g.m);
    const long long o64 = (long long)r + (long long)c * g.ld64, o32 = (long long)r + (long long)c * g.ld32;
    m32 = fmax(m32, fmax(fabs((double)g.f32[2 * o32]), fabs((double)g.f32[2 * o32 + 1])));
    m64 = fmax(m64, fmax(fabs(g.f64[2 * o64]), fabs(g.f64[2 * o64 + 1])));
  }
  s32[threadIdx.x] = m32;
  s64[threadIdx.x] = m64;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s32[threadIdx.x] = fmax(s32[threadIdx.x], s32[threadIdx.x + st]);
      s64[threadIdx.x] = fmax(s64[threadIdx.x], s64[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s32[0] / (s64[0] > 0.0 ? s64[0] : 1.0);
}
hipError_t correction_ratio_grouped(hipStream_t s, const ConvDesc* d, int count, double* out) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_correction_ratio_g, dim3(count), dim3(256), 0, s, d, out);
  return hipGetLastError();
}

// fill n ints with a value (graph-capturable, unlike a memset node)
__global__ void k_fill_i32(int* p, long long n, int v) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}
hipError_t fill_i32(hipStream_t s, int* p, long long n, int v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// complex128 -> complex64 copy of an m x n column-major block (ld may differ)
__global__ void k_demote(const double* __restrict__ src, int lds, long long ss, float* __restrict__ dst, int ldd,
                         long long sd, int m, int n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  const double* s = src + 2 * ss * blockIdx.z + 2 * ((long long)r + (long long)c * lds);
  float* d = dst + 2 * sd * blockIdx.z + 2 * ((long long)r + (long long)c * ldd);
  d[0] = (float)s[0];
  d[1] = (float)s[1];
}
__global__ void k_promote_add(const float* __restrict__ src, int lds, long long ss, double* __restrict__ dst,
                              int ldd, long long sd, int m, int n, int overwrite) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  const float* s = src + 2 * ss * blockIdx.z + 2 * ((long long)r + (long long)c * lds);
  double* d = dst + 2 * sd * blockIdx.z + 2 * ((long long)r + (long long)c * ldd);
  if (overwrite) { d[0] = s[0]; d[1] = s[1]; }
  else { d[0] += s[0]; d[1] += s[1]; }
}

hipError_t demote(hipStream_t s, const double* src, int lds, long long ss, float* dst, int ldd, long long sd, int m,
                  int n, int batch) {
  const long long tot = (long long)m * n;
  hipLaunchKernelGGL(k_demote, dim3((unsigned)((tot + 255) / 256), 1, batch), dim3(256), 0, s, src, lds, ss, dst, ldd, sd, m, n);
  return hipGetLastError();
}
hipError_t promote_add(hipStream_t s, const float* src, int lds, long long ss, double* dst, int ldd, long long sd,
                       int m, int n, int batch, bool overwrite) {
  const long long tot = (long long)m * n;
  hipLaunchKernelGGL(k_promote_add, dim3((unsigned)((tot + 255) / 256), 1, batch), dim3(256), 0, s, src, lds, ss, dst, ldd, sd, m, n, overwrite ? 1 : 0);
  return hipGetLastError();
}

template hipError_t cgemm<double>(hipStream_t, bool, int, int, int, double, double, const double*, int, long long,
                                  const double*, int, long long, double, double, double*, int, long long, int);
template hipError_t cgemm<float>(hipStream_t, bool, int, int, int, float, float, const float*, int, long long,
                                 const float*, int, long long, float, float, float*, int, long long, int);
template hipError_t cgemm_grouped<double>(hipStream_t, bool, const GemmDesc<double>*, int, int, int, int);
template hipError_t cgemm_grouped<float>(hipStream_t, bool, const GemmDesc<float>*, int, int, int, int);
template hipError_t lu_factor_grouped<double>(hipStream_t, const LuDesc<double>*, int, int, int);
template hipError_t lu_factor_grouped<float>(hipStream_t, const LuDesc<float>*, int, int, int);
template hipError_t lu_solve_grouped<double>(hipStream_t, const LuDesc<double>*, int, int, int);
template hipError_t lu_solve_grouped<float>(hipStream_t, const LuDesc<float>*, int, int, int);
template hipError_t lu_factor<double>(hipStream_t, double*, int, long long, int, int*, int*, unsigned*, int, void*);
template hipError_t lu_factor<float>(hipStream_t, float*, int, long long, int, int*, int*, unsigned*, int, void*);
template hipError_t lu_solve<double>(hipStream_t, const double*, int, long long, int, const int*, double*, int,
                                     long long, int, int);
template hipError_t lu_solve<float>(hipStream_t, const float*, int, long long, int, const int*, float*, int,
                                    long long, int, int);
template hipError_t lu_inverse_nopiv<float>(hipStream_t, const float*, int, long long, int, float*, int, long long,
                                            int);

// ---------------------------------------------------------------------------
// Explicit inverse by blocked Gauss-Jordan elimination, no row interchanges
// ---------------------------------------------------------------------------
// For matrices whose elimination needs no pivoting (the diagonal blocks of the
// influence matrix: 2 pi on the diagonal plus a compact part; growth factor
// ~1): step k, with the 64 rows and columns R_k = [k0, k0 + w),
//   Pinv = (A[R_k, R_k])^-1                        (every workgroup, in LDS)
//   T    = Pinv A[R_k, :], T[:, R_k] = Pinv        (k_gj_row, one 64-column tile each)
//   C    = A[:, R_k] with rows R_k zeroed; A[:, R_k] = 0 off R_k; A[R_k, :] = T   (k_gj_col)
//   A   -= C T                                     (one n x n x w complex GEMM on MFMA)
// after the last step A holds A^-1. n^3 complex multiply-adds in all, nearly all
// in the GEMMs, against LU + triangular inverses whose panels are latency bound.
constexpr int kGj = 64;

// the pivot tile inverted by one wave, lane r holding row r in registers: per
// column j the pivot row is broadcast from lane j with readlane (j is a
// constant of the unrolled loop) and every lane applies one rank-1 update,
//   g = f / p (lane != j, f = its entry in column j) or 1 - 1/p (lane j),
//   x[c] -= g row_j[c] (c != j),   x[j] = -g (lane != j) or 1/p (lane j),
// which leaves lane j's row scaled by 1/p and column j eliminated elsewhere.
// *bad = 1 on a zero pivot.
__device__ __forceinline__ float lane_bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// one pivot step; J is a compile-time constant, so column J of the row and the
// broadcast lane are static (no dynamic register index, no selects)
template <int J>
__device__ __forceinline__ void gj_wave_step(float (&xr)[kGj], float (&xi)[kGj], int lane, int* bad) {
  const float fr = xr[J], fi = xi[J];
  const float dr = lane_bcast(fr, J), di = lane_bcast(fi, J);
  const float den = dr * dr + di * di;
  float rr = 0.0f, ri = 0.0f;
  if (den > 0.0f) { rr = dr / den; ri = -di / den; }
  else if (lane == 0) *bad = 1;
  float gr = fr * rr - fi * ri, gi = fr * ri + fi * rr;
  if (lane == J) { gr = 1.0f - rr; gi = -ri; }
#pragma unroll
  for (int c = 0; c < kGj; ++c) {
    if (c == J) continue;
    const float ur = lane_bcast(xr[c], J), ui = lane_bcast(xi[c], J);
    xr[c] -= gr * ur - gi * ui;
    xi[c] -= gr * ui + gi * ur;
  }
  xr[J] = lane == J ? rr : -gr;
  xi[J] = lane == J ? ri : -gi;
}

template <int... J>
__device__ __forceinline__ void gj_wave_steps(float (&xr)[kGj], float (&xi)[kGj], int lane, int* bad,
                                              std::integer_sequence<int, J...>) {
  (gj_wave_step<J>(xr, xi, lane, bad), ...);
}

__device__ __forceinline__ void gj_wave_invert(float (&xr)[kGj], float (&xi)[kGj], int lane, int* bad) {
  gj_wave_steps(xr, xi, lane, bad, std::make_integer_sequence<int, kGj>{});
}

// T[:, c0 .. c0 + 63] of step k0 (blockIdx.x = column tile, blockIdx.z = matrix):
// wave 0 inverts the pivot tile, then the four waves form Pinv A[R_k, tile]
__global__ __launch_bounds__(256) void k_gj_row(const float* __restrict__ A, int lda, long long sA, int n, int k0,
                                                float* __restrict__ Tb, long long sT, int* info) {
  __shared__ float pr[kGj][kGj + 1], pi[kGj][kGj + 1], ar[kGj][kGj + 1], ai[kGj][kGj + 1];
  __shared__ int bad;
  const float* a = A + 2 * sA * blockIdx.z;
  float* tb = Tb + 2 * sT * blockIdx.z;
  const int w = n - k0 < kGj ? n - k0 : kGj, c0 = kGj * blockIdx.x;
  const int wc = n - c0 < kGj ? n - c0 : kGj;
  const bool diag = c0 == k0;
  if (threadIdx.x == 0) bad = 0;
  if (!diag)
    for (int e = threadIdx.x; e < kGj * kGj; e += blockDim.x) {
      const int r = e % kGj, c = e / kGj;
      const bool ina = r < w && c < wc;
      const long long oa = (long long)(k0 + r) + (long long)(c0 + c) * lda;
      ar[r][c] = ina ? a[2 * oa] : 0.0f;
      ai[r][c] = ina ? a[2 * oa + 1] : 0.0f;
    }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    float xr[kGj], xi[kGj];
#pragma unroll
    for (int c = 0; c < kGj; ++c) {
      const bool in = r < w && c < w;
      const long long o = (long long)(k0 + r) + (long long)(k0 + c) * lda;
      xr[c] = in ? a[2 * o] : (r == c ? 1.0f : 0.0f);
      xi[c] = in ? a[2 * o + 1] : 0.0f;
    }
    gj_wave_invert(xr, xi, r, &bad);
#pragma unroll
    for (int c = 0; c < kGj; ++c) {
      pr[r][c] = xr[c];
      pi[r][c] = xi[c];
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0 && bad && info[blockIdx.z] == 0) info[blockIdx.z] = k0 + 1;
  const int i = threadIdx.x >> 2, q0 = 16 * (threadIdx.x & 3);
  if (i >= w) return;
  for (int q = 0; q < 16; ++q) {
    const int c = q0 + q;
    if (c >= wc) break;
    float sr = 0.0f, si = 0.0f;
    if (diag) {
      sr = pr[i][c];
      si = pi[i][c];
    } else {
      for (int k = 0; k < w; ++k) {
        sr += pr[i][k] * ar[k][c] - pi[i][k] * ai[k][c];
        si += pr[i][k] * ai[k][c] + pi[i][k] * ar[k][c];
      }
    }
    const long long o = (long long)i + (long long)(c0 + c) * kGj;
    tb[2 * o] = sr;
    tb[2 * o + 1] = si;
  }
}

// C = A[:, R_k] with rows R_k zero, A[:, R_k] = 0 off R_k, A[R_k, :] = T
__global__ __launch_bounds__(256) void k_gj_col(float* __restrict__ A, int lda, long long sA, int n, int k0,
                                                const float* __restrict__ Tb, long long sT, float* __restrict__ Cb,
                                                long long sC) {
  float* a = A + 2 * sA * blockIdx.z;
  const float* tb = Tb + 2 * sT * blockIdx.z;
  float* cb = Cb + 2 * sC * blockIdx.z;
  const int w = n - k0 < kGj ? n - k0 : kGj;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x, half = (long long)n * kGj;
  if (e < half) {
    const int r = (int)(e % n), c = (int)(e / n);
    if (c >= w) return;
    const long long o = (long long)r + (long long)(k0 + c) * lda, oc = (long long)r + (long long)c * n;
    if (r >= k0 && r < k0 + w) {
      cb[2 * oc] = 0.0f;
      cb[2 * oc + 1] = 0.0f;
    } else {
      cb[2 * oc] = a[2 * o];
      cb[2 * oc + 1] = a[2 * o + 1];
      a[2 * o] = 0.0f;
      a[2 * o + 1] = 0.0f;
    }
  } else if (e < 2 * half) {
    const long long f = e - half;
    const int r = (int)(f % kGj), c = (int)(f / kGj);
    if (r >= w) return;
    const long long o = (long long)(k0 + r) + (long long)c * lda, ot = (long long)r + (long long)c * kGj;
    a[2 * o] = tb[2 * ot];
    a[2 * o + 1] = tb[2 * ot + 1];
  }
}

__global__ void k_iota(int* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

hipError_t fill_iota(hipStream_t s, int* p, long long n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

size_t gj_inverse_ws_bytes(int n) { return 2 * 2 * sizeof(float) * (size_t)n * kGj; }

hipError_t gj_inverse_nopiv(hipStream_t s, float* A, int lda, long long sA, int n, void* ws, int* info, int batch) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  float* T = static_cast<float*>(ws);
  float* C = T + 2 * (size_t)batch * n * kGj;
  const long long sT = (long long)n * kGj, sC = (long long)n * kGj;
  const int ntile = (n + kGj - 1) / kGj;
  const long long colw = 2LL * n * kGj;
  hipError_t e;
  for (int k0 = 0; k0 < n; k0 += kGj) {
    const int w = n - k0 < kGj ? n - k0 : kGj;
    hipLaunchKernelGGL(k_gj_row, dim3(ntile, 1, batch), dim3(256), 0, s, A, lda, sA, n, k0, T, sT, info);
    hipLaunchKernelGGL(k_gj_col, dim3((unsigned)((colw + 255) / 256), 1, batch), dim3(256), 0, s, A, lda, sA, n, k0,
                       T, sT, C, sC);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = cgemm_3m<float>(s, false, n, n, w, -1.0f, 0.0f, C, n, sC, T, kGj, sT, 1.0f, 0.0f, A, lda, sA, batch)) !=
        hipSuccess)
      return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Complex GEMM with FP16 operands and FP32 accumulation (MFMA f32 16x16x32 f16)
// ---------------------------------------------------------------------------
// A workgroup of 4 waves per 128 x 64 output tile, each wave 64 x 32 (4 x 2
// MFMA tiles of 16 x 16). Per 32-k chunk the workgroup stages the tile's rows
// of A and columns of B in LDS (the packed layout, rows padded to 72 halves),
// so that each is read from memory once for the two waves that share it; the
// next chunk is loaded into registers while the current one is multiplied.
// Both operands give lane group q (lanes 16q..16q+15) the k of group q of the
// chunk, so the products pair equal k whatever order the matrix core sums them
// in. A complex product is four real MFMAs into Re / Im accumulators (Re gets
// Ar Br + Ai (-Bi)).
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
constexpr int kH16Ld = 72;   // halves per LDS row: 4 groups of 16 and a pad of 8

__global__ __launch_bounds__(256) void k_cgemm_h16(int m, int ncol, int kp, const _Float16* __restrict__ A,
                                                   long long sA, const _Float16* __restrict__ B, long long ldb,
                                                   long long sB, float* __restrict__ C, long long ldc, long long sC,
                                                   int cnt, int joint) {
  __shared__ __attribute__((aligned(16))) _Float16 la[128 * kH16Ld];
  __shared__ __attribute__((aligned(16))) _Float16 lb[64 * kH16Ld];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, q = lane >> 4;
  const int row0 = blockIdx.x * 128, col0 = blockIdx.y * 64;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 32;
  // joint: the members' columns tile together (one A); else member blockIdx.z
  const int ntot = joint ? ncol * cnt : ncol;
  const int mb = joint ? 0 : (int)blockIdx.z;
  const long long kg = kp / 8;
  // this thread's 16-byte pieces of the A tile (4: rows (tid + 256 i) / 8) and
  // of the B tile (2: columns (tid + 256 i) / 8), part (tid % 8) of the row's
  // 64 halves of the chunk
  const int part = tid & 7;
  const _Float16* ga[4];
  const _Float16* gb[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = min(row0 + (tid + 256 * i) / 8, m - 1);
    ga[i] = A + sA * mb + (long long)r * kg * 16 + part * 8;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int vc = min(col0 + (tid + 256 * i) / 8, ntot - 1);
    const int bb = joint ? vc / ncol : mb, c = joint ? vc % ncol : vc;
    gb[i] = B + sB * bb + ldb * c + part * 8;
  }
  half8_t pa[4], pb[2];
  auto load = [&](int kc) {
    const long long o = (long long)kc * 64;
#pragma unroll
    for (int i = 0; i < 4; ++i) pa[i] = *(const half8_t*)(ga[i] + o);
#pragma unroll
    for (int i = 0; i < 2; ++i) pb[i] = *(const half8_t*)(gb[i] + o);
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) *(half8_t*)(la + ((tid + 256 * i) / 8) * kH16Ld + part * 8) = pa[i];
#pragma unroll
    for (int i = 0; i < 2; ++i) *(half8_t*)(lb + ((tid + 256 * i) / 8) * kH16Ld + part * 8) = pb[i];
  };
  float4_t cre[4][2], cim[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) { cre[t][u][v] = 0.0f; cim[t][u][v] = 0.0f; }
  const int nk = kp / 32;
  load(0);
  stash();
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    if (kc + 1 < nk) load(kc + 1);
    half8_t ar[4], ai[4], br[2], bi[2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const _Float16* x = la + (wr + 16 * t + li) * kH16Ld + q * 16;
      ar[t] = *(const half8_t*)x;
      ai[t] = *(const half8_t*)(x + 8);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const _Float16* y = lb + (wc + 16 * u + li) * kH16Ld + q * 16;
      br[u] = *(const half8_t*)y;
      bi[u] = *(const half8_t*)(y + 8);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const half8_t bn = -bi[u];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        cre[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar[t], br[u], cre[t][u], 0, 0, 0);
        cre[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ai[t], bn, cre[t][u], 0, 0, 0);
        cim[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar[t], bi[u], cim[t][u], 0, 0, 0);
        cim[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ai[t], br[u], cim[t][u], 0, 0, 0);
      }
    }
    __syncthreads();
    if (kc + 1 < nk) {
      stash();
      __syncthreads();
    }
  }
  // lane holds rows 4q..4q+3 of each 16 x 16 tile at its column li
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int vc = col0 + wc + 16 * u + li;
    if (vc >= ntot) continue;
    const int bb = joint ? vc / ncol : mb, c = joint ? vc % ncol : vc;
    float* cc = C + 2 * (sC * bb + ldc * c);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + wr + 16 * t + 4 * q + v;
        if (r < m) {
          cc[2 * (long long)r] = cre[t][u][v];
          cc[2 * (long long)r + 1] = cim[t][u][v];
        }
      }
  }
}

hipError_t cgemm_h16(hipStream_t s, int m, int ncol, int kp, const _Float16* A, long long sA, const _Float16* B,
                     long long ldb, long long sB, float* C, long long ldc, long long sC, int cnt) {
  if (m <= 0 || ncol <= 0 || cnt <= 0) return hipSuccess;
  if (kp <= 0 || kp % 32) return hipErrorInvalidValue;
  const int joint = sA == 0 ? 1 : 0;
  const int cols = joint ? ncol * cnt : ncol;
  hipLaunchKernelGGL(k_cgemm_h16, dim3((m + 127) / 128, (cols + 63) / 64, joint ? 1 : cnt), dim3(256), 0, s, m, ncol,
                     kp, A, sA, B, ldb, sB, C, ldc, sC, cnt, joint);
  return hipGetLastError();
}

// thread: row r of member blockIdx.z, k group blockIdx.y (8 k)
__global__ void k_pack_a16(const float* __restrict__ src, int lda, long long ssrc, int m, int k, int kp,
                           _Float16* __restrict__ dst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (r >= m) return;
  const float* a = src + 2 * ssrc * blockIdx.z;
  half8_t re, im;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int kk = 8 * g + e;
    float xr = 0.0f, xi = 0.0f;
    if (kk < k) {
      const long long o = 2 * ((long long)r + (long long)kk * lda);
      xr = a[o];
      xi = a[o + 1];
    }
    re[e] = (_Float16)xr;
    im[e] = (_Float16)xi;
  }
  _Float16* d = dst + (long long)blockIdx.z * m * kp * 2 + ((long long)r * (kp / 8) + g) * 16;
  *(half8_t*)d = re;
  *(half8_t*)(d + 8) = im;
}

hipError_t pack_a16(hipStream_t s, const float* src, int lda, long long ssrc, int m, int k, _Float16* dst, int cnt) {
  if (m <= 0 || k <= 0 || cnt <= 0) return hipSuccess;
  const int kp = h16_kp(k);
  hipLaunchKernelGGL(k_pack_a16, dim3((m + 127) / 128, kp / 8, cnt), dim3(128), 0, s, src, lda, ssrc, m, k, kp, dst);
  return hipGetLastError();
}

// thread: group gq (8 rows) of the members' padded rows, column blockIdx.y
template <typename T>
__global__ void k_pack_b16(const T* __restrict__ V, long long ldv, const int* __restrict__ perm, int r0, int nb,
                           int kp, int cnt, _Float16* __restrict__ W, long long ldw) {
  const int gq = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y, kg = kp / 8;
  if (gq >= kg * cnt) return;
  const int b = gq / kg, g = gq % kg;
  half8_t re, im;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int i = 8 * g + e;
    T xr = 0, xi = 0;
    if (i < nb) {
      const int row = r0 + b * nb + i;
      const long long o = 2 * ((long long)(perm ? perm[row] : row) + (long long)c * ldv);
      xr = V[o];
      xi = V[o + 1];
    }
    re[e] = (_Float16)xr;
    im[e] = (_Float16)xi;
  }
  _Float16* d = W + ldw * c + (long long)gq * 16;
  *(half8_t*)d = re;
  *(half8_t*)(d + 8) = im;
}

template <typename T>
static hipError_t pack_b16_t(hipStream_t s, const T* V, long long ldv, const int* perm, int r0, int nb, int cnt,
                             int ncol, _Float16* W, long long ldw) {
  if (nb <= 0 || cnt <= 0 || ncol <= 0) return hipSuccess;
  const int kp = h16_kp(nb), groups = kp / 8 * cnt;
  hipLaunchKernelGGL(k_pack_b16<T>, dim3((groups + 255) / 256, ncol), dim3(256), 0, s, V, ldv, perm, r0, nb, kp, cnt,
                     W, ldw);
  return hipGetLastError();
}

hipError_t pack_b16(hipStream_t s, const double* V, long long ldv, const int* perm, int r0, int nb, int cnt, int ncol,
                    _Float16* W, long long ldw) {
  return pack_b16_t(s, V, ldv, perm, r0, nb, cnt, ncol, W, ldw);
}
hipError_t pack_b16(hipStream_t s, const float* V, long long ldv, const int* perm, int r0, int nb, int cnt, int ncol,
                    _Float16* W, long long ldw) {
  return pack_b16_t(s, V, ldv, perm, r0, nb, cnt, ncol, W, ldw);
}

}  // namespace hm
