// hm_gmres.hip -- restarted GMRES for the dense BEM system on gfx950.
//
// Replaces ZITER_SOLVE_MULTI -> ZBLOCK_GMRES_SOLVE -> ZGMRES_SOLVE
// (IterativeSolvers.f90:574, :535, :324) with the block-diagonal preconditioner
// of SETUP_BLOCK_PRECONDITIONER / APPLY_PRECONDITIONER (:190-307).
//
// The reference solves the right-hand sides one after another. Here all columns
// advance in lockstep, so the Krylov step of every column is one complex GEMM:
// W = A32 * Z, with the FP32 copy of the matrix on MFMA. Columns that have
// converged are frozen (their basis vectors are zero from then on). Each
// restart cycle solves for a correction of the current FP64 residual
// R = B - A*X, which is computed with the FP64 matrix. This is iterative
// refinement with GMRES as the inner solver: the FP32 products set the inner
// accuracy, and the FP64 residual decides convergence (||R_c|| <= tol*||B_c||
// for every column c, the reference's criterion at :388-391).
//
// Orthogonalisation is classical Gram-Schmidt applied twice (the reference uses
// one modified Gram-Schmidt pass). The complex Givens rotation is the unitary
// one, c = conj(a)/d, s = conj(b)/d. The reference's ZGIVENS_ROTATION
// (:490-506) takes c = a/d, s = b/d, which annihilates the subdiagonal only for
// real a, so its residual estimate is inexact. The true residual is the same
// criterion in both.
//
// Right preconditioning as in the reference: the Krylov vectors are
// v -> M^{-1} v before the product, and the correction is M^{-1} (V y). M is
// block diagonal with one block per body, LU-factored in FP32; M^{-1} is applied
// by the blocks' triangular solves, or, when the explicit FP32 inverses of the
// blocks are given, by one batched MFMA GEMM after the pivot permutation (the
// triangular solves of a hundred columns are latency bound: 64 small GEMMs and
// 64 block solves per application for a block of 2048).
#include <hip/hip_runtime.h>

#include <vector>

#include "hm_gmres.hpp"
#include "hm_linalg.hpp"

namespace hm {

namespace {

struct cplx { double re, im; };
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx conjmul(cplx a, cplx b) {  // conj(a) * b
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cplx ld(const double* p) { return {p[0], p[1]}; }
__device__ __forceinline__ void st(double* p, cplx a) { p[0] = a.re; p[1] = a.im; }

// block (256 threads) sum of a double
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// out[c] = ||X[:, c]||_2 for an n x ncol block with leading dimension ldx
__global__ __launch_bounds__(256) void k_colnorm(const double* __restrict__ X, long long ldx, int n, double* out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  const double* x = X + 2 * ldx * c;
  double acc = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) acc += x[2 * r] * x[2 * r] + x[2 * r + 1] * x[2 * r + 1];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[c] = sqrt(acc);
}

// start of a cycle: V0 = R / ||R|| for the columns still above tolerance
__global__ __launch_bounds__(256) void k_start(const double* __restrict__ R, long long ldr, int n, int ncol,
                                               const double* rnorm, const double* bnorm, double tol, int m1,
                                               double* __restrict__ V0, double* g, int* active, int* steps) {
  const int c = blockIdx.y;
  const double beta = rnorm[c];
  const bool act = beta > tol * bnorm[c] && bnorm[c] > 1e-30;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < n) {
    const double s = act ? 1.0 / beta : 0.0;
    V0[2 * ((long long)c * n + r)] = act ? R[2 * (ldr * c + r)] * s : 0.0;
    V0[2 * ((long long)c * n + r) + 1] = act ? R[2 * (ldr * c + r) + 1] * s : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* gc = g + 2 * (long long)c * m1;
    for (int k = 0; k < m1; ++k) { gc[2 * k] = 0.0; gc[2 * k + 1] = 0.0; }
    gc[0] = act ? beta : 0.0;
    active[c] = act ? 1 : 0;
    steps[c] = 0;
  }
}

// hc[c][i] = <V_i[:, c], w[:, c]> for i = 0..j (conjugated first argument)
__global__ __launch_bounds__(256) void k_dots(const double* __restrict__ V, long long vstride, const double* __restrict__ w,
                                              int n, const int* active, int m1, double* hc) {
  __shared__ double sr[256], si[256];
  const int i = blockIdx.x, c = blockIdx.y;
  if (!active[c]) {
    if (threadIdx.x == 0) { hc[2 * ((long long)c * m1 + i)] = 0.0; hc[2 * ((long long)c * m1 + i) + 1] = 0.0; }
    return;
  }
  const double* v = V + 2 * (vstride * i + (long long)c * n);
  const double* x = w + 2 * (long long)c * n;
  double ar = 0.0, ai = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    const cplx p = conjmul(ld(v + 2 * r), ld(x + 2 * r));
    ar += p.re;
    ai += p.im;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { hc[2 * ((long long)c * m1 + i)] = sr[0]; hc[2 * ((long long)c * m1 + i) + 1] = si[0]; }
}

// w[:, c] -= sum_{i<=j} hc[c][i] V_i[:, c]
__global__ __launch_bounds__(256) void k_axpy(double* __restrict__ w, const double* __restrict__ V, long long vstride,
                                              int j, int n, const int* active, int m1, const double* hc) {
  __shared__ double hr[kGmresMaxRestart + 1], hi[kGmresMaxRestart + 1];
  const int c = blockIdx.y;
  if (!active[c]) return;
  for (int i = threadIdx.x; i <= j; i += 256) {
    hr[i] = hc[2 * ((long long)c * m1 + i)];
    hi[i] = hc[2 * ((long long)c * m1 + i) + 1];
  }
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double* x = w + 2 * ((long long)c * n + r);
  double xr = x[0], xi = x[1];
  for (int i = 0; i <= j; ++i) {
    const double* v = V + 2 * (vstride * i + (long long)c * n + r);
    xr -= hr[i] * v[0] - hi[i] * v[1];
    xi -= hr[i] * v[1] + hi[i] * v[0];
  }
  x[0] = xr;
  x[1] = xi;
}

// V_{j+1} = w / ||w|| (zero for frozen columns or an exact breakdown)
__global__ __launch_bounds__(256) void k_normalize(const double* __restrict__ w, const double* wnorm, int n,
                                                   const int* active, double* __restrict__ Vn) {
  const int c = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const double nr = wnorm[c];
  const bool ok = active[c] && nr > 1e-30;
  const double s = ok ? 1.0 / nr : 0.0;
  const long long o = 2 * ((long long)c * n + r);
  Vn[o] = ok ? w[o] * s : 0.0;
  Vn[o + 1] = ok ? w[o + 1] * s : 0.0;
}

// Hessenberg column j of every active column: H(0:j, j) = hc1 + hc2,
// H(j+1, j) = ||w||; previous rotations, a new unitary rotation, the
// least-squares right-hand side g and the residual estimate |g(j+1)|.
// One workgroup; writes the number of columns still active to *nactive.
__global__ __launch_bounds__(256) void k_givens(int j, int ncol, int m1, double* H, double* cs, double* sn, double* g,
                                                const double* hc1, const double* hc2, const double* wnorm,
                                                const double* bnorm, double tol, int* active, int* steps,
                                                double* est, int* nactive) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int m = m1 - 1;
  int mine = 0;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    if (!active[c]) continue;
    double* hcol = H + 2 * ((long long)c * m1 * m + (long long)j * m1);
    double* csc = cs + 2 * (long long)c * m;
    double* snc = sn + 2 * (long long)c * m;
    double* gc = g + 2 * (long long)c * m1;
    for (int i = 0; i <= j; ++i) {
      const cplx a = ld(hc1 + 2 * ((long long)c * m1 + i)), b = ld(hc2 + 2 * ((long long)c * m1 + i));
      st(hcol + 2 * i, {a.re + b.re, a.im + b.im});
    }
    st(hcol + 2 * (j + 1), {wnorm[c], 0.0});
    for (int i = 0; i < j; ++i) {
      const cplx ci = ld(csc + 2 * i), si = ld(snc + 2 * i), hi = ld(hcol + 2 * i), hn = ld(hcol + 2 * (i + 1));
      const cplx t0 = cmul(ci, hi), t1 = cmul(si, hn);
      const cplx u0 = conjmul(si, hi), u1 = conjmul(ci, hn);
      st(hcol + 2 * i, {t0.re + t1.re, t0.im + t1.im});
      st(hcol + 2 * (i + 1), {u1.re - u0.re, u1.im - u0.im});
    }
    const cplx a = ld(hcol + 2 * j), b = ld(hcol + 2 * (j + 1));
    cplx cj, sj;
    if (sqrt(cabs2(b)) < 1e-30) { cj = {1.0, 0.0}; sj = {0.0, 0.0}; }
    else if (sqrt(cabs2(a)) < 1e-30) { cj = {0.0, 0.0}; sj = {1.0, 0.0}; }
    else {
      const double d = sqrt(cabs2(a) + cabs2(b));
      cj = {a.re / d, -a.im / d};
      sj = {b.re / d, -b.im / d};
    }
    st(csc + 2 * j, cj);
    st(snc + 2 * j, sj);
    {
      const cplx t0 = cmul(cj, a), t1 = cmul(sj, b);
      st(hcol + 2 * j, {t0.re + t1.re, t0.im + t1.im});
      st(hcol + 2 * (j + 1), {0.0, 0.0});
    }
    const cplx gj = ld(gc + 2 * j), gn = ld(gc + 2 * (j + 1));
    const cplx t0 = cmul(cj, gj), t1 = cmul(sj, gn);
    const cplx u0 = conjmul(sj, gj), u1 = conjmul(cj, gn);
    st(gc + 2 * j, {t0.re + t1.re, t0.im + t1.im});
    st(gc + 2 * (j + 1), {u1.re - u0.re, u1.im - u0.im});
    steps[c] = j + 1;
    const double e = sqrt((u1.re - u0.re) * (u1.re - u0.re) + (u1.im - u0.im) * (u1.im - u0.im));
    est[c] = e;
    if (e < tol * bnorm[c] || !(wnorm[c] > 1e-30)) active[c] = 0;
    else ++mine;
  }
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) *nactive = cnt;
}

// y = H(0:k,0:k)^{-1} g(0:k), back substitution as ZGMRES_SOLVE :447-458
__global__ void k_backsolve(int ncol, int m1, const double* H, const double* g, const int* steps, double* y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int m = m1 - 1, k = steps[c];
  const double* Hc = H + 2 * (long long)c * m1 * m;
  double* yc = y + 2 * (long long)c * m;
  for (int i = 0; i < k; ++i) st(yc + 2 * i, ld(g + 2 * ((long long)c * m1 + i)));
  for (int i = k - 1; i >= 0; --i) {
    const cplx d = ld(Hc + 2 * (i + (long long)i * m1));
    const double dd = cabs2(d);
    cplx yi = ld(yc + 2 * i);
    if (sqrt(dd) < 1e-30) yi = {0.0, 0.0};
    else yi = {(yi.re * d.re + yi.im * d.im) / dd, (yi.im * d.re - yi.re * d.im) / dd};
    st(yc + 2 * i, yi);
    for (int kk = 0; kk < i; ++kk) {
      const cplx p = cmul(ld(Hc + 2 * (kk + (long long)i * m1)), yi);
      const cplx o = ld(yc + 2 * kk);
      st(yc + 2 * kk, {o.re - p.re, o.im - p.im});
    }
  }
}

// U[:, c] = sum_{i < steps[c]} y[c][i] V_i[:, c]   (or X += ... when add)
__global__ __launch_bounds__(256) void k_combine(const double* __restrict__ V, long long vstride, int n, int m,
                                                 const double* y, const int* steps, double* __restrict__ U,
                                                 long long ldu, int add) {
  __shared__ double yr[kGmresMaxRestart], yi[kGmresMaxRestart];
  const int c = blockIdx.y, k = steps[c];
  for (int i = threadIdx.x; i < k; i += 256) {
    yr[i] = y[2 * ((long long)c * m + i)];
    yi[i] = y[2 * ((long long)c * m + i) + 1];
  }
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double ur = 0.0, ui = 0.0;
  for (int i = 0; i < k; ++i) {
    const double* v = V + 2 * (vstride * i + (long long)c * n + r);
    ur += yr[i] * v[0] - yi[i] * v[1];
    ui += yr[i] * v[1] + yi[i] * v[0];
  }
  double* u = U + 2 * (ldu * c + r);
  if (add) { u[0] += ur; u[1] += ui; }
  else { u[0] = ur; u[1] = ui; }
}

}  // namespace

size_t gmres_workspace_bytes(int n, int ncol, int restart, int nblk16) {
  const size_t m = (size_t)restart, m1 = m + 1, cx = 2 * sizeof(double);
  size_t b = 0;
  b += cx * m1 * ncol * n;          // V
  b += cx * ncol * m1 * m;          // H
  b += 2 * cx * ncol * m;           // cs, sn
  b += cx * ncol * m1;              // g
  b += cx * ncol * m;               // y
  b += 2 * cx * ncol * m1;          // hc1, hc2
  b += 4 * sizeof(double) * ncol;   // bnorm, rnorm, wnorm, est
  b += 3 * sizeof(int) * ncol + 64; // active, steps, nactive
  b += cx * (size_t)n * ncol * 2;   // R, W64
  b += 2 * 2 * sizeof(float) * (size_t)n * ncol;  // Z32, W32
  b += 2 * sizeof(_Float16) * (size_t)(n + 31 * nblk16) * ncol + 256;   // W16 (rows padded to 32 per block)
  return b + 64 * 16;               // alignment slack
}

// runs of consecutive blocks of equal size, handled as one batched call each
template <typename F>
static hipError_t for_block_runs(int nblk, const int* blk_start, F&& f) {
  size_t off = 0;
  for (int b = 0; b < nblk;) {
    const int r0 = blk_start[b], nb = blk_start[b + 1] - r0;
    int cnt = 1;
    while (b + cnt < nblk && blk_start[b + cnt + 1] - blk_start[b + cnt] == nb) ++cnt;
    hipError_t e = f(r0, nb, cnt, off);
    if (e != hipSuccess) return e;
    off += (size_t)nb * nb * cnt;
    b += cnt;
  }
  return hipSuccess;
}

// R = B - W (column c of B at B + c ldb, of W and R at + c n)
__global__ void k_sub(const double* __restrict__ B, long long ldb, const double* __restrict__ W, int n,
                      double* __restrict__ R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i >= n) return;
  const long long o = (long long)c * n + i, ob = (long long)c * ldb + i;
  R[2 * o] = B[2 * ob] - W[2 * o];
  R[2 * o + 1] = B[2 * ob + 1] - W[2 * o + 1];
}

hipError_t gmres_block_precond_setup(hipStream_t s, const double* A, int lda, int nblk, const int* blk_start,
                                     float* P, int* pipiv, int* info, unsigned* counter) {
  return for_block_runs(nblk, blk_start, [&](int r0, int nb, int cnt, size_t off) {
    hipError_t e = demote(s, A + 2 * ((long long)r0 + (long long)r0 * lda), lda, (long long)nb * (lda + 1),
                          P + 2 * off, nb, (long long)nb * nb, nb, nb, cnt);
    if (e != hipSuccess) return e;
    return lu_factor<float>(s, P + 2 * off, nb, (long long)nb * nb, nb, pipiv + r0, info, counter, cnt);
  });
}

// batch of nb x nb identity matrices, one after the other
__global__ void k_identity(float* __restrict__ I, int nb) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x, nn = (long long)nb * nb;
  if (e >= nn) return;
  float* b = I + 2 * nn * blockIdx.y;
  const int r = (int)(e % nb), c = (int)(e / nb);
  b[2 * e] = r == c ? 1.f : 0.f;
  b[2 * e + 1] = 0.f;
}

hipError_t gmres_block_inverse(hipStream_t s, int nblk, const int* blk_start, const float* P, float* Pinv) {
  return for_block_runs(nblk, blk_start, [&](int r0, int nb, int cnt, size_t off) {
    const long long nn = (long long)nb * nb;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((nn + 255) / 256), cnt), dim3(256), 0, s, Pinv + 2 * off, nb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return lu_inverse_nopiv<float>(s, P + 2 * off, nb, nn, nb, Pinv + 2 * off, nb, nn, cnt);
  });
}

// row permutation of each block from its LAPACK-style pivot sequence: perm =
// the sequential swaps j <-> ipiv[j] applied to the identity, stored as
// absolute rows (row r of the permuted vector is row perm[r] of the vector)
__global__ void k_pivot_perm(const int* __restrict__ ipiv, int* __restrict__ perm, int nb, int cnt, int r0) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= cnt) return;
  const int* pv = ipiv + (long long)nb * z;
  int* pm = perm + (long long)nb * z;
  for (int j = 0; j < nb; ++j) pm[j] = j;
  for (int j = 0; j < nb; ++j) {
    const int q = pv[j];
    if (q != j) {
      const int a = pm[j];
      pm[j] = pm[q];
      pm[q] = a;
    }
  }
  const int base = r0 + nb * z;
  for (int j = 0; j < nb; ++j) pm[j] += base;
}

hipError_t gmres_block_perm(hipStream_t s, int nblk, const int* blk_start, const int* pipiv, int* perm) {
  return for_block_runs(nblk, blk_start, [&](int r0, int nb, int cnt, size_t) {
    hipLaunchKernelGGL(k_pivot_perm, dim3((cnt + 63) / 64), dim3(64), 0, s, pipiv + r0, perm + r0, nb, cnt, r0);
    return hipGetLastError();
  });
}

// W = FP32 copy of the rows perm of V (n x ncol, leading dimensions ldv and n)
__global__ void k_demote_perm(const double* __restrict__ V, long long ldv, const int* __restrict__ perm, int n,
                              float* __restrict__ W) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= n) return;
  const long long o = 2 * ((long long)perm[r] + (long long)c * ldv), d = 2 * ((long long)r + (long long)c * n);
  W[d] = (float)V[o];
  W[d + 1] = (float)V[o + 1];
}

// Z = M^{-1} Z in place with the blocks' LU factors
static hipError_t apply_precond(hipStream_t s, const GmresArgs& a, float* Z, int ncol) {
  return for_block_runs(a.nblk, a.blk_start, [&](int r0, int nb, int cnt, size_t off) {
    return lu_solve<float>(s, a.P + 2 * off, nb, (long long)nb * nb, nb, a.pipiv + r0, Z + 2 * (long long)r0, a.n,
                           nb, ncol, cnt);
  });
}

// Z = M^{-1} v = (U^{-1} L^{-1}) (P v): the pivot permutation applied while v is
// demoted to FP32 (into W), or to FP16 (into W16) with the FP16 inverses (half:
// Arnoldi vectors, of unit norm, so that no entry that matters falls below
// FP16's normal range), then one batched GEMM per run of equal blocks
static hipError_t apply_inverse(hipStream_t s, const GmresArgs& a, const double* v, float* W, _Float16* W16, float* Z,
                                int ncol, bool half) {
  if (half && a.Pinv16) {
    // blocks of one size (gmres_solve checks): packed FP16 rows, kp per block
    const int r0 = a.blk_start[0], nb = a.blk_start[1] - r0, kp = h16_kp(nb);
    const long long ldw = 2LL * kp * a.nblk;
    hipError_t e = pack_b16(s, v, (long long)a.n, a.pperm, r0, nb, a.nblk, ncol, W16, ldw);
    if (e != hipSuccess) return e;
    return cgemm_h16(s, nb, ncol, kp, a.Pinv16, a.pinv16_shared ? 0 : 2LL * nb * kp, W16, ldw, 2LL * kp,
                     Z + 2 * (long long)r0, a.n, nb, a.nblk);
  }
  hipLaunchKernelGGL(k_demote_perm, dim3((a.n + 255) / 256, ncol), dim3(256), 0, s, v, (long long)a.n, a.pperm, a.n, W);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.pinv_shared) {
    // one inverse for all blocks (gmres_solve checks the layout): W and Z
    // (n x ncol, ld n = nblk nb) read as nb x (nblk ncol) matrices of leading
    // dimension nb put block b's rows of column c in column b + c nblk
    const int nb = a.blk_start[1];
    return cgemm_3m<float>(s, false, nb, a.nblk * ncol, nb, 1.f, 0.f, a.Pinv, nb, 0, W, nb, 0, 0.f, 0.f, Z, nb, 0, 1);
  }
  return for_block_runs(a.nblk, a.blk_start, [&](int r0, int nb, int cnt, size_t off) {
    return cgemm_3m<float>(s, false, nb, ncol, nb, 1.f, 0.f, a.Pinv + 2 * off, nb, (long long)nb * nb,
                        W + 2 * (long long)r0, a.n, nb, 0.f, 0.f, Z + 2 * (long long)r0, a.n, nb, cnt);
  });
}

hipError_t gmres_solve(hipStream_t s, const GmresArgs& a, void* work, GmresStats* stats) {
  const int n = a.n, ncol = a.ncol, m = a.restart, m1 = m + 1;
  const size_t cx = 2 * sizeof(double);
  char* p = static_cast<char*>(work);
  auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~(size_t)255; return q; };
  double* V = (double*)take(cx * (size_t)m1 * ncol * n);
  double* H = (double*)take(cx * (size_t)ncol * m1 * m);
  double* cs = (double*)take(cx * (size_t)ncol * m);
  double* sn = (double*)take(cx * (size_t)ncol * m);
  double* g = (double*)take(cx * (size_t)ncol * m1);
  double* y = (double*)take(cx * (size_t)ncol * m);
  double* hc1 = (double*)take(cx * (size_t)ncol * m1);
  double* hc2 = (double*)take(cx * (size_t)ncol * m1);
  double* bnorm = (double*)take(sizeof(double) * ncol);
  double* rnorm = (double*)take(sizeof(double) * ncol);
  double* wnorm = (double*)take(sizeof(double) * ncol);
  double* est = (double*)take(sizeof(double) * ncol);
  int* active = (int*)take(sizeof(int) * ncol);
  int* steps = (int*)take(sizeof(int) * ncol);
  int* nactive = (int*)take(sizeof(int) * 16);
  double* R = (double*)take(cx * (size_t)n * ncol);
  double* W = (double*)take(cx * (size_t)n * ncol);
  float* Z32 = (float*)take(2 * sizeof(float) * (size_t)n * ncol);
  float* W32 = (float*)take(2 * sizeof(float) * (size_t)n * ncol);
  _Float16* W16 = (_Float16*)take(2 * sizeof(_Float16) * (size_t)(n + 31 * (a.Pinv16 ? a.nblk : 0)) * ncol);
  if (a.Pinv16 || a.pinv_shared) {
    for (int b = 1; b < a.nblk; ++b)
      if (a.blk_start[b + 1] - a.blk_start[b] != a.blk_start[1] - a.blk_start[0]) return hipErrorInvalidValue;
  }
  if (a.pinv_shared && (a.blk_start[0] != 0 || a.blk_start[a.nblk] != n)) return hipErrorInvalidValue;
  const long long vstride = (long long)ncol * n;
  const dim3 gcol((n + 255) / 256, ncol);
  hipError_t e;
#define GM_CHECK(x) do { e = (x); if (e != hipSuccess) return e; } while (0)

  std::vector<double> hb(ncol), hr(ncol);
  GM_CHECK(hipMemsetAsync(a.X, 0, cx * (size_t)n * ncol, s));
  hipLaunchKernelGGL(k_colnorm, dim3(ncol), dim3(256), 0, s, a.B, (long long)a.ldb, n, bnorm);
  GM_CHECK(hipGetLastError());
  GM_CHECK(hipMemcpyAsync(hb.data(), bnorm, sizeof(double) * ncol, hipMemcpyDeviceToHost, s));
  GM_CHECK(hipMemcpy2DAsync(R, cx * n, a.B, cx * a.ldb, cx * n, ncol, hipMemcpyDeviceToDevice, s));
  GM_CHECK(hipMemcpyAsync(rnorm, bnorm, sizeof(double) * ncol, hipMemcpyDeviceToDevice, s));
  GM_CHECK(hipStreamSynchronize(s));
  int total = 0, cycles = 0;
  bool done = false;
  double worst = 0.0;
  for (;;) {
    // convergence on the FP64 residual
    GM_CHECK(hipMemcpyAsync(hr.data(), rnorm, sizeof(double) * ncol, hipMemcpyDeviceToHost, s));
    GM_CHECK(hipStreamSynchronize(s));
    worst = 0.0;
    for (int c = 0; c < ncol; ++c)
      if (hb[c] > 1e-30) worst = fmax(worst, hr[c] / hb[c]);
    if (worst <= a.tol) { done = true; break; }
    if (total >= a.maxiter || cycles >= a.max_cycles) break;
    ++cycles;
    // this cycle's Arnoldi target: inner_tol, or a reduction by cycle_rtol of the
    // largest relative residual when the inner operator cannot get further
    const double target = fmax(a.inner_tol, a.cycle_rtol * worst);
    hipLaunchKernelGGL(k_start, gcol, dim3(256), 0, s, R, (long long)n, n, ncol, rnorm, bnorm, a.inner_tol, m1, V, g,
                       active, steps);
    GM_CHECK(hipGetLastError());
    for (int j = 0; j < m && total < a.maxiter; ++j) {
      ++total;
      const double* vj = V + 2 * vstride * j;
      // W = A * M^{-1} v_j on MFMA in FP32
      if (a.nblk > 0 && a.Pinv) {
        GM_CHECK(apply_inverse(s, a, vj, W32, W16, Z32, ncol, true));
      } else {
        GM_CHECK(demote(s, vj, n, 0, Z32, n, 0, n, ncol, 1));
        if (a.nblk > 0) GM_CHECK(apply_precond(s, a, Z32, ncol));
      }
      if (a.nblk > 0 && a.op32_off) {
        // the diagonal blocks times their own inverses: v itself
        GM_CHECK(demote(s, vj, n, 0, W32, n, 0, n, ncol, 1));
        GM_CHECK(a.op32_off(s, Z32, W32));
      } else if (a.op32) {
        GM_CHECK(a.op32(s, Z32, W32));
      } else {
        GM_CHECK(cgemm<float>(s, false, n, ncol, n, 1.f, 0.f, a.A32, n, 0, Z32, n, 0, 0.f, 0.f, W32, n, 0, 1));
      }
      GM_CHECK(promote_add(s, W32, n, 0, W, n, 0, n, ncol, 1, true));
      // classical Gram-Schmidt, twice
      hipLaunchKernelGGL(k_dots, dim3(j + 1, ncol), dim3(256), 0, s, V, vstride, W, n, active, m1, hc1);
      hipLaunchKernelGGL(k_axpy, gcol, dim3(256), 0, s, W, V, vstride, j, n, active, m1, hc1);
      hipLaunchKernelGGL(k_dots, dim3(j + 1, ncol), dim3(256), 0, s, V, vstride, W, n, active, m1, hc2);
      hipLaunchKernelGGL(k_axpy, gcol, dim3(256), 0, s, W, V, vstride, j, n, active, m1, hc2);
      hipLaunchKernelGGL(k_colnorm, dim3(ncol), dim3(256), 0, s, W, (long long)n, n, wnorm);
      hipLaunchKernelGGL(k_normalize, gcol, dim3(256), 0, s, W, wnorm, n, active, V + 2 * vstride * (j + 1));
      hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, s, j, ncol, m1, H, cs, sn, g, hc1, hc2, wnorm, bnorm,
                         target, active, steps, est, nactive);
      GM_CHECK(hipGetLastError());
      int na = 0;
      GM_CHECK(hipMemcpyAsync(&na, nactive, sizeof(int), hipMemcpyDeviceToHost, s));
      GM_CHECK(hipStreamSynchronize(s));
      if (na == 0) break;
    }
    // X += M^{-1} (V y)
    hipLaunchKernelGGL(k_backsolve, dim3((ncol + 63) / 64), dim3(64), 0, s, ncol, m1, H, g, steps, y);
    if (a.nblk > 0) {
      hipLaunchKernelGGL(k_combine, gcol, dim3(256), 0, s, V, vstride, n, m, y, steps, W, (long long)n, 0);
      GM_CHECK(hipGetLastError());
      if (a.Pinv) {
        // the update itself in FP32: its scale follows the residual's
        GM_CHECK(apply_inverse(s, a, W, W32, W16, Z32, ncol, false));
      } else {
        GM_CHECK(demote(s, W, n, 0, Z32, n, 0, n, ncol, 1));
        GM_CHECK(apply_precond(s, a, Z32, ncol));
      }
      GM_CHECK(promote_add(s, Z32, n, 0, a.X, n, 0, n, ncol, 1, false));
    } else {
      hipLaunchKernelGGL(k_combine, gcol, dim3(256), 0, s, V, vstride, n, m, y, steps, a.X, (long long)n, 1);
      GM_CHECK(hipGetLastError());
    }
    // FP64 residual R = B - A X
    if (a.op64) {
      GM_CHECK(a.op64(s, a.X, W));
      hipLaunchKernelGGL(k_sub, gcol, dim3(256), 0, s, a.B, (long long)a.ldb, W, n, R);
    } else {
      GM_CHECK(hipMemcpy2DAsync(R, cx * n, a.B, cx * a.ldb, cx * n, ncol, hipMemcpyDeviceToDevice, s));
      GM_CHECK(cgemm<double>(s, false, n, ncol, n, -1.0, 0.0, a.A64, a.lda64, 0, a.X, n, 0, 1.0, 0.0, R, n, 0, 1));
    }
    hipLaunchKernelGGL(k_colnorm, dim3(ncol), dim3(256), 0, s, R, (long long)n, n, rnorm);
    GM_CHECK(hipGetLastError());
  }
#undef GM_CHECK
  if (stats) {
    stats->iterations = total;
    stats->cycles = cycles;
    stats->converged = done ? 1 : 0;
    stats->rel_residual = worst;
  }
  return hipSuccess;
}

}  // namespace hm
