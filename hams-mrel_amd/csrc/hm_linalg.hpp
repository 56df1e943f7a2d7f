// hm_linalg.hpp -- host-side launchers of the dense complex kernels (hm_linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace hm {

constexpr int kLuNb = 64;   // LU panel width (also the triangular-solve block)

// C = alpha*op(A)*B + beta*C, complex interleaved, column-major; op(A) = A or A^T.
template <typename T>
hipError_t cgemm(hipStream_t s, bool transA, int m, int n, int k, T ar, T ai, const T* A, int lda, long long sA,
                 const T* B, int ldb, long long sB, T br, T bi, T* C, int ldc, long long sC, int batch);
// The same with each complex product formed from three real ones
// (Re = Ar Br - Ai Bi, Im = (Ar + Ai)(Br + Bi) - Ar Br - Ai Bi): a quarter
// fewer MFMAs; the imaginary part's error is bounded by
// (|Ar| + |Ai|)(|Br| + |Bi|) instead of by its own terms.
template <typename T>
hipError_t cgemm_3m(hipStream_t s, bool transA, int m, int n, int k, T ar, T ai, const T* A, int lda, long long sA,
                    const T* B, int ldb, long long sB, T br, T bi, T* C, int ldc, long long sC, int batch);

// In-place partial-pivot LU of `batch` n x n matrices (0-based ipiv, n per matrix;
// info[b] = first zero-pivot column + 1 or 0). counter: `batch` zeroed uints.
template <typename T>
hipError_t lu_factor(hipStream_t s, T* A, int lda, long long sA, int n, int* ipiv, int* info, unsigned* counter,
                     int batch, void* coop_ws = nullptr);
// workspace of lu_factor's cooperative panel kernel (one matrix of order <= n):
// with it, the panels of a single matrix are factored one cooperative launch
// each (the panel in LDS across workgroups, one grid barrier a column) where
// the device can hold the launch
size_t lu_coop_ws_bytes(int n);

// Solve A X = B with the factors of lu_factor; B is overwritten by X.
template <typename T>
hipError_t lu_solve(hipStream_t s, const T* A, int lda, long long sA, int n, const int* ipiv, T* B, int ldb,
                    long long sB, int nrhs, int batch);

// B = (L U)^{-1} B without the row interchanges, for B = the identity on entry
// (the caller sets it): the forward sweep touches only the columns that
// L^{-1} fills (column block k0 needs columns < k0 + 64), 2/3 of the work of
// lu_solve with n right-hand sides
template <typename T>
hipError_t lu_inverse_nopiv(hipStream_t s, const T* A, int lda, long long sA, int n, T* B, int ldb, long long sB,
                            int batch);

// In-place explicit inverse of `batch` n x n FP32 matrices (stride sA) by
// blocked Gauss-Jordan elimination without row interchanges (for matrices that
// need none: no pivot search, nearly all the work in n x n x 64 GEMMs on MFMA).
// ws: gj_inverse_ws_bytes(n) per matrix; info[b] = first column + 1 of a tile
// with a zero pivot, or left as it is (zeroed by the caller).
size_t gj_inverse_ws_bytes(int n);
hipError_t fill_iota(hipStream_t s, int* p, long long n);   // p[i] = i

// Complex GEMM with FP16 operands and FP32 accumulation on MFMA, over packed
// operands: the contraction index k (padded with zeros to kp, a multiple of
// 32) in groups of 8, each group 16 halves [re of k..k+7 | im of k..k+7].
// - A (m x kp): row r, group g at (r * kp / 8 + g) * 16 halves (pack_a16);
// - B (kp x ncol): column c, group g at c * ldb + g * 16 halves (pack_b16).
// Member b of cnt: C_b (m x ncol, FP32 interleaved complex, column-major, ldc
// in complex entries) at C + 2 b sC, A_b at A + b sA, B_b at B + b sB (halves).
// sA = 0: one A for every member (the members' columns then tile together).
hipError_t cgemm_h16(hipStream_t s, int m, int ncol, int kp, const _Float16* A, long long sA, const _Float16* B,
                     long long ldb, long long sB, float* C, long long ldc, long long sC, int cnt);
inline int h16_kp(int k) { return (k + 31) / 32 * 32; }
// A packed from FP32 interleaved complex column-major matrices (m x k, lda;
// cnt of them at stride ssrc complex entries) into dst (stride m * kp * 2 halves)
hipError_t pack_a16(hipStream_t s, const float* src, int lda, long long ssrc, int m, int k, _Float16* dst, int cnt);
// B packed from column-major complex vectors: member b's rows
// [r0 + b nb, r0 + (b + 1) nb) of each of ncol columns (row i of the member is
// row perm[r0 + b nb + i] of V when perm is given), into W at column stride ldw
// halves, member b at row offset b kp (W's rows kp * cnt, zero padded)
hipError_t pack_b16(hipStream_t s, const double* V, long long ldv, const int* perm, int r0, int nb, int cnt, int ncol,
                    _Float16* W, long long ldw);
hipError_t pack_b16(hipStream_t s, const float* V, long long ldv, const int* perm, int r0, int nb, int cnt, int ncol,
                    _Float16* W, long long ldw);
hipError_t gj_inverse_nopiv(hipStream_t s, float* A, int lda, long long sA, int n, void* ws, int* info, int batch);

// ---- grouped operations: members of different sizes, one launch per step ----
template <typename T>
struct GemmDesc {            // C = alpha op(A) B + beta C for one member
  int m, n, k;
  T ar, ai, br, bi;
  const T* A;
  long long lda;
  const T* B;
  long long ldb;
  T* C;
  long long ldc;
};
template <typename T>
struct LuDesc {              // one member of a grouped factorisation / solve
  T* A;                      // n x n, leading dimension lda, factored in place
  long long lda;
  int n;
  int* ipiv;                 // n (0-based)
  int* info;                 // one word: first zero-pivot column + 1, or 0 (zeroed by the caller)
  unsigned* counter;         // one word, zeroed by the caller
  T* B;                      // right-hand sides of lu_solve_grouped (n x nrhs, ld ldb), overwritten by X
  long long ldb;
  int nrhs;
};
struct ConvDesc {            // FP64 <-> FP32 copy of an m x n block
  double* f64;
  long long ld64;
  float* f32;
  long long ld32;
  int m, n;
};
struct CopyDesc {            // FP64 m x n block copy
  const double* src;
  long long lds;
  double* dst;
  long long ldd;
  int m, n;
};
hipError_t copy_grouped(hipStream_t s, const CopyDesc* d, int count, long long max_elems);
// out[z] = max |correction (f32)| / max |solution (f64)| of member z
hipError_t correction_ratio_grouped(hipStream_t s, const ConvDesc* d, int count, double* out);
// kb: the GEMM tiles' K chunk (8, 16 or 32), the shape the batch autotuner picks
template <typename T>
hipError_t cgemm_grouped(hipStream_t s, bool transA, const GemmDesc<T>* d, int count, int mmax, int nmax, int kb = 16);
template <typename T>
hipError_t lu_factor_grouped(hipStream_t s, const LuDesc<T>* d, int count, int nmax, int kb = 16);
template <typename T>
hipError_t lu_solve_grouped(hipStream_t s, const LuDesc<T>* d, int count, int nmax, int nrhs_max);
hipError_t convert_grouped(hipStream_t s, const ConvDesc* d, int count, long long max_elems, bool to_single,
                           bool overwrite);
hipError_t fill_i32(hipStream_t s, int* p, long long n, int v);

hipError_t demote(hipStream_t s, const double* src, int lds, long long ss, float* dst, int ldd, long long sd, int m,
                  int n, int batch);
hipError_t promote_add(hipStream_t s, const float* src, int lds, long long ss, double* dst, int ldd, long long sd,
                       int m, int n, int batch, bool overwrite);

}  // namespace hm
