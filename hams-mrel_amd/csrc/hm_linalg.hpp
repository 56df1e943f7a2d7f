// hm_linalg.hpp -- host-side launchers of the dense complex kernels (hm_linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace hm {

constexpr int kLuNb = 64;   // LU panel width (also the triangular-solve block)

// C = alpha*op(A)*B + beta*C, complex interleaved, column-major; op(A) = A or A^T.
template <typename T>
hipError_t cgemm(hipStream_t s, bool transA, int m, int n, int k, T ar, T ai, const T* A, int lda, long long sA,
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
