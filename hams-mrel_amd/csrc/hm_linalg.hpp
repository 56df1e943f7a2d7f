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
                     int batch);

// Solve A X = B with the factors of lu_factor; B is overwritten by X.
template <typename T>
hipError_t lu_solve(hipStream_t s, const T* A, int lda, long long sA, int n, const int* ipiv, T* B, int ldb,
                    long long sB, int nrhs, int batch);

hipError_t demote(hipStream_t s, const double* src, int lds, long long ss, float* dst, int ldd, long long sd, int m,
                  int n, int batch);
hipError_t promote_add(hipStream_t s, const float* src, int lds, long long ss, double* dst, int ldd, long long sd,
                       int m, int n, int batch, bool overwrite);

}  // namespace hm
