// hm_gmres.hpp -- restarted GMRES with FP32 MFMA products and FP64 residuals.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <functional>

namespace hm {

constexpr int kGmresMaxRestart = 128;

struct GmresArgs {
  int n = 0, ncol = 0;
  int restart = 50, maxiter = 500, max_cycles = 50;
  double tol = 1e-6;        // on the FP64 residual: ||B - A X|| <= tol ||B|| per column
  double inner_tol = 5e-7;  // on the Arnoldi residual estimate of a cycle
  // > 0: a cycle's Arnoldi steps stop once the estimate is below cycle_rtol
  // times the largest relative residual at the cycle's start (if that is above
  // inner_tol): for an inner operator whose rounding limits what a cycle reaches
  double cycle_rtol = 0.0;
  const double* A64 = nullptr;  // FP64 system matrix (n x n, leading dimension lda64)
  int lda64 = 0;
  const float* A32 = nullptr;   // FP32 copy, leading dimension n
  const double* B = nullptr;    // right-hand sides, leading dimension ldb
  int ldb = 0;
  double* X = nullptr;          // solution (n x ncol, leading dimension n)
  // block-diagonal preconditioner: nblk blocks [blk_start[b], blk_start[b+1]) (host array)
  int nblk = 0;
  const int* blk_start = nullptr;
  const float* P = nullptr;     // LU factors of the blocks, packed one after the other
  const int* pipiv = nullptr;   // block-local pivots at the block's row offset
  const float* Pinv = nullptr;  // (U^-1 L^-1) of the blocks, packed like P (gmres_block_inverse), or null
  const int* pperm = nullptr;   // the blocks' pivot permutations (gmres_block_perm), needed with Pinv
  // FP16 copy of Pinv, packed for cgemm_h16 (pack_a16: one block if
  // pinv16_shared, else nblk of them; the blocks must all have one size): the
  // Arnoldi steps then apply the preconditioner by FP16 MFMA GEMMs with FP32
  // accumulation and result instead of FP32 ones (the update of X keeps the
  // FP32 inverses). With op32_off, D M^-1 then differs from the identity by the
  // FP16 rounding of M^-1 (~1e-3), which the FP64 residual of each restart
  // corrects like any other error of the FP32 inner solve. The workspace must
  // be sized with nblk16 = nblk.
  const _Float16* Pinv16 = nullptr;
  bool pinv16_shared = false;
  // every block's inverse in Pinv is the same (blocks of one size covering
  // 0..n): the FP32 application is one GEMM over all blocks' columns
  bool pinv_shared = false;
  // operator form (used instead of A32 / A64 when set): op32 writes W = A Z in
  // FP32, op64 writes W = A X in FP64, both n x ncol with leading dimension n
  std::function<hipError_t(hipStream_t, const float*, float*)> op32;
  // the blocks of P are the FP32 LU factors of A's own diagonal blocks and
  // op32_off adds the rest of A (W += (A - D) Z): a Krylov step then forms
  // A M^-1 v as v + (A - D) M^-1 v, without the diagonal blocks' product
  std::function<hipError_t(hipStream_t, const float*, float*)> op32_off;
  std::function<hipError_t(hipStream_t, const double*, double*)> op64;
};

struct GmresStats {
  int iterations = 0, cycles = 0, converged = 0;
  double rel_residual = 0.0;
};

size_t gmres_workspace_bytes(int n, int ncol, int restart, int nblk16 = 0);
// FP32 copies of the diagonal blocks of A, LU-factored in place
hipError_t gmres_block_precond_setup(hipStream_t s, const double* A, int lda, int nblk, const int* blk_start,
                                     float* P, int* pipiv, int* info, unsigned* counter);
// row permutations of the factored blocks from their pivots, as absolute rows:
// row r of P v is row perm[r] of v
hipError_t gmres_block_perm(hipStream_t s, int nblk, const int* blk_start, const int* pipiv, int* perm);
// explicit inverses of the blocks without their row interchanges, U^-1 L^-1,
// from the LU factors P, packed like P (the block's inverse is U^-1 L^-1 P_b)
hipError_t gmres_block_inverse(hipStream_t s, int nblk, const int* blk_start, const float* P, float* Pinv);
hipError_t gmres_solve(hipStream_t s, const GmresArgs& a, void* work, GmresStats* stats);

}  // namespace hm
