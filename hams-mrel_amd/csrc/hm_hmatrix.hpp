// hm_hmatrix.hpp -- hierarchical (H-) matrix form of the influence matrices.
//
// Replaces HMATRIX_BUILD / HMATVEC (HMatrix.f90:135-196, :760-820) for the
// multi-body arrays: one cluster per body (bodies of equal panel count m). The
// diagonal (same-body) blocks are assembled densely; every off-diagonal block
// is admissible (bodies are separated) and is compressed by adaptive cross
// approximation with partial pivoting (ADD_LOWRANK_BLOCK :560-748) to
// A_blk ~= U V^T (plain transpose, as the influence matrices are complex
// symmetric-free), stopping when |u_k| |v_k| < eps |U V^T|_F (ACA_TOL = 1e-4,
// RANK_MAX = 50 by default, :58-60). The Frobenius norm estimate carries the
// exact cross terms 2 Re sum_l (u_l^H u_k)(v_l^H v_k); a zero residual row
// moves the pivot to the next unused row instead of ending the block.
//
// Unlike the reference, which builds the H-matrix from the assembled dense
// AMAT, the cross approximation here evaluates only the rows and columns it
// samples, directly with the Green function and Rankine kernels, so the
// off-diagonal blocks are never assembled. A and S (for the right-hand sides
// S W) are compressed independently. All processes (matrix x block) advance in
// lockstep: one launch evaluates the pivot row of every unconverged process,
// the next its pivot column.
//
// Products: y = A x for n x ncol blocks x is
//   y_b  = D_b x_b                                   (batched GEMM, diagonal blocks)
//   T_b2 = [V_{b1,b2}]_b1^T x_b2                      (batched GEMM per column block)
//   y_b1 += [U_{b1,b2}]_b2 [T_{b1,b2}]_b2               (batched GEMM per row block)
// with each block's factors at its own rank (hmat_pack), in FP32 on MFMA for the GMRES Krylov products and in FP64 for the residuals
// and the right-hand sides.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

#include "hm_assembly.hpp"

namespace hm {

struct HmatDev {
  int nb, m, kmax;
  int nmat;          // 2: A and S from the Green function; 1: A only, from `dense`
  const double* dense;  // when set: entries A(i, j) = dense[i + j ldd] of an assembled matrix
  long long ldd;
  double eps, eps_s;   // stopping tolerance for A (mat 0) and S (mat 1)
  double* U;         // [mat][b1][b2'] m x kmax, ld m   (b2' = b2 without b1)
  double* V;         // [mat][b2][b1'] m x kmax, ld m   (b1' = b1 without b2)
  double* rowbuf;    // [process][m] complex
  double* norm2;     // [process]
  int *rank, *conv, *next_row, *pivot, *skip;   // [process]
  int *row_used, *col_used;                     // [process][m]
  int* counter;                                 // one int for host polls
  // [process][chunk of ech entries][kmax] complex: the row and column launches'
  // partial inner products with the earlier factor columns (the norm estimate's
  // cross terms); room for chunks of 64 entries
  double *prow, *pcol;
  int ech;           // entries per chunk (workgroup) of the row and column launches
  // share[p] (nullptr: none): the process whose block equals process p's bit
  // for bit (equal bodies at an equal offset); p itself for the processes
  // that run. The others are skipped (conv = 3) and copied at the end.
  const int* share;
  // run (nullptr: all processes): the nrun processes that run (share[p] == p),
  // so that the step launches' grids cover those only
  const int* run;
  int nrun;
  // psym (nullptr: none): for a process that copies another's block, 0 for a
  // plain copy or the index s of the body symmetry whose panel map relates the
  // two blocks; row k of the copy's U and V factors is then row sperm[s m + k]
  // of the source's (sperm: the panel each map takes onto panel k, m per map)
  const int* psym;
  const int* sperm;
  // 1: the bodies' horizontal bounding circles are >= 0.05 h apart, so no
  // off-diagonal entry is PIDCOCK or LINTON; the entry kernels then take the
  // EIGENE-only lane groups of 8 (any such entry would still be redone per lane)
  int narrow;
};

inline int hmat_processes(int nb, int nmat = 2) { return nmat * nb * (nb - 1); }
size_t hmat_bytes(int nb, int m, int kmax, int nmat = 2);   // device memory for HmatDev arrays
HmatDev hmat_carve(void* base, int nb, int m, int kmax, double eps, double eps_s, int nmat = 2,
                   const double* dense = nullptr, long long ldd = 0);

// optional marks around each entry-kernel launch (row and column evaluation):
// mark(user, 1) before the launch, mark(user, 0) after it, on the same stream
struct AcaTimer {
  void* user;
  void (*mark)(void* user, int begin);
};

// cross approximation of all off-diagonal blocks of A (mat 0) and S (mat 1), or
// of the assembled matrix `dense` (nmat = 1); returns the number of
// unconverged processes in *unconverged
hipError_t aca_build(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int* unconverged,
                     int* steps, const AcaTimer* tm = nullptr);
// the same in two parts: aca_begin clears the state and enqueues `blind` steps
// without waiting for the device (*steps = the steps enqueued); aca_end polls,
// runs further steps while a process is open, then counts the unconverged
// processes (*steps = all steps). The host returns from aca_begin at once, so
// other streams' work can be enqueued while the steps run.
hipError_t aca_begin(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int blind, int* steps,
                     const AcaTimer* tm = nullptr);
hipError_t aca_end(hipStream_t s, const GeomDev& g, const FreqDev& f, const HmatDev& h, int* steps, int* unconverged,
                   const AcaTimer* tm = nullptr);

// One matrix's low-rank factors packed at their ranks: block row b1's U
// factors side by side, [U_{b1,b2}(:, 1:r_{b1,b2})]_{b2} (m x ru, ru the
// largest such column count, zero padded), and block column b2's V factors
// likewise (m x rv). The products then run over the ranks the blocks have
// instead of the largest rank of all.
struct LowRank {
  int ru = 0, rv = 0;
  // [b1][t < ru]: b2 rv + row of the V products (T_b2 = V_b2^T x_b2) that row
  // t of block row b1's stack takes, or -1 (padding)
  const int* tmap = nullptr;
};
// host part: column offsets of each block in its block row (offu, by b1 (nb-1)
// + b2') and block column (offv, by b2 (nb-1) + b1'), ru, rv and the map
struct LowRankPlan {
  int ru = 0, rv = 0;
  std::vector<int> offu, offv, tmap;
};
// rank[q]: the rank of block q = b1 (nb-1) + b2' of the matrix
void lowrank_plan(int nb, const int* rank, LowRankPlan& plan);
// ints of device metadata hmat_pack needs
inline size_t lowrank_meta_ints(int nb, int ru) { return 2 * (size_t)nb * (nb - 1) + (size_t)nb * (ru > 0 ? ru : 1); }
// packs matrix `mat`'s factors by `plan` into Uc (nb m ru complex) and Vc (nb
// m rv complex), with the plan's metadata uploaded to meta (synchronises s)
hipError_t hmat_pack(hipStream_t s, const HmatDev& h, int mat, const LowRankPlan& plan, int* meta, double* Uc,
                     double* Vc, LowRank* out);

// Y = D X for the right-hand sides W of the H-matrix form (column c < 6 nb is
// zero outside body c / 6, the nbeta columns after them are full), with one
// diagonal block D shared by all nb bodies: body b's rows need its own 6
// columns and the diffraction ones only, so the products of all bodies are one
// GEMM over nb (6 + nbeta) columns (xc, yc: m nb (6 + nbeta) complex each).
// Every entry of Y is written; add the low-rank part with hmat_apply (D null).
hipError_t hmat_diag_rhs(hipStream_t s, int nb, int m, int nbeta, const double* D, long long ldd, const double* X,
                         long long ldx, double* Y, long long ldy, double* xc, double* yc);

// Y = H X: diagonal blocks D (block b at D + b sD complex elements,
// column-major, ld ldd; sD = 0: one block for every body) plus the low-rank
// blocks of one matrix, packed by
// hmat_pack. X, Y: n x ncol (n = nb m); tbuf, tstack: nb max(ru, rv) ncol
// complex each. D null: Y += the low-rank blocks' product only.
template <typename T>
hipError_t hmat_apply(hipStream_t s, int nb, int m, const LowRank& lr, const T* D, long long ldd, long long sD,
                      const T* U, const T* V, const T* X, long long ldx, T* Y, long long ldy, int ncol, T* tbuf,
                      T* tstack);

}  // namespace hm
