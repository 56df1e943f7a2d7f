// hm_htree.hpp -- H-matrix over a geometric cluster tree, for any panel set.
//
// The cluster tree and block partition follow HMATRIX_BUILD (HMatrix.f90:135-196):
// BUILD_CLUSTER_TREE / SPLIT_CLUSTER_ITERATIVE (:203-363) bisect a cluster's
// bounding box at the midpoint of its longest side until a cluster holds at
// most `leaf` points (NLEAF_DEFAULT = 32), the centre being the points' mean
// and the diameter the bounding box's diagonal; BUILD_BLOCK_TREE_ITERATIVE
// (:372-497) descends from (root, root): a block is admissible when the
// centres are farther apart than eta times the larger diameter (ETA_DEFAULT =
// 2) and both clusters hold more than one point, and is then compressed by
// cross approximation (ADD_LOWRANK_BLOCK :560-748; stored dense when a side
// holds at most 4 points, when the rank does not pay, rank (nr + nc) >= nr nc,
// or, unlike the reference, when the tolerance was not reached within the rank
// cap); an inadmissible pair of leaves is stored dense (ADD_DENSE_BLOCK
// :503-545); otherwise the children are paired.
//
// On the device every low-rank block's cross approximation advances in
// lockstep with the others (one launch evaluates the pivot rows of all open
// blocks, the next their pivot columns), reading the entries of an assembled
// matrix through the cluster permutation. Products y = H x (HMATVEC :760-820)
// run in three launches: x permuted into cluster order, t_b = V_b^T x for the
// low-rank blocks, then one workgroup per leaf cluster sums, for its rows, the
// dense blocks' and the low-rank blocks' (U_b t_b) contributions of every
// block whose row cluster contains the leaf and writes them to y in the
// original order: no atomics, a fixed summation order.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace hm {

struct HTreeNode {
  int s, e;             // cluster positions [s, e) in the permuted order
  int left, right;      // children (-1: leaf)
  double c[3];          // mean of the points
  double diam;          // bounding-box diagonal
};

struct HTreeBlock {
  int rs, nr, cs, nc;   // rows [rs, rs + nr) x columns [cs, cs + nc), permuted order
  int lowrank;          // 1: admissible, compressed by cross approximation; 0: dense
};

struct HTreeHost {
  std::vector<int> perm;             // cluster position -> original index
  std::vector<HTreeNode> nodes;      // nodes[0] is the root
  std::vector<HTreeBlock> blocks;
  int nleaves = 0;
};

// the reference's cluster tree and block partition over the n points
// (x[i], y[i], z[i]); leaf >= 2, eta > 0
void htree_partition(int n, const double* x, const double* y, const double* z, int leaf, double eta,
                     HTreeHost& t);

struct HTreeStats {
  int clusters = 0, leaves = 0, dense_blocks = 0, lowrank_blocks = 0, fallback_dense = 0, max_rank = 0, steps = 0;
  double mean_rank = 0.0, stored_fraction = 0.0;
};

// device form, built from an assembled n x n complex FP64 matrix
class HTree {
 public:
  ~HTree() { release(); }
  // compress `A` (column-major, leading dimension lda, interleaved re/im) over
  // the partition `t`: eps = the cross approximation's stopping tolerance,
  // kmax = its rank cap; also keeps FP32 copies for the Krylov products
  hipError_t build(hipStream_t s, const HTreeHost& t, const double* A, long long lda, double eps, int kmax);
  // Y = H X (n x ncol, leading dimensions ldx, ldy): FP64 factors for double,
  // the FP32 copies for float. `work` holds workspace_bytes(ncol, sizeof(T))
  template <typename T>
  hipError_t apply(hipStream_t s, const T* X, long long ldx, T* Y, long long ldy, int ncol, void* work) const;
  size_t workspace_bytes(int ncol, size_t real_bytes) const;
  const HTreeStats& stats() const { return st_; }
  int n() const { return n_; }
  void release();

 private:
  struct Blk {                       // a stored block, as the product kernels read it
    int rs, nr, cs, nc, rank;        // rank < 0: dense
    long long off, voff, toff;       // dense: D offset; low rank: U, V offsets; t offset (complex elements)
  };
  int n_ = 0, nlr_ = 0, nleaves_ = 0;
  long long trows_ = 0;              // sum of the low-rank blocks' ranks
  HTreeStats st_;
  std::vector<void*> allocs_;        // device allocations owned by this form
  int* perm_ = nullptr;
  Blk* blk_ = nullptr;               // low-rank blocks first (nlr_), then dense
  int2* leaf_ = nullptr;             // per leaf: (first row, rows)
  int* cov_start_ = nullptr;         // per leaf: start of its coverage list (nleaves + 1)
  int2* cov_ = nullptr;              // (block, row offset of the leaf inside the block)
  double *D64_ = nullptr, *U64_ = nullptr, *V64_ = nullptr;
  float *D32_ = nullptr, *U32_ = nullptr, *V32_ = nullptr;
};

}  // namespace hm
