// hm_htree.hip -- cluster tree, block partition, lockstep cross approximation
// of the admissible blocks and the products of the tree H-matrix (see
// hm_htree.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "hm_htree.hpp"

namespace hm {

// ---------------------------------------------------------------------------
// host: cluster tree and block partition
// ---------------------------------------------------------------------------

void htree_partition(int n, const double* x, const double* y, const double* z, int leaf, double eta,
                     HTreeHost& t) {
  t.perm.resize(n);
  std::iota(t.perm.begin(), t.perm.end(), 0);
  t.nodes.clear();
  t.blocks.clear();
  t.nleaves = 0;
  if (n <= 0) return;
  // working coordinates, permuted together with perm
  std::vector<double> w[3] = {std::vector<double>(x, x + n), std::vector<double>(y, y + n),
                              std::vector<double>(z, z + n)};
  struct Item { int s, e, parent, side; };   // inclusive range [s, e]
  std::vector<Item> stack{{0, n - 1, -1, 0}};
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const int ns = it.e - it.s + 1;
    if (ns <= 0) continue;
    const int id = (int)t.nodes.size();
    t.nodes.push_back(HTreeNode{it.s, it.e + 1, -1, -1, {0.0, 0.0, 0.0}, 0.0});
    if (it.parent >= 0) (it.side == 1 ? t.nodes[it.parent].left : t.nodes[it.parent].right) = id;
    double lo[3], hi[3], sum[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = w[d][it.s];
    for (int j = it.s; j <= it.e; ++j)
      for (int d = 0; d < 3; ++d) {
        const double v = w[d][j];
        if (v < lo[d]) lo[d] = v;
        if (v > hi[d]) hi[d] = v;
        sum[d] += v;
      }
    HTreeNode& nd = t.nodes[id];
    double len[3];
    for (int d = 0; d < 3; ++d) {
      nd.c[d] = sum[d] / (double)ns;
      len[d] = hi[d] - lo[d];
    }
    nd.diam = std::sqrt(len[0] * len[0] + len[1] * len[1] + len[2] * len[2]);
    if (ns <= leaf) {
      ++t.nleaves;
      continue;
    }
    int dim = 0;
    if (len[1] > len[dim]) dim = 1;
    if (len[2] > len[dim]) dim = 2;
    const double split = 0.5 * (lo[dim] + hi[dim]);
    // two-pointer partition: points at or below the midpoint first
    int i = it.s, j = it.e;
    while (i <= j) {
      if (w[dim][i] <= split) {
        ++i;
      } else {
        for (int d = 0; d < 3; ++d) std::swap(w[d][i], w[d][j]);
        std::swap(t.perm[i], t.perm[j]);
        --j;
      }
    }
    int left_end = i - 1;
    if (left_end < it.s) left_end = it.s;
    if (left_end >= it.e) left_end = it.e - 1;
    stack.push_back(Item{left_end + 1, it.e, id, 2});
    stack.push_back(Item{it.s, left_end, id, 1});
  }
  // block partition
  std::vector<std::pair<int, int>> bs{{0, 0}};
  while (!bs.empty()) {
    const auto [r, c] = bs.back();
    bs.pop_back();
    const HTreeNode& a = t.nodes[r];
    const HTreeNode& b = t.nodes[c];
    const int nr = a.e - a.s, nc = b.e - b.s;
    if (nr <= 0 || nc <= 0) continue;
    const bool leaf_r = a.left < 0, leaf_c = b.left < 0;
    const double dx = a.c[0] - b.c[0], dy = a.c[1] - b.c[1], dz = a.c[2] - b.c[2];
    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double dmax = std::max(a.diam, b.diam);
    if (dist > eta * dmax && nr > 1 && nc > 1) {
      t.blocks.push_back(HTreeBlock{a.s, nr, b.s, nc, (nr <= 4 || nc <= 4) ? 0 : 1});
      continue;
    }
    if (leaf_r && leaf_c) {
      t.blocks.push_back(HTreeBlock{a.s, nr, b.s, nc, 0});
      continue;
    }
    if (leaf_r) {
      bs.push_back({r, b.right});
      bs.push_back({r, b.left});
    } else if (leaf_c) {
      bs.push_back({a.right, c});
      bs.push_back({a.left, c});
    } else {
      bs.push_back({a.right, b.right});
      bs.push_back({a.right, b.left});
      bs.push_back({a.left, b.right});
      bs.push_back({a.left, b.left});
    }
  }
}

// ---------------------------------------------------------------------------
// device: cross approximation of the admissible blocks
// ---------------------------------------------------------------------------

namespace {

struct TProc {
  int rs, nr, cs, nc, kmax, pad;
  long long uoff, voff, roff, ruoff, cuoff;   // U, V, row buffer (complex); row_used, col_used (int)
};

struct TAca {
  const double* A;
  long long lda;
  const int* perm;
  const TProc* proc;
  double *U, *V, *rowbuf, *norm2;
  int *rank, *conv, *next_row, *pivot, *skip, *row_used, *col_used;
  double eps;
};

__device__ void t_argmax(double v, int idx, double* sv, int* si, double& best, int& bidx) {
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

__device__ double t_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// residual of entry (i, j) of process q: A(perm[rs+i], perm[cs+j]) - sum_l U(i,l) V(j,l)
__device__ __forceinline__ void t_residual(const TAca& h, const TProc& q, int k, int i, int j, double& er,
                                           double& ei) {
  const long long o = (long long)h.perm[q.rs + i] + (long long)h.perm[q.cs + j] * h.lda;
  er = h.A[2 * o];
  ei = h.A[2 * o + 1];
  const double* U = h.U + 2 * q.uoff;
  const double* V = h.V + 2 * q.voff;
  for (int l = 0; l < k; ++l) {
    const double ur = U[2 * (i + (long long)l * q.nr)], ui = U[2 * (i + (long long)l * q.nr) + 1];
    const double vr = V[2 * (j + (long long)l * q.nc)], vi = V[2 * (j + (long long)l * q.nc) + 1];
    er -= ur * vr - ui * vi;
    ei -= ur * vi + ui * vr;
  }
}

// residual pivot rows; one workgroup per (process, 256 columns)
__global__ __launch_bounds__(256) void k_t_eval_row(TAca h, const int2* chunks) {
  const int2 ch = chunks[blockIdx.x];
  const int p = ch.x;
  if (h.conv[p]) return;
  const TProc q = h.proc[p];
  const int j = ch.y + threadIdx.x;
  if (j >= q.nc) return;
  double er, ei;
  t_residual(h, q, h.rank[p], h.next_row[p], j, er, ei);
  h.rowbuf[2 * (q.roff + j)] = er;
  h.rowbuf[2 * (q.roff + j) + 1] = ei;
}

// pivot column = argmax |r| over unused columns, v_k = r / r(pivot); a zero
// residual row is retired and the pivot row moves to the next unused row
__global__ __launch_bounds__(256) void k_t_pivot_row(TAca h) {
  const int p = blockIdx.x;
  if (h.conv[p]) return;
  __shared__ double sv[256];
  __shared__ int si[256];
  const TProc q = h.proc[p];
  const int k = h.rank[p], i = h.next_row[p];
  const double* rb = h.rowbuf + 2 * q.roff;
  const int* cu = h.col_used + q.cuoff;
  double best = -1.0;
  int bj = q.nc;
  for (int j = threadIdx.x; j < q.nc; j += blockDim.x) {
    const double a = rb[2 * j] * rb[2 * j] + rb[2 * j + 1] * rb[2 * j + 1];
    if (!cu[j] && a > best) { best = a; bj = j; }
  }
  double emax;
  int js;
  t_argmax(best, bj, sv, si, emax, js);
  if (!(emax > 0.0) || js >= q.nc) {
    if (threadIdx.x == 0) {
      int* ru = h.row_used + q.ruoff;
      ru[i] = 1;
      int nr = -1;
      for (int t = 1; t <= q.nr; ++t)
        if (!ru[(i + t) % q.nr]) { nr = (i + t) % q.nr; break; }
      if (nr < 0) h.conv[p] = 1;
      else h.next_row[p] = nr;
      h.skip[p] = 1;
    }
    return;
  }
  const double pr = rb[2 * js], pi = rb[2 * js + 1], den = pr * pr + pi * pi;
  const double ir = pr / den, ii = -pi / den;
  double* V = h.V + 2 * (q.voff + (long long)k * q.nc);
  for (int j = threadIdx.x; j < q.nc; j += blockDim.x) {
    const double er = rb[2 * j], ei = rb[2 * j + 1];
    V[2 * j] = er * ir - ei * ii;
    V[2 * j + 1] = er * ii + ei * ir;
  }
  if (threadIdx.x == 0) {
    h.pivot[p] = js;
    h.col_used[q.cuoff + js] = 1;
    h.skip[p] = 0;
  }
}

// residual pivot columns u_k; one workgroup per (process, 256 rows)
__global__ __launch_bounds__(256) void k_t_eval_col(TAca h, const int2* chunks) {
  const int2 ch = chunks[blockIdx.x];
  const int p = ch.x;
  if (h.conv[p] || h.skip[p]) return;
  const TProc q = h.proc[p];
  const int i = ch.y + threadIdx.x;
  if (i >= q.nr) return;
  const int k = h.rank[p];
  double er, ei;
  t_residual(h, q, k, i, h.pivot[p], er, ei);
  double* U = h.U + 2 * (q.uoff + (long long)k * q.nr);
  U[2 * i] = er;
  U[2 * i + 1] = ei;
}

// norm estimate |U V^T|_F^2 with the exact cross terms, stopping test, next
// pivot row = argmax |u_k| over unused rows
__global__ __launch_bounds__(256) void k_t_finish_col(TAca h) {
  const int p = blockIdx.x;
  if (h.conv[p] || h.skip[p]) return;
  __shared__ double sh[256];
  __shared__ double sv[256];
  __shared__ int si[256];
  const TProc q = h.proc[p];
  const int k = h.rank[p], ir0 = h.next_row[p];
  const double* U = h.U + 2 * q.uoff;
  const double* V = h.V + 2 * q.voff;
  const double* uk = U + 2LL * k * q.nr;
  const double* vk = V + 2LL * k * q.nc;
  double a = 0.0, b = 0.0;
  for (int t = threadIdx.x; t < q.nr; t += blockDim.x) a += uk[2 * t] * uk[2 * t] + uk[2 * t + 1] * uk[2 * t + 1];
  for (int t = threadIdx.x; t < q.nc; t += blockDim.x) b += vk[2 * t] * vk[2 * t] + vk[2 * t + 1] * vk[2 * t + 1];
  const double nu2 = t_sum(a, sh), nv2 = t_sum(b, sh);
  // 2 Re sum_{l<k} (u_l^H u_k)(v_l^H v_k), one wave per l
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double cross = 0.0;
  for (int l = wv; l < k; l += nw) {
    const double* ul = U + 2LL * l * q.nr;
    const double* vl = V + 2LL * l * q.nc;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    for (int t = lane; t < q.nr; t += 64) {
      c0 += ul[2 * t] * uk[2 * t] + ul[2 * t + 1] * uk[2 * t + 1];
      c1 += ul[2 * t] * uk[2 * t + 1] - ul[2 * t + 1] * uk[2 * t];
    }
    for (int t = lane; t < q.nc; t += 64) {
      c2 += vl[2 * t] * vk[2 * t] + vl[2 * t + 1] * vk[2 * t + 1];
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
  cross = t_sum(lane == 0 ? cross : 0.0, sh);
  const double n2 = h.norm2[p] + nu2 * nv2 + 2.0 * cross;
  const int* ru = h.row_used + q.ruoff;
  double best = -1.0;
  int bi = q.nr;
  for (int i = threadIdx.x; i < q.nr; i += blockDim.x) {
    const double v = uk[2 * i] * uk[2 * i] + uk[2 * i + 1] * uk[2 * i + 1];
    if (!ru[i] && i != ir0 && v > best) { best = v; bi = i; }
  }
  double bmax;
  int bidx;
  t_argmax(best, bi, sv, si, bmax, bidx);
  if (threadIdx.x == 0) {
    h.row_used[q.ruoff + ir0] = 1;
    h.norm2[p] = n2;
    h.rank[p] = k + 1;
    if (sqrt(nu2 * nv2) < h.eps * sqrt(fmax(n2, 1e-30))) h.conv[p] = 1;
    else if (bidx >= q.nr) h.conv[p] = 1;        // every row used: the approximation is exact
    else if (k + 1 >= q.kmax) h.conv[p] = 2;     // rank cap reached before the tolerance
    else h.next_row[p] = bidx;
  }
}

__global__ void k_t_count_active(const int* conv, int n, int* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) mine += conv[p] == 0;
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) *out = cnt;
}

__global__ void k_t_fill(int* p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_t_fill_d(double* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// dense blocks in cluster order: one workgroup per (block, column)
struct DJob { int blk, col; };
template <typename B>
__global__ __launch_bounds__(256) void k_t_pack_dense(const double* A, long long lda, const int* perm, const B* blk,
                                                      const DJob* jobs, double* D64, float* D32) {
  const DJob jb = jobs[blockIdx.x];
  const B b = blk[jb.blk];
  const long long gj = perm[b.cs + jb.col];
  for (int i = threadIdx.x; i < b.nr; i += blockDim.x) {
    const long long o = (long long)perm[b.rs + i] + gj * lda;
    const long long d = b.off + i + (long long)jb.col * b.nr;
    const double re = A[2 * o], im = A[2 * o + 1];
    D64[2 * d] = re;
    D64[2 * d + 1] = im;
    D32[2 * d] = (float)re;
    D32[2 * d + 1] = (float)im;
  }
}

__global__ void k_t_demote(const double* src, float* dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// ---------------------------------------------------------------------------
// products
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_t_perm_in(const int* perm, int n, int ncol, const T* X, long long ldx, T* Xp) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * ncol) return;
  const int c = (int)(e / n), i = (int)(e % n);
  const long long src = (long long)perm[i] + (long long)c * ldx;
  Xp[2 * e] = X[2 * src];
  Xp[2 * e + 1] = X[2 * src + 1];
}

// t_b(r, c) = sum_j V_b(j, r) Xp(cs + j, c): one workgroup per (low-rank block,
// 4 columns), a wave per rank index, lanes along j
template <typename T, typename B>
__global__ __launch_bounds__(256) void k_t_lowrank_t(const B* blk, const T* V, const T* Xp, int n, int ncol, T* tb) {
  const B b = blk[blockIdx.x];
  const int c0 = blockIdx.y * 4;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const T* Vb = V + 2 * b.voff;
  T* t = tb + 2 * b.toff * (long long)ncol;
  for (int r = wv; r < b.rank; r += 4) {
    T acc[4][2] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
    const T* vr = Vb + 2LL * r * b.nc;
    for (int j = lane; j < b.nc; j += 64) {
      const T a = vr[2 * j], ai = vr[2 * j + 1];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (c0 + q >= ncol) break;
        const T* x = Xp + 2 * ((long long)b.cs + j + (long long)(c0 + q) * n);
        acc[q][0] += a * x[0] - ai * x[1];
        acc[q][1] += a * x[1] + ai * x[0];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      for (int o = 32; o > 0; o >>= 1) {
        acc[q][0] += __shfl_xor(acc[q][0], o);
        acc[q][1] += __shfl_xor(acc[q][1], o);
      }
    if (lane == 0)
      for (int q = 0; q < 4 && c0 + q < ncol; ++q) {
        t[2 * (r + (long long)(c0 + q) * b.rank)] = acc[q][0];
        t[2 * (r + (long long)(c0 + q) * b.rank) + 1] = acc[q][1];
      }
  }
}

// one workgroup per (leaf, 4 columns): 64 rows x 4 columns; every block whose
// row cluster contains the leaf adds its rows' part; the sums go to Y in the
// original order
template <typename T, typename B>
__global__ __launch_bounds__(256) void k_t_rows(const int2* leaf, const int* cov_start, const int2* cov, const B* blk,
                                                const T* D, const T* U, const T* Xp, const T* tb, int n, int ncol,
                                                const int* perm, T* Y, long long ldy) {
  const int2 lf = leaf[blockIdx.x];
  const int i = threadIdx.x & 63, c = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= lf.y || c >= ncol) return;
  T ar = T(0), ai = T(0);
  for (int e = cov_start[blockIdx.x]; e < cov_start[blockIdx.x + 1]; ++e) {
    const int2 cv = cov[e];
    const B b = blk[cv.x];
    const long long row = cv.y + i;
    if (b.rank < 0) {
      const T* d = D + 2 * (b.off + row);
      const T* x = Xp + 2 * ((long long)b.cs + (long long)c * n);
      for (int j = 0; j < b.nc; ++j) {
        const T dr = d[2LL * j * b.nr], di = d[2LL * j * b.nr + 1];
        const T xr = x[2 * j], xi = x[2 * j + 1];
        ar += dr * xr - di * xi;
        ai += dr * xi + di * xr;
      }
    } else {
      const T* u = U + 2 * (b.off + row);
      const T* t = tb + 2 * (b.toff * (long long)ncol + (long long)c * b.rank);
      for (int r = 0; r < b.rank; ++r) {
        const T ur = u[2LL * r * b.nr], ui = u[2LL * r * b.nr + 1];
        const T tr = t[2 * r], ti = t[2 * r + 1];
        ar += ur * tr - ui * ti;
        ai += ur * ti + ui * tr;
      }
    }
  }
  const long long o = (long long)perm[lf.x + i] + (long long)c * ldy;
  Y[2 * o] = ar;
  Y[2 * o + 1] = ai;
}

}  // namespace

// ---------------------------------------------------------------------------
// HTree
// ---------------------------------------------------------------------------

void HTree::release() {
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
  n_ = nlr_ = nleaves_ = 0;
  trows_ = 0;
  st_ = HTreeStats{};
  perm_ = nullptr;
  blk_ = nullptr;
  leaf_ = nullptr;
  cov_start_ = nullptr;
  cov_ = nullptr;
  D64_ = U64_ = V64_ = nullptr;
  D32_ = U32_ = V32_ = nullptr;
}

hipError_t HTree::build(hipStream_t s, const HTreeHost& t, const double* A, long long lda, double eps, int kmax) {
  release();
  const int n = (int)t.perm.size();
  n_ = n;
  hipError_t e;
  std::vector<void*> temp;
  auto dalloc = [&](size_t bytes, bool keep) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    (keep ? allocs_ : temp).push_back(p);
    return p;
  };
  auto free_temp = [&]() { for (void* p : temp) (void)hipFree(p); temp.clear(); };
  auto upload = [&](const void* src, size_t bytes, bool keep, void** dst) -> hipError_t {
    *dst = dalloc(bytes, keep);
    if (!*dst) return hipErrorOutOfMemory;
    return bytes ? hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  // the admissible blocks' processes
  std::vector<TProc> proc;
  long long usz = 0, vsz = 0, rsz = 0, rusz = 0, cusz = 0;
  std::vector<int2> rchunks, cchunks;
  int max_nr = 1;
  for (const HTreeBlock& b : t.blocks) {
    if (!b.lowrank) continue;
    TProc q{};
    q.rs = b.rs; q.nr = b.nr; q.cs = b.cs; q.nc = b.nc;
    q.kmax = std::min(kmax, std::min(b.nr, b.nc));
    q.uoff = usz; usz += (long long)b.nr * q.kmax;
    q.voff = vsz; vsz += (long long)b.nc * q.kmax;
    q.roff = rsz; rsz += b.nc;
    q.ruoff = rusz; rusz += b.nr;
    q.cuoff = cusz; cusz += b.nc;
    const int p = (int)proc.size();
    for (int c0 = 0; c0 < b.nc; c0 += 256) rchunks.push_back(make_int2(p, c0));
    for (int c0 = 0; c0 < b.nr; c0 += 256) cchunks.push_back(make_int2(p, c0));
    max_nr = std::max(max_nr, b.nr);
    proc.push_back(q);
  }
  const int P = (int)proc.size();
  void* vp = nullptr;
  if ((e = upload(t.perm.data(), sizeof(int) * (size_t)n, true, &vp)) != hipSuccess) return free_temp(), e;
  perm_ = static_cast<int*>(vp);
  U64_ = static_cast<double*>(dalloc(16 * (size_t)usz, true));
  V64_ = static_cast<double*>(dalloc(16 * (size_t)vsz, true));
  if (!U64_ || !V64_) return free_temp(), hipErrorOutOfMemory;
  std::vector<int> rank(P, 0), conv(P, 0);
  int steps = 0;
  if (P > 0) {
    TAca h{};
    h.A = A;
    h.lda = lda;
    h.perm = perm_;
    h.eps = eps;
    h.U = U64_;
    h.V = V64_;
    void *dproc, *drc, *dcc;
    if ((e = upload(proc.data(), sizeof(TProc) * proc.size(), false, &dproc)) != hipSuccess) return free_temp(), e;
    if ((e = upload(rchunks.data(), sizeof(int2) * rchunks.size(), false, &drc)) != hipSuccess) return free_temp(), e;
    if ((e = upload(cchunks.data(), sizeof(int2) * cchunks.size(), false, &dcc)) != hipSuccess) return free_temp(), e;
    h.proc = static_cast<const TProc*>(dproc);
    h.rowbuf = static_cast<double*>(dalloc(16 * (size_t)rsz, false));
    h.norm2 = static_cast<double*>(dalloc(8 * (size_t)P, false));
    int* ints = static_cast<int*>(dalloc(4 * (5 * (size_t)P + rusz + cusz + 1), false));
    if (!h.rowbuf || !h.norm2 || !ints) return free_temp(), hipErrorOutOfMemory;
    h.rank = ints;
    h.conv = ints + P;
    h.next_row = ints + 2 * P;
    h.pivot = ints + 3 * P;
    h.skip = ints + 4 * P;
    h.row_used = ints + 5 * P;
    h.col_used = h.row_used + rusz;
    int* cnt = h.col_used + cusz;
    const long long nints = 5LL * P + rusz + cusz + 1;
    // clear with fill kernels (no memset nodes)
    hipLaunchKernelGGL(k_t_fill, dim3((unsigned)((nints + 255) / 256)), dim3(256), 0, s, ints, nints, 0);
    hipLaunchKernelGGL(k_t_fill_d, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, h.norm2, (long long)P);
    hipLaunchKernelGGL(k_t_fill_d, dim3((unsigned)((2 * usz + 255) / 256)), dim3(256), 0, s, U64_, 2 * usz);
    hipLaunchKernelGGL(k_t_fill_d, dim3((unsigned)((2 * vsz + 255) / 256)), dim3(256), 0, s, V64_, 2 * vsz);
    if ((e = hipGetLastError()) != hipSuccess) return free_temp(), e;
    int open = P;
    // each row step adds a rank or retires a zero row: at most kmax + nr steps
    const int max_steps = kmax + max_nr;
    for (; steps < max_steps && open > 0; ++steps) {
      hipLaunchKernelGGL(k_t_eval_row, dim3((unsigned)rchunks.size()), dim3(256), 0, s, h, (const int2*)drc);
      hipLaunchKernelGGL(k_t_pivot_row, dim3(P), dim3(256), 0, s, h);
      hipLaunchKernelGGL(k_t_eval_col, dim3((unsigned)cchunks.size()), dim3(256), 0, s, h, (const int2*)dcc);
      hipLaunchKernelGGL(k_t_finish_col, dim3(P), dim3(256), 0, s, h);
      if ((e = hipGetLastError()) != hipSuccess) return free_temp(), e;
      if (steps % 4 == 3) {
        hipLaunchKernelGGL(k_t_count_active, dim3(1), dim3(256), 0, s, h.conv, P, cnt);
        if ((e = hipMemcpyAsync(&open, cnt, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return free_temp(), e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return free_temp(), e;
      }
    }
    if ((e = hipMemcpyAsync(rank.data(), h.rank, sizeof(int) * P, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return free_temp(), e;
    if ((e = hipMemcpyAsync(conv.data(), h.conv, sizeof(int) * P, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return free_temp(), e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return free_temp(), e;
  }
  // stored blocks: the low-rank ones that converged and pay, then the dense ones
  std::vector<Blk> lr, dn;
  long long toff = 0, doff = 0, stored = 0;
  int p = 0, fallback = 0, rmax = 0;
  long long rsum = 0;
  for (const HTreeBlock& b : t.blocks) {
    if (b.lowrank) {
      const TProc& q = proc[p];
      const int k = rank[p], cv = conv[p];
      ++p;
      if (cv == 1 && k > 0 && (long long)k * (b.nr + b.nc) < (long long)b.nr * b.nc) {
        lr.push_back(Blk{b.rs, b.nr, b.cs, b.nc, k, q.uoff, q.voff, toff});
        toff += k;
        stored += (long long)k * (b.nr + b.nc);
        rsum += k;
        rmax = std::max(rmax, k);
        continue;
      }
      ++fallback;
    }
    dn.push_back(Blk{b.rs, b.nr, b.cs, b.nc, -1, doff, 0, 0});
    doff += (long long)b.nr * b.nc;
  }
  stored += doff;
  nlr_ = (int)lr.size();
  trows_ = toff;
  std::vector<Blk> all(lr);
  all.insert(all.end(), dn.begin(), dn.end());
  if ((e = upload(all.data(), sizeof(Blk) * all.size(), true, &vp)) != hipSuccess) return free_temp(), e;
  blk_ = static_cast<Blk*>(vp);
  // dense blocks packed from A
  D64_ = static_cast<double*>(dalloc(16 * (size_t)doff, true));
  D32_ = static_cast<float*>(dalloc(8 * (size_t)doff, true));
  U32_ = static_cast<float*>(dalloc(8 * (size_t)usz, true));
  V32_ = static_cast<float*>(dalloc(8 * (size_t)vsz, true));
  if (!D64_ || !D32_ || !U32_ || !V32_) return free_temp(), hipErrorOutOfMemory;
  std::vector<DJob> jobs;
  for (size_t b = 0; b < dn.size(); ++b)
    for (int j = 0; j < dn[b].nc; ++j) jobs.push_back(DJob{(int)(lr.size() + b), j});
  if (!jobs.empty()) {
    if ((e = upload(jobs.data(), sizeof(DJob) * jobs.size(), false, &vp)) != hipSuccess) return free_temp(), e;
    hipLaunchKernelGGL((k_t_pack_dense<Blk>), dim3((unsigned)jobs.size()), dim3(256), 0, s, A, lda, perm_, blk_,
                       (const DJob*)vp, D64_, D32_);
  }
  if (usz) hipLaunchKernelGGL(k_t_demote, dim3((unsigned)((2 * usz + 255) / 256)), dim3(256), 0, s, U64_, U32_, 2 * usz);
  if (vsz) hipLaunchKernelGGL(k_t_demote, dim3((unsigned)((2 * vsz + 255) / 256)), dim3(256), 0, s, V64_, V32_, 2 * vsz);
  if ((e = hipGetLastError()) != hipSuccess) return free_temp(), e;
  // leaves in cluster order and the blocks covering each
  std::vector<int2> leaves;
  for (const HTreeNode& nd : t.nodes)
    if (nd.left < 0) leaves.push_back(make_int2(nd.s, nd.e - nd.s));
  std::sort(leaves.begin(), leaves.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
  nleaves_ = (int)leaves.size();
  std::vector<std::vector<int2>> per(leaves.size());
  for (size_t b = 0; b < all.size(); ++b) {
    const Blk& bl = all[b];
    auto it = std::lower_bound(leaves.begin(), leaves.end(), bl.rs, [](const int2& a, int v) { return a.x < v; });
    for (; it != leaves.end() && it->x < bl.rs + bl.nr; ++it) per[it - leaves.begin()].push_back(make_int2((int)b, it->x - bl.rs));
  }
  std::vector<int> cstart(1, 0);
  std::vector<int2> cov;
  for (auto& v : per) {
    cov.insert(cov.end(), v.begin(), v.end());
    cstart.push_back((int)cov.size());
  }
  if ((e = upload(leaves.data(), sizeof(int2) * leaves.size(), true, &vp)) != hipSuccess) return free_temp(), e;
  leaf_ = static_cast<int2*>(vp);
  if ((e = upload(cstart.data(), sizeof(int) * cstart.size(), true, &vp)) != hipSuccess) return free_temp(), e;
  cov_start_ = static_cast<int*>(vp);
  if ((e = upload(cov.data(), sizeof(int2) * cov.size(), true, &vp)) != hipSuccess) return free_temp(), e;
  cov_ = static_cast<int2*>(vp);
  e = hipStreamSynchronize(s);
  free_temp();
  if (e != hipSuccess) return e;
  st_.clusters = (int)t.nodes.size();
  st_.leaves = nleaves_;
  st_.dense_blocks = (int)dn.size();
  st_.lowrank_blocks = nlr_;
  st_.fallback_dense = fallback;
  st_.max_rank = rmax;
  st_.steps = steps;
  st_.mean_rank = nlr_ ? (double)rsum / nlr_ : 0.0;
  st_.stored_fraction = n ? (double)stored / ((double)n * n) : 0.0;
  return hipSuccess;
}

size_t HTree::workspace_bytes(int ncol, size_t real_bytes) const {
  const size_t a = 2 * real_bytes * (size_t)n_ * ncol;
  const size_t b = 2 * real_bytes * (size_t)(trows_ ? trows_ : 1) * ncol;
  return ((a + 255) & ~(size_t)255) + b;
}

template <typename T>
hipError_t HTree::apply(hipStream_t s, const T* X, long long ldx, T* Y, long long ldy, int ncol, void* work) const {
  if (ncol < 1 || n_ < 1) return hipSuccess;
  const bool dbl = sizeof(T) == sizeof(double);
  const T* D = dbl ? (const T*)D64_ : (const T*)D32_;
  const T* U = dbl ? (const T*)U64_ : (const T*)U32_;
  const T* V = dbl ? (const T*)V64_ : (const T*)V32_;
  T* Xp = static_cast<T*>(work);
  T* tb = reinterpret_cast<T*>(static_cast<char*>(work) + ((2 * sizeof(T) * (size_t)n_ * ncol + 255) & ~(size_t)255));
  const long long tot = (long long)n_ * ncol;
  hipLaunchKernelGGL((k_t_perm_in<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, perm_, n_, ncol, X, ldx, Xp);
  const unsigned cy = (unsigned)((ncol + 3) / 4);
  if (nlr_) hipLaunchKernelGGL((k_t_lowrank_t<T, Blk>), dim3(nlr_, cy), dim3(256), 0, s, blk_, V, Xp, n_, ncol, tb);
  hipLaunchKernelGGL((k_t_rows<T, Blk>), dim3(nleaves_, cy), dim3(256), 0, s, leaf_, cov_start_, cov_, blk_, D, U, Xp,
                     tb, n_, ncol, perm_, Y, ldy);
  return hipGetLastError();
}

template hipError_t HTree::apply<double>(hipStream_t, const double*, long long, double*, long long, int, void*) const;
template hipError_t HTree::apply<float>(hipStream_t, const float*, long long, float*, long long, int, void*) const;

}  // namespace hm
