// hm_pairs.hpp -- the write of one assembled pair (shared by the assembly kernels).
#pragma once
#include <hip/hip_runtime.h>

#include "hm_assembly.hpp"
#include "hm_device.hpp"

namespace hm {

// Rankine terms of a near pair and the entries A[i,j], S[i,j]
//   A = (i==j ? 2*pi : 0) + sum_k (dS_k + dG_k*ds_j) n_jk  (flag = 1)  or  (sum_k dG_k n_jk) ds_j
//   S = S_ij + G*ds_j  (flag = 1)  or  G*ds_j
__device__ __forceinline__ void pair_values(const GeomDev& g, int i, int j, const double* xp, int flag, const Grn4& gr,
                                            double& ar, double& ai, double& sr, double& si) {
  Rk rk{0.0, 0.0, 0.0, 0.0};
  if (flag) rankine(g.ncn[j], g.verts + 12 * (long long)j, xp, rk);
  const double ear = g.ds[j];
  const double* nv = g.dxyz + 6 * (long long)j;
  if (flag) {
    ar = (rk.dx + gr.re[1] * ear) * nv[0] + (rk.dy + gr.re[2] * ear) * nv[1] + (rk.dz + gr.re[3] * ear) * nv[2];
    ai = (gr.im[1] * ear) * nv[0] + (gr.im[2] * ear) * nv[1] + (gr.im[3] * ear) * nv[2];
    sr = rk.s + gr.re[0] * ear;
    si = gr.im[0] * ear;
  } else {
    ar = (gr.re[1] * nv[0] + gr.re[2] * nv[1] + gr.re[3] * nv[2]) * ear;
    ai = (gr.im[1] * nv[0] + gr.im[2] * nv[1] + gr.im[3] * nv[2]) * ear;
    sr = gr.re[0] * ear;
    si = gr.im[0] * ear;
  }
  if (i == j) ar = 2.0 * kPi + ar;
}

// the write of A[i,j], S[i,j] into column-major matrices with leading dimension ld
__device__ __forceinline__ void write_pair_g(const GeomDev& g, int i, int j, const double* xp, int flag, const Grn4& gr,
                                           double* A, double* S, long long ld) {
  double ar, ai, sr, si;
  pair_values(g, i, j, xp, flag, gr, ar, ai, sr, si);
  const long long off = (long long)i + (long long)j * ld;
  A[2 * off] = ar;
  A[2 * off + 1] = ai;
  S[2 * off] = sr;
  S[2 * off + 1] = si;
}

// one entry pair (A[i,j], S[i,j]) of any regime, for the entry-wise consumers
// (cross approximation); wvn may be global or LDS memory
__device__ __forceinline__ void pair_entry(const GeomDev& g, const FreqDev& f, const double* wvn, int i, int j,
                                           double& ar, double& ai, double& sr, double& si) {
  const double* xp = g.fp + 3 * (long long)i;
  const double* xq = g.xyz_p + 3 * (long long)j;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  if (f.h < 0.0) infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gr);
  else fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, wvn, f.h, flag, gr);
  pair_values(g, i, j, xp, flag, gr, ar, ai, sr, si);
}

}  // namespace hm
