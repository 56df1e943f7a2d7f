// hm_pairs.hpp -- the write of one assembled pair (shared by the assembly kernels).
#pragma once
#include <hip/hip_runtime.h>

#include "hm_assembly.hpp"
#include "hm_coop.hpp"
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
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  if (f.h < 0.0) infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gr);
  else fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, wvn, f.h, flag, gr);
  pair_values(g, i, j, xp, flag, gr, ar, ai, sr, si);
}

// pair_entry for a whole workgroup of lanes, one entry per lane: the PIDCOCK
// and EIGENE series are evaluated by lane groups of G (a group runs the series
// of its entries in turn, in registers), EIGEN, the frequency limits and
// infinite depth per lane as in pair_entry. All lanes call it; a lane with
// active = false (i, j must still name panels) takes part in its group's
// series and gets zeros. LINTON entries (R/h < 5e-4, whose private arrays
// would spill the whole kernel) are left out: zeros and deferred_entry() true;
// callers evaluate them with pair_entry. PID = false: PIDCOCK entries are left
// out too (an EIGENE-only kernel, for entries that are PIDCOCK only by
// exception). The cross approximation (mixed path) is its caller: the series
// take K0/K1 from sf_k01_fast.
__device__ __forceinline__ bool deferred_entry(const GeomDev& g, const FreqDev& f, int i, int j, bool pid) {
  if (!(f.h > 0.0) || fabs(f.v) < F32(1.e-8) || fabs(f.v + 1.0) < F32(1.e-8)) return false;
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const int reg = fin_regime(horiz_dist(xq[0], xp[0], xq[1], xp[1]), f.h);
  return reg == kLinton || (!pid && reg == kPidcock);
}

// DIRECT: EIGENE entries take the evanescent sums of fin_evanescent_direct,
// each lane its own entry (no epsilon table), instead of the lane groups.
template <int G, bool PID = true, bool DIRECT = false>
__device__ __forceinline__ void pair_entry_coop(const GeomDev& g, const FreqDev& f, const double* wvn, bool active,
                                                int i, int j, double& ar, double& ai, double& sr, double& si) {
  constexpr int KP = (kPidcockMaxTerms + G - 1) / G, KE = (kEigeneMaxTerms + G - 1) / G;
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2];
  const bool zero = fabs(f.v) < F32(1.e-8), inf = !zero && fabs(f.v + 1.0) < F32(1.e-8);
  const int reg = h > 0.0 && !zero && !inf ? fin_regime(rr, h) : -1;
  const int nt = !active ? 0
                 : reg == kPidcock ? (PID ? fin_pidcock_terms(rr, h) : 0)
                 : reg == kEigene && !DIRECT ? fin_eigene_terms(rr, h) : 0;
  const int l = threadIdx.x & (G - 1);
  double L1 = 0.0, L2 = 0.0, L3 = 0.0;
  // with DIRECT and no PIDCOCK entries no lane group runs a series
  if constexpr (PID || !DIRECT)
  for (int q = 0; q < G; ++q) {
    const int ntq = group_bcast<G>(nt, q);
    if (ntq == 0) continue;                  // uniform over the group
    const int rq = group_bcast<G>(reg, q);
    const double r = group_bcast<G>(rr, q);
    const double zfh = group_bcast<G>(zf, q) + h, zph = group_bcast<G>(zp, q) + h;
    double l1, l2, l3;
    if (PID && rq == kPidcock) coop_series<kPidcock, G, KP, true>(ntq, r, zfh, zph, h, f.v, wvn, l, l1, l2, l3);
    else coop_series<kEigene, G, KE, true>(ntq, r, zfh, zph, h, f.v, wvn, l, l1, l2, l3);
    if (l == q) { L1 = l1; L2 = l2; L3 = l3; }
  }
  ar = ai = sr = si = 0.0;
  if (!active) return;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  const double xpq = xq[0] - xp[0], ypq = xq[1] - xp[1];
  Grn4 gr;
  if (h < 0.0) {
    infgreen3d(xq[0], xp[0], xq[1], xp[1], zf, zp, f.v, flag, gr);
  } else if (inf) {
    fin_limit_inf(xpq, ypq, zf, zp, rr, flag, gr);
  } else {
    FinOut fo{0, 0, 0, 0, 0, 0};
    if (!zero) {
      if (reg == kPidcock) {
        if (!PID) return;                     // deferred
        fin_pidcock_head(rr, zf, zp, wvn, h, fo);
        fin_pidcock_tail(L1, L2, L3, rr, zf, zp, h, fo);
      } else if (reg == kEigene) {
        fin_propagating(rr, zf + h, zp + h, wvn[0], h, fo);
        if (DIRECT && f.evt)
          fin_evanescent_direct_tab(rr, wvn, f.evt, f.evt_ld, f.pmt, i, j, fin_direct_terms(rr, h), L1, L2, L3);
        else if (DIRECT)
          fin_evanescent_direct(rr, zf + h, zp + h, wvn, h, fin_direct_terms(rr, h), L1, L2, L3);
        fo.g_re += L1;
        fo.gr_re += L2;
        fo.gz_re += L3;
        fin_strip(rr, zf, zp, fo);
      } else if (reg == kLinton) {
        return;
      } else {
        fin_eigen(rr, zf, zp, wvn, h, fo);
      }
      if (rr == 0.0) { fo.gr_re = 0.0; fo.gr_im = 0.0; }
    }
    fin_to_grn(fo, xpq, ypq, zf, zp, rr, !zero, flag, gr);
  }
  pair_values(g, i, j, xp, flag, gr, ar, ai, sr, si);
}

// pair_entry for the cross approximation between bodies at least 0.05 h apart
// in finite depth (no PIDCOCK or LINTON entries; the depth-factor tables
// f.evt, f.pmt set), with S consecutive lanes per entry: lane `sub` of the
// entry's group sums the evanescent modes 1 + sub, 1 + sub + S, ... and the
// group combines the partial sums by a butterfly (every lane gets the same
// value); the rest of the entry is formed by every lane of the group alike. All
// S lanes of a group call it with the same (i, j).
template <int S>
__device__ __forceinline__ void pair_entry_split(const GeomDev& g, const FreqDev& f, const double* wvn, int i, int j,
                                                 int sub, double& ar, double& ai, double& sr, double& si) {
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2];
  const bool zero = fabs(f.v) < F32(1.e-8), inf = !zero && fabs(f.v + 1.0) < F32(1.e-8);
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  const double xpq = xq[0] - xp[0], ypq = xq[1] - xp[1];
  Grn4 gr;
  if (inf) {
    fin_limit_inf(xpq, ypq, zf, zp, rr, flag, gr);
  } else {
    FinOut fo{0, 0, 0, 0, 0, 0};
    if (!zero) {
      const int reg = fin_regime(rr, h);
      if (reg == kEigene) {
        fin_propagating(rr, zf + h, zp + h, wvn[0], h, fo);
        double L1, L2, L3;
        fin_evanescent_direct_tab_part(rr, wvn, f.evt, f.evt_ld, f.pmt, i, j, fin_direct_terms(rr, h), 1 + sub, S,
                                       L1, L2, L3);
#pragma unroll
        for (int o = 1; o < S; o <<= 1) {
          L1 += __shfl_xor(L1, o);
          L2 += __shfl_xor(L2, o);
          L3 += __shfl_xor(L3, o);
        }
        fo.g_re += L1;
        fo.gr_re += L2;
        fo.gz_re += L3;
        fin_strip(rr, zf, zp, fo);
      } else {
        fin_eigen(rr, zf, zp, wvn, h, fo);
      }
      if (rr == 0.0) { fo.gr_re = 0.0; fo.gr_im = 0.0; }
    }
    fin_to_grn(fo, xpq, ypq, zf, zp, rr, !zero, flag, gr);
  }
  pair_values(g, i, j, xp, flag, gr, ar, ai, sr, si);
}

}  // namespace hm
