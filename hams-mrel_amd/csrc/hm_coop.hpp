// hm_coop.hpp -- lane-cooperative evaluation of the epsilon-accelerated series.
//
// The EIGENE (0.05 <= R/h < 0.5) and PIDCOCK (5e-4 <= R/h < 0.05) regimes of
// FINGREEN3D sum up to 50 resp. 99 series terms and take the Wynn epsilon limit
// of the partial sums (Limes, FinGrnExtSubs.f90:370-503). One pair per lane
// needs 4 x 51 (4 x 100) doubles of private arrays, which the compiler places in
// scratch memory: that traffic, not FP64 issue, bounded those kernels.
//
// Here a group of G lanes (G = 4 or 8, a DPP quad or half-row) works on one
// pair at a time. Lane l of the group holds table rows i = l + G*k in register
// slot k (k < K), so the partial sums and the epsilon table stay in VGPRs:
//   * terms: each lane evaluates the terms of its rows; the partial sums are an
//     inclusive scan over the group plus a carry per slot;
//   * epsilon table: column by column as in Limes; the value of row i+1 comes
//     from the next lane (DPP), or for the group's last lane from lane 0 of the
//     next slot;
//   * the selection of Limes (the running maximum of |a| in row order, exits on
//     an exact repeat) is evaluated a slot at a time with a group prefix
//     maximum and ballots, which visits the rows in the reference's order.
// Each lane owns one pair of the list; the group loops over its G pairs.
#pragma once
#include <hip/hip_runtime.h>

#include "hm_device.hpp"

namespace hm {

// DPP reads of a lane that is switched off in EXEC return 0 here, so every call
// below is made by all lanes of the group in straight-line code, and lane
// choices are selects of values already moved (never `c ? dpp(x) : dpp(y)`,
// which evaluates each move under a partial EXEC mask).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

template <int G> struct LaneGroup;

// a DPP quad: lanes 4q .. 4q+3
template <> struct LaneGroup<4> {
  // value of row i+1 for a lane holding row i in slot k: x = slot k, xn = slot k+1
  static __device__ __forceinline__ double next(double x, double xn, int l) {
    // quad_perm [1,2,3,0]: lane l reads lane (l+1)&3
    const double a = dpp_d<0x39>(x), b = dpp_d<0x39>(xn);
    return l == 3 ? b : a;
  }
  // value of lane l-d (d = 1, 2) within the quad; meaningless for l < d
  template <int D> static __device__ __forceinline__ double prev(double x) {
    return D == 1 ? dpp_d<0x90>(x)    // quad_perm [0,0,1,2]
                  : dpp_d<0x44>(x);   // quad_perm [0,1,0,1]
  }
  static __device__ __forceinline__ double incl_sum(double x, int l) {
    double y = prev<1>(x);
    x = l >= 1 ? x + y : x;
    y = prev<2>(x);
    return l >= 2 ? x + y : x;
  }
  // max over lanes < l of v (ident for lane 0)
  static __device__ __forceinline__ double excl_max(double v, int l, double ident) {
    double u = prev<1>(v);
    u = l >= 1 ? u : ident;
    double y = prev<1>(u);
    u = l >= 1 ? fmax(u, y) : u;
    y = prev<2>(u);
    return l >= 2 ? fmax(u, y) : u;
  }
};

// half a DPP row: lanes 8q .. 8q+7
template <> struct LaneGroup<8> {
  static __device__ __forceinline__ double next(double x, double xn, int l) {
    // row_shl:1 reads lane l+1; the last lane reads lane l-7 of the next slot (row_shr:7)
    const double a = dpp_d<0x101>(x), b = dpp_d<0x117>(xn);
    return l == 7 ? b : a;
  }
  template <int D> static __device__ __forceinline__ double prev(double x) {
    return dpp_d<0x110 + D>(x);   // row_shr:D
  }
  static __device__ __forceinline__ double incl_sum(double x, int l) {
    double y = prev<1>(x);
    x = l >= 1 ? x + y : x;
    y = prev<2>(x);
    x = l >= 2 ? x + y : x;
    y = prev<4>(x);
    return l >= 4 ? x + y : x;
  }
  static __device__ __forceinline__ double excl_max(double v, int l, double ident) {
    double u = prev<1>(v);
    u = l >= 1 ? u : ident;
    double y = prev<1>(u);
    u = l >= 1 ? fmax(u, y) : u;
    y = prev<2>(u);
    u = l >= 2 ? fmax(u, y) : u;
    y = prev<4>(u);
    return l >= 4 ? fmax(u, y) : u;
  }
};

// a full DPP row: lanes 16q .. 16q+15
template <> struct LaneGroup<16> {
  static __device__ __forceinline__ double next(double x, double xn, int l) {
    // row_shr:15 brings lane 0 of the next slot to the last lane; row_shl:1 then
    // brings lane l+1 to lanes 0 .. 14, and the last lane, whose source lies
    // outside the row, keeps that value (bound_ctrl off): two moves a half, no
    // select
    (void)l;
    const int blo = __builtin_amdgcn_update_dpp(0, __double2loint(xn), 0x11F, 0xF, 0xF, true);
    const int bhi = __builtin_amdgcn_update_dpp(0, __double2hiint(xn), 0x11F, 0xF, 0xF, true);
    const int lo = __builtin_amdgcn_update_dpp(blo, __double2loint(x), 0x101, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(bhi, __double2hiint(x), 0x101, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
  }
  template <int D> static __device__ __forceinline__ double prev(double x) {
    return dpp_d<0x110 + D>(x);   // row_shr:D
  }
  static __device__ __forceinline__ double incl_sum(double x, int l) {
    double y = prev<1>(x);
    x = l >= 1 ? x + y : x;
    y = prev<2>(x);
    x = l >= 2 ? x + y : x;
    y = prev<4>(x);
    x = l >= 4 ? x + y : x;
    y = prev<8>(x);
    return l >= 8 ? x + y : x;
  }
  static __device__ __forceinline__ double excl_max(double v, int l, double ident) {
    double u = prev<1>(v);
    u = l >= 1 ? u : ident;
    double y = prev<1>(u);
    u = l >= 1 ? fmax(u, y) : u;
    y = prev<2>(u);
    u = l >= 2 ? fmax(u, y) : u;
    y = prev<4>(u);
    u = l >= 4 ? fmax(u, y) : u;
    y = prev<8>(u);
    return l >= 8 ? fmax(u, y) : u;
  }
};

// G-bit mask of the lanes of this lane's group where p holds
template <int G>
__device__ __forceinline__ unsigned group_ballot(bool p) {
  const unsigned long long b = __ballot(p);
  const int base = (threadIdx.x & 63) & ~(G - 1);
  return (unsigned)(b >> base) & ((1u << G) - 1u);
}

// value of x at lane q of this lane's group
template <int G>
__device__ __forceinline__ double group_bcast(double x, int q) {
  return __shfl(x, q, G);
}
template <int G>
__device__ __forceinline__ int group_bcast(int x, int q) {
  return __shfl(x, q, G);
}

// 1/x from the hardware estimate and two Newton steps (within an ulp or two of
// the correctly rounded quotient; the callers never pass 0)
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// max over the group (every lane receives it)
template <int G> __device__ __forceinline__ double group_max(double v);
template <> __device__ __forceinline__ double group_max<4>(double v) {
  v = fmax(v, dpp_d<0xB1>(v));   // quad_perm [1,0,3,2]
  return fmax(v, dpp_d<0x4E>(v)); // quad_perm [2,3,0,1]
}
template <> __device__ __forceinline__ double group_max<8>(double v) {
  v = group_max<4>(v);
  return fmax(v, dpp_d<0x141>(v)); // row_half_mirror
}
template <> __device__ __forceinline__ double group_max<16>(double v) {
  v = group_max<4>(v);
  v = fmax(v, dpp_d<0x141>(v));
  return fmax(v, dpp_d<0x140>(v)); // row_mirror
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) { return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, true); }
template <int G> __device__ __forceinline__ int group_min(int v);
template <> __device__ __forceinline__ int group_min<4>(int v) {
  v = min(v, dpp_i<0xB1>(v));
  return min(v, dpp_i<0x4E>(v));
}
template <> __device__ __forceinline__ int group_min<8>(int v) {
  v = group_min<4>(v);
  return min(v, dpp_i<0x141>(v));
}
template <> __device__ __forceinline__ int group_min<16>(int v) {
  v = group_min<4>(v);
  v = min(v, dpp_i<0x141>(v));
  return min(v, dpp_i<0x140>(v));
}

// column pairs of the epsilon table in the cross approximation's entries (the
// plain kernels and the mirrored diagonal-block kernels build the whole table,
// as Limes does)
constexpr int kMixedEpsPairs = 12;

// Wynn epsilon limit of the partial sums s[0..n] held as rows l + G*k of the
// group (Limes :370-503: same columns, same exact-equality tests, and the same
// result of its row-order running-maximum selection). s is overwritten.
//
// Selection per even column: with M0 the running maximum before the column, a
// row updates the result when |a| exceeds M0 and every |a| before it in the
// column. Without an exit the last update is the first row holding the column
// maximum, found by a group max and a group min of row numbers. An exit needs a
// row with an exact repeat (s_old[i+1] == s_new[i]) and |a| > M0; only columns
// that have such a row are scanned row by row.
template <int G, int K, bool FAST_RCP = true, int KKMAX = 1 << 20>
__device__ double coop_limes(int n, double (&s)[K], int l) {
  typedef LaneGroup<G> LG;
  auto inv = [](double x) { return FAST_RCP ? rcp_nr(x) : 1.0 / x; };
  double a[K], so[K];   // a column; s of the next row before the current column
  double rl_own = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    a[k] = 0.0;
    so[k] = 0.0;
    if (k == n / G) rl_own = s[k];
  }
  double rl = group_bcast<G>(rl_own, n % G);
  double amax = 0.0;
  for (int kk = 1; kk <= KKMAX; ++kk) {
    const int lim_a = n - 2 * kk + 1;
    if (lim_a < 0) break;
    // odd column: a[i] = a[i+1] + 1/(s[i+1]-s[i]) for i <= lim_a (slot k+1 is
    // still old when slot k is updated)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (G * k > lim_a) break;
      const double sk1 = k + 1 < K ? s[k + 1] : 0.0;
      const double ak1 = k + 1 < K ? a[k + 1] : 0.0;
      const double sn = LG::next(s[k], sk1, l);
      const double an = LG::next(a[k], ak1, l);
      so[k] = sn;
      const int i = l + G * k;
      if (i <= lim_a) a[k] = (sn != s[k]) ? an + inv(sn - s[k]) : an;
    }
    const int lim_s = n - 2 * kk;
    if (lim_s < 0) break;
    // even column: s[i] = s[i+1] + 1/(a[i+1]-a[i]) for i <= lim_s
    const double m0 = amax;
    double best_v = -1.0, best_s = 0.0;
    int best_k = 0;
    unsigned eqbits = 0;
    bool cand = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (G * k > lim_s) break;
      const double ak1 = k + 1 < K ? a[k + 1] : 0.0;
      const double an = LG::next(a[k], ak1, l);
      const int i = l + G * k;
      const bool act = i <= lim_s;
      const double sn = so[k];
      const double snew = act ? ((an != a[k]) ? sn + inv(an - a[k]) : sn) : s[k];
      const double v = act ? fabs(a[k]) : -1.0;
      const bool eq = act && sn == snew;
      if (v > best_v) { best_v = v; best_k = k; best_s = snew; }
      cand = cand || (eq && v > m0);
      eqbits |= (eq ? 1u : 0u) << k;
      s[k] = snew;
    }
    if (group_ballot<G>(cand)) {
      // a possible exit: the selection row by row, as Limes does it
      bool done = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (G * k > lim_s) break;
        const int i = l + G * k;
        const bool act = i <= lim_s;
        const double v = act ? fabs(a[k]) : -1.0;
        const double e = LG::excl_max(v, l, -1.0);
        const bool upd = act && v > fmax(amax, e);
        const unsigned mex = group_ballot<G>(upd && ((eqbits >> k) & 1u));
        if (mex) {
          rl = group_bcast<G>(s[k], __builtin_ctz(mex));
          done = true;
          break;
        }
        const unsigned mup = group_ballot<G>(upd);
        if (mup) rl = group_bcast<G>(s[k], 31 - __builtin_clz(mup));
        amax = fmax(amax, group_bcast<G>(fmax(e, v), G - 1));
      }
      if (done) break;
    } else {
      const double cmax = group_max<G>(best_v);
      if (cmax > m0) {
        const int row = group_min<G>(best_v == cmax ? l + G * best_k : 0x7fffffff);
        rl = group_bcast<G>(best_s, row % G);
        amax = cmax;
      }
    }
  }
  return rl;
}

// series of one pair (broadcast inputs) evaluated by the whole group; every lane
// returns the three epsilon limits. FAST: K0/K1 by sf_k01_fast (mixed paths)
template <int REG, int G, int K, bool FAST = false>
__device__ __forceinline__ void coop_series(int ntq, double r, double zfh, double zph, double h, double v,
                                            const double* wvn, int l, double& l1, double& l2, double& l3) {
  typedef LaneGroup<G> LG;
  double s1[K], s2[K], s3[K];
  double c1 = 0.0, c2 = 0.0, c3 = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int row = l + G * k;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (G * k < ntq) {
      if (row >= 1 && row < ntq) {
        if (REG == kEigene) fin_evanescent<FAST>(wvn[row], r, zfh, zph, h, t1, t2, t3);
        else fin_pidcock_term<FAST>(row, wvn[row], r, zfh, zph, v, h, t1, t2, t3);
      }
      t1 = LG::incl_sum(t1, l) + c1;
      t2 = LG::incl_sum(t2, l) + c2;
      t3 = LG::incl_sum(t3, l) + c3;
      c1 = group_bcast<G>(t1, G - 1);
      c2 = group_bcast<G>(t2, G - 1);
      c3 = group_bcast<G>(t3, G - 1);
    } else {
      t1 = c1;
      t2 = c2;
      t3 = c3;
    }
    s1[k] = t1;
    s2[k] = t2;
    s3[k] = t3;
  }
  const int n = ntq - 1;
  l1 = coop_limes<G, K, true, FAST ? kMixedEpsPairs : (1 << 20)>(n, s1, l);
  l2 = coop_limes<G, K, true, FAST ? kMixedEpsPairs : (1 << 20)>(n, s2, l);
  l3 = coop_limes<G, K, true, FAST ? kMixedEpsPairs : (1 << 20)>(n, s3, l);
}

// PIDCOCK or EIGENE series of one pair and of its mirror (depths swapped) by
// the whole group: the three limits of the pair and l4, the limit of the series
// of the zph-derivatives (the mirror's l3; its l1, l2 are the pair's)
// tc (PIDCOCK, optional): the pair's cached i pi / h halves, 4 doubles per row
template <int REG, int G, int K, bool CACHED = false>
__device__ __forceinline__ void coop_series_sym(int ntq, double r, double zfh, double zph, double h, double v,
                                                const double* wvn, int l, double& l1, double& l2, double& l3,
                                                double& l4, const double* tc = nullptr) {
  typedef LaneGroup<G> LG;
  double s1[K], s2[K], s3[K], s4[K];
  double c1 = 0.0, c2 = 0.0, c3 = 0.0, c4 = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int row = l + G * k;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
    if (G * k < ntq) {
      if (row >= 1 && row < ntq) {
        if (REG == kEigene) fin_evanescent_sym(wvn[row], r, zfh, zph, h, t1, t2, t3, t4);
        else if (CACHED) fin_pidcock_term_cached(wvn[row], r, zfh, zph, v, h, tc + 4 * row, t1, t2, t3, t4);
        else fin_pidcock_term_sym(row, wvn[row], r, zfh, zph, v, h, t1, t2, t3, t4);
      }
      t1 = LG::incl_sum(t1, l) + c1;
      t2 = LG::incl_sum(t2, l) + c2;
      t3 = LG::incl_sum(t3, l) + c3;
      t4 = LG::incl_sum(t4, l) + c4;
      c1 = group_bcast<G>(t1, G - 1);
      c2 = group_bcast<G>(t2, G - 1);
      c3 = group_bcast<G>(t3, G - 1);
      c4 = group_bcast<G>(t4, G - 1);
    } else {
      t1 = c1;
      t2 = c2;
      t3 = c3;
      t4 = c4;
    }
    s1[k] = t1;
    s2[k] = t2;
    s3[k] = t3;
    s4[k] = t4;
  }
  const int n = ntq - 1;
  l1 = coop_limes<G, K, true>(n, s1, l);
  l2 = coop_limes<G, K, true>(n, s2, l);
  l3 = coop_limes<G, K, true>(n, s3, l);
  l4 = coop_limes<G, K, true>(n, s4, l);
}

}  // namespace hm
