// hm_surrogate.hip -- training and inference kernels of the evanescent-sum
// surrogate (see hm_surrogate.hpp for the model and its error map).
#include <hip/hip_runtime.h>
#include <math.h>

#include "hm_device.hpp"
#include "hm_pairs.hpp"
#include "hm_surrogate.hpp"

namespace hm {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(uint64_t seed, long long i, int k) {
  return (double)(mix64(seed ^ mix64((uint64_t)i * 8u + (uint64_t)k)) >> 11) * (1.0 / 9007199254740992.0);
}

// round to the nearest BF16 value (ties to even), returned as float
__device__ __forceinline__ float bf16r(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// normalized network inputs of (R/h, (z+zeta)/h, (z-zeta)/h, Kh)
__device__ __forceinline__ void sur_inputs(double rh, double ssum, double sdif, double kh, float* x) {
  x[0] = (float)(2.0 * (log(rh) - log(kSurRhMin)) / (log(kSurRhMax) - log(kSurRhMin)) - 1.0);
  x[1] = (float)(ssum / kSurDepth + 1.0);
  x[2] = (float)(sdif / kSurDepth);
  x[3] = (float)(2.0 * (log(kh) - log(kSurKhMin)) / (log(kSurKhMax) - log(kSurKhMin)) - 1.0);
}
__device__ __forceinline__ int bin(float x, int n) {
  int b = (int)((x + 1.0f) * 0.5f * (float)n);
  return b < 0 ? 0 : (b >= n ? n - 1 : b);
}
__device__ __forceinline__ int sur_cell(const float* x) {
  return ((bin(x[0], kMapR) * kMapS + bin(x[1], kMapS)) * kMapD + bin(x[2], kMapD)) * kMapK + bin(x[3], kMapK);
}
// Network target: Y = T + S e^{-k1 R} per component, with T the evanescent
// sums, S the strip terms (-(1/r + 1/r1) and its R- and z-derivatives) and k1
// the first evanescent wave number. Near R = 0 the strip cancels the 1/r
// singularity of T (Y is the smooth regular part); for large R both terms
// decay like e^{-k1 R}. Scaled target U = Y R^p e^{k1 R} (p = 1, 2, 2), which
// is dimensionless for any depth.
__device__ __forceinline__ void sur_strip(double r, double zf, double zp, double* sv) {
  const double sr = sqrt(r * r + (zf - zp) * (zf - zp)), sr1 = sqrt(r * r + (zf + zp) * (zf + zp));
  sv[0] = -(1.0 / sr + 1.0 / sr1);
  sv[1] = (r / sr) / (sr * sr) + (r / sr1) / (sr1 * sr1);
  sv[2] = ((zf - zp) / sr) / (sr * sr) + ((zf + zp) / sr1) / (sr1 * sr1);
}
__device__ __forceinline__ void sur_target(double r, double zf, double zp, double k1, double tg, double tr, double tz,
                                           double* u) {
  double sv[3];
  sur_strip(r, zf, zp, sv);
  const double e = exp(-k1 * r), ie = exp(k1 * r);
  u[0] = (tg + sv[0] * e) * r * ie;
  u[1] = (tr + sv[1] * e) * r * r * ie;
  u[2] = (tz + sv[2] * e) * r * r * ie;
}
// regular part V = T + S from a scaled target U (dimensional r, k1)
__device__ __forceinline__ void sur_regular_from(double r, double zf, double zp, double k1, const double* u, double* v) {
  double sv[3];
  sur_strip(r, zf, zp, sv);
  const double e = exp(-k1 * r), om = -expm1(-k1 * r);
  v[0] = u[0] * e / r + sv[0] * om;
  v[1] = u[1] * e / (r * r) + sv[1] * om;
  v[2] = u[2] * e / (r * r) + sv[2] * om;
}

// forward pass of one sample; w = parameters (hidden and output weights
// BF16-rounded, first layer and biases FP32). Optional per-layer storage.
__device__ __forceinline__ void mlp_eval(const float* w, const float* x, float* y, float* Z, float* A, long long zs,
                                         bool quant = true) {
  float a[kSurH], t[kSurH];
  for (int i = 0; i < kSurH; ++i) {
    float z = w[kSurOffB0 + i];
    for (int j = 0; j < kSurIn; ++j) z = fmaf(w[kSurOffW0 + i * kSurIn + j], x[j], z);
    if (Z) Z[i * zs] = z;
    a[i] = quant ? bf16r(z * sigm(z)) : z * sigm(z);
    if (A) A[i * zs] = a[i];
  }
  for (int l = 0; l < 3; ++l) {
    const float* W = w + kSurOffW1 + l * kSurHiddenStride;
    const float* b = W + kSurH * kSurH;
    for (int i = 0; i < kSurH; ++i) {
      float z = b[i];
      for (int j = 0; j < kSurH; ++j) z = fmaf(W[i * kSurH + j], a[j], z);
      t[i] = z;
    }
    for (int i = 0; i < kSurH; ++i) {
      const float z = t[i];
      if (Z) Z[((long long)(l + 1) * kSurH + i) * zs] = z;
      a[i] = quant ? bf16r(z * sigm(z)) : z * sigm(z);
      if (A) A[((long long)(l + 1) * kSurH + i) * zs] = a[i];
    }
  }
  for (int o = 0; o < kSurOut; ++o) {
    float z = w[kSurOffB4 + o];
    for (int j = 0; j < kSurH; ++j) z = fmaf(w[kSurOffW4 + o * kSurH + j], a[j], z);
    y[o] = z;
  }
}

// ---- data ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_generate(long long n, uint64_t seed, int nkh, const double* kh,
                                                  const double* wvn, float* X, float* Y, float* M, int* cell) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double rh = exp(log(kSurRhMin) + unit(seed, i, 0) * (log(kSurRhMax) - log(kSurRhMin)));
    // depths in [-0.25 h, 0], denser near the free surface (density ~ 1/sqrt(-z/h))
    const double u1 = unit(seed, i, 1), u2 = unit(seed, i, 2);
    const double zf = -kSurDepth * u1 * u1, zp = -kSurDepth * u2 * u2;
    int ik = (int)(unit(seed, i, 3) * nkh);
    ik = ik >= nkh ? nkh - 1 : ik;
    const double* w = wvn + (long long)kNk * ik;
    double tg, tr, tz, u[3];
    if (fin_regime(rh, 1.0) == kEigene) fin_eigene_evan(rh, zf, zp, w, 1.0, tg, tr, tz);
    else fin_eigen_evan(rh, zf, zp, w, 1.0, tg, tr, tz);
    sur_target(rh, zf, zp, w[1], tg, tr, tz, u);
    Y[3 * i] = (float)u[0];
    Y[3 * i + 1] = (float)u[1];
    Y[3 * i + 2] = (float)u[2];
    const double ie = exp(w[1] * rh);
    const double sc[3] = {rh * ie, rh * rh * ie, rh * rh * ie};
    float x[4];
    sur_inputs(rh, zf + zp, zf - zp, kh[ik], x);
    for (int k = 0; k < 4; ++k) X[4 * i + k] = x[k];
    if (M) {
      // size of each full component: propagating mode + evanescent sum + Rankine terms
      FinOut p{0, 0, 0, 0, 0, 0};
      fin_propagating(rh, zf + 1.0, zp + 1.0, w[0], 1.0, p);
      const double sr = sqrt(rh * rh + (zf - zp) * (zf - zp)), sr1 = sqrt(rh * rh + (zf + zp) * (zf + zp));
      const double mg = hypot(p.g_re, p.g_im) + fabs(tg) + 1.0 / sr + 1.0 / sr1;
      const double mr = hypot(p.gr_re, p.gr_im) + fabs(tr) + rh / (sr * sr * sr) + rh / (sr1 * sr1 * sr1);
      const double mz = hypot(p.gz_re, p.gz_im) + fabs(tz) + fabs(zf - zp) / (sr * sr * sr) +
                        fabs(zf + zp) / (sr1 * sr1 * sr1);
      M[3 * i] = (float)(mg * sc[0]);
      M[3 * i + 1] = (float)(mr * sc[1]);
      M[3 * i + 2] = (float)(mz * sc[2]);
      cell[i] = sur_cell(x);
    }
  }
}

// ---- training -------------------------------------------------------------------
__global__ void k_round(const float* p, float* q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kSurParams) return;
  const bool hidden_w = (i >= kSurOffW1 && i < kSurOffW4 && (i - kSurOffW1) % kSurHiddenStride < kSurH * kSurH) ||
                        (i >= kSurOffW4 && i < kSurOffB4);
  q[i] = hidden_w ? bf16r(p[i]) : p[i];
}

// Z, A: [layer 0..3][feature][sample] with sample stride 1
__global__ __launch_bounds__(256) void k_forward(int B, const float* X, const float* q, float* Z, float* A, float* Yh,
                                                 int quant) {
  __shared__ float w[kSurParams];
  for (int i = threadIdx.x; i < kSurParams; i += blockDim.x) w[i] = q[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  float y[kSurOut];
  mlp_eval(w, X + 4 * (long long)s, y, Z ? Z + s : nullptr, A ? A + s : nullptr, B, quant != 0);
  for (int o = 0; o < kSurOut; ++o) Yh[3 * (long long)s + o] = y[o];
}

// D: d loss / d z for layers 0..3 ([layer][feature][sample]), then the output
// deltas [sample][3]; loss accumulates sum of squared normalized errors
__global__ __launch_bounds__(256) void k_backward(int B, const float* Y, const float* yscale, const float* q,
                                                  const float* Z, const float* Yh, float* D, double* loss) {
  __shared__ float w[kSurParams];
  __shared__ double red[256];
  for (int i = threadIdx.x; i < kSurParams; i += blockDim.x) w[i] = q[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (s < B) {
    const float inv = 1.0f / (3.0f * (float)B);
    float dy[kSurOut];
    for (int o = 0; o < kSurOut; ++o) {
      const float e = Yh[3 * (long long)s + o] - Y[3 * (long long)s + o] / yscale[o];
      sq += (double)e * e;
      dy[o] = 2.0f * e * inv;
      D[4 * (long long)kSurH * B + 3 * (long long)s + o] = dy[o];
    }
    float da[kSurH], dz[kSurH];
    for (int j = 0; j < kSurH; ++j) {
      float t = 0.0f;
      for (int o = 0; o < kSurOut; ++o) t = fmaf(w[kSurOffW4 + o * kSurH + j], dy[o], t);
      da[j] = t;
    }
    for (int l = 3; l >= 0; --l) {
      for (int j = 0; j < kSurH; ++j) {
        const float z = Z[((long long)l * kSurH + j) * B + s];
        const float sg = sigm(z);
        dz[j] = da[j] * (sg + z * sg * (1.0f - sg));
        D[((long long)l * kSurH + j) * B + s] = dz[j];
      }
      if (l == 0) break;
      const float* W = w + kSurOffW1 + (l - 1) * kSurHiddenStride;
      for (int j = 0; j < kSurH; ++j) {
        float t = 0.0f;
        for (int i = 0; i < kSurH; ++i) t = fmaf(W[i * kSurH + j], dz[i], t);
        da[j] = t;
      }
    }
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// one thread per parameter: sum over the batch of delta x input
__global__ __launch_bounds__(256) void k_wgrad(int B, const float* X, const float* A, const float* D, float* grad) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= kSurParams) return;
  const float* dv;      // deltas of the parameter's output feature, sample stride ds
  const float* iv;      // inputs of the parameter's input feature (nullptr for a bias), stride is_
  long long ds = 1, is_ = 1;
  if (pi < kSurOffB0) {
    const int i = pi / kSurIn, j = pi % kSurIn;
    dv = D + (long long)i * B;
    iv = X + j;
    is_ = 4;
  } else if (pi < kSurOffW1) {
    dv = D + (long long)(pi - kSurOffB0) * B;
    iv = nullptr;
  } else if (pi < kSurOffW4) {
    const int l = (pi - kSurOffW1) / kSurHiddenStride, r = (pi - kSurOffW1) % kSurHiddenStride;
    if (r < kSurH * kSurH) {
      const int i = r / kSurH, j = r % kSurH;
      dv = D + ((long long)(l + 1) * kSurH + i) * B;
      iv = A + ((long long)l * kSurH + j) * B;
    } else {
      dv = D + ((long long)(l + 1) * kSurH + (r - kSurH * kSurH)) * B;
      iv = nullptr;
    }
  } else if (pi < kSurOffB4) {
    const int o = (pi - kSurOffW4) / kSurH, j = (pi - kSurOffW4) % kSurH;
    dv = D + 4 * (long long)kSurH * B + o;
    ds = 3;
    iv = A + (3 * (long long)kSurH + j) * B;
  } else {
    dv = D + 4 * (long long)kSurH * B + (pi - kSurOffB4);
    ds = 3;
    iv = nullptr;
  }
  double acc = 0.0;
  for (int s = 0; s < B; ++s) acc += (double)dv[s * ds] * (iv ? (double)iv[s * is_] : 1.0);
  grad[pi] = (float)acc;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kSurParams) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

__global__ __launch_bounds__(256) void k_errmap(long long n, const float* Y, const float* M, const int* cell,
                                                const float* U, unsigned* map_bits, double* sq) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    // error of the Green function value, and of its gradient vector (R and z
    // components together), each against the size of the full component
    const float d0 = fabsf(U[3 * i] - Y[3 * i]);
    const float d1 = U[3 * i + 1] - Y[3 * i + 1], d2 = U[3 * i + 2] - Y[3 * i + 2];
    const float e = fmaxf(d0 / M[3 * i], sqrtf(d1 * d1 + d2 * d2) / sqrtf(M[3 * i + 1] * M[3 * i + 1] +
                                                                             M[3 * i + 2] * M[3 * i + 2]));
    acc += (double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2;
    atomicMax(&map_bits[cell[i]], __float_as_uint(e));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(sq, red[0]);
}

__global__ __launch_bounds__(256) void k_predict(long long n, const float* X, SurDev m, float* U) {
  __shared__ float w[kSurParams];
  const float* src = m.quant ? m.q : m.p;
  for (int i = threadIdx.x; i < kSurParams; i += blockDim.x) w[i] = src[i];
  __syncthreads();
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float y[kSurOut];
    mlp_eval(w, X + 4 * i, y, nullptr, nullptr, 0, m.quant != 0);
    for (int o = 0; o < kSurOut; ++o) U[3 * i + o] = y[o] * m.yscale[o];
  }
}

// ---- assembly -----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_assemble_sur(GeomDev g, FreqDev f, double kh, const uint32_t* __restrict__ pairs,
                                                      long long np, SurDev m, float tol, double* A, double* S,
                                                      long long ld, uint32_t* fb_e, uint32_t* fb_g,
                                                      unsigned long long* fb_count) {
  __shared__ float w[kSurParams];
  __shared__ double wvn[kNk];
  const float* src = m.quant ? m.q : m.p;
  for (int i = threadIdx.x; i < kSurParams; i += blockDim.x) w[i] = src[i];
  for (int i = threadIdx.x; i < kNk; i += blockDim.x) wvn[i] = f.wvn[i];
  __syncthreads();
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < np;
  const uint32_t p = valid ? pairs[t] : 0u;
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  const double* xp = g.fp + 3 * (long long)i;
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2], rh = rr / h;
  float x[4];
  sur_inputs(rh > kSurRhMin ? rh : kSurRhMin, (zf + zp) / h, (zf - zp) / h, kh, x);
  // R/h >= 0.5 (EIGEN, at most 9 terms) is cheaper as plain series than as a network evaluation
  const bool inside = rh >= kSurRhMin && rh < 0.5 && zf >= -kSurDepth * h && zp >= -kSurDepth * h;
  const bool net = inside && m.map[sur_cell(x)] <= tol;
  const bool plain_far = rh >= 0.5;
  // pairs handed to the plain kernels: wave-aggregated append per regime
  const bool to_e = valid && !net && !plain_far && rh < 0.5, to_g = valid && !net && !plain_far && rh >= 0.5;
  const unsigned long long be = __ballot(to_e), bg = __ballot(to_g), bf = __ballot(valid && plain_far);
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long base_e = 0, base_g = 0;
  if (lane == 0) {
    if (be) base_e = atomicAdd(&fb_count[0], (unsigned long long)__popcll(be));
    if (bg) base_g = atomicAdd(&fb_count[1], (unsigned long long)__popcll(bg));
    if (bf) atomicAdd(&fb_count[2], (unsigned long long)__popcll(bf));   // plain series in this kernel
  }
  base_e = __shfl(base_e, 0);
  base_g = __shfl(base_g, 0);
  if (to_e) fb_e[base_e + __popcll(be & lt)] = p;
  if (to_g) fb_g[base_g + __popcll(bg & lt)] = p;
  if (!valid || !(net || plain_far)) return;
  FinOut fo{0, 0, 0, 0, 0, 0};
  if (plain_far) {
    fin_eigen(rr, zf, zp, wvn, h, fo);
  } else {
    float y[kSurOut];
    mlp_eval(w, x, y, nullptr, nullptr, 0, m.quant != 0);
    // U = Y R^p e^{k1 R} is dimensionless for any depth (R, k1 dimensional)
    const double u[3] = {(double)(y[0] * m.yscale[0]), (double)(y[1] * m.yscale[1]), (double)(y[2] * m.yscale[2])};
    double v[3];
    sur_regular_from(rr, zf, zp, wvn[1], u, v);
    fin_propagating(rr, zf + h, zp + h, wvn[0], h, fo);
    fo.g_re += v[0];
    fo.gr_re += v[1];
    fo.gz_re += v[2];
  }
  if (rr == 0.0) { fo.gr_re = 0.0; fo.gr_im = 0.0; }
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  fin_to_grn(fo, xq[0] - xp[0], xq[1] - xp[1], zf, zp, rr, true, flag, gr);
  write_pair_g(g, i, j, xp, flag, gr, A, S, ld);
}


// network and plain evanescent sums (unscaled, h = 1) at given points, with the
// error-map bound of each point's cell
__global__ __launch_bounds__(256) void k_eval_points(int n, const double* rh, const double* zs, const double* zd,
                                                     const double* kh, const double* wvn, SurDev m, double* tn,
                                                     double* tp, double* bound) {
  __shared__ float w[kSurParams];
  const float* src = m.quant ? m.q : m.p;
  for (int i = threadIdx.x; i < kSurParams; i += blockDim.x) w[i] = src[i];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* wv = wvn + (long long)kNk * i;
  const double zf = 0.5 * (zs[i] + zd[i]), zp = 0.5 * (zs[i] - zd[i]);
  double tg, tr, tz, u[3], v[3];
  if (fin_regime(rh[i], 1.0) == kEigene) fin_eigene_evan(rh[i], zf, zp, wv, 1.0, tg, tr, tz);
  else fin_eigen_evan(rh[i], zf, zp, wv, 1.0, tg, tr, tz);
  sur_target(rh[i], zf, zp, wv[1], tg, tr, tz, u);
  sur_regular_from(rh[i], zf, zp, wv[1], u, v);
  for (int o = 0; o < kSurOut; ++o) tp[3 * i + o] = v[o];
  float x[4], y[kSurOut];
  sur_inputs(rh[i], zs[i], zd[i], kh[i], x);
  mlp_eval(w, x, y, nullptr, nullptr, 0, m.quant != 0);
  for (int o = 0; o < kSurOut; ++o) u[o] = (double)(y[o] * m.yscale[o]);
  sur_regular_from(rh[i], zf, zp, wv[1], u, v);
  for (int o = 0; o < kSurOut; ++o) tn[3 * i + o] = v[o];
  bound[i] = m.map[sur_cell(x)];
}

}  // namespace

hipError_t sur_generate(hipStream_t s, long long n, uint64_t seed, int nkh, const double* kh, const double* wvn,
                        float* X, float* Y, float* M, int* cell) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, n, seed, nkh, kh,
                     wvn, X, Y, M, cell);
  return hipGetLastError();
}
hipError_t sur_round(hipStream_t s, const float* p, float* q) {
  hipLaunchKernelGGL(k_round, dim3((kSurParams + 255) / 256), dim3(256), 0, s, p, q);
  return hipGetLastError();
}
hipError_t sur_forward(hipStream_t s, int B, const float* X, const float* q, float* Z, float* A, float* Yh, bool quant) {
  hipLaunchKernelGGL(k_forward, dim3((B + 255) / 256), dim3(256), 0, s, B, X, q, Z, A, Yh, quant ? 1 : 0);
  return hipGetLastError();
}
hipError_t sur_loss_grad(hipStream_t s, int B, const float* X, const float* Y, const float* yscale, const float* q,
                         const float* Z, const float* A, const float* Yh, float* D, float* grad, float* loss) {
  hipLaunchKernelGGL(k_backward, dim3((B + 255) / 256), dim3(256), 0, s, B, Y, yscale, q, Z, Yh, D, (double*)loss);
  hipLaunchKernelGGL(k_wgrad, dim3((kSurParams + 255) / 256), dim3(256), 0, s, B, X, A, D, grad);
  return hipGetLastError();
}
hipError_t sur_adam(hipStream_t s, float* p, const float* grad, float* m, float* v, float lr, int t) {
  const float c1 = 1.0f / (1.0f - powf(0.9f, (float)t)), c2 = 1.0f / (1.0f - powf(0.999f, (float)t));
  hipLaunchKernelGGL(k_adam, dim3((kSurParams + 255) / 256), dim3(256), 0, s, p, grad, m, v, lr, c1, c2);
  return hipGetLastError();
}
hipError_t sur_error_map(hipStream_t s, long long n, const float* Y, const float* M, const int* cell, const float* U,
                         unsigned* map_bits, double* sq_err) {
  hipLaunchKernelGGL(k_errmap, dim3(1024), dim3(256), 0, s, n, Y, M, cell, U, map_bits, sq_err);
  return hipGetLastError();
}
hipError_t sur_predict(hipStream_t s, long long n, const float* X, SurDev m, float* U) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_predict, dim3((unsigned)(blocks < 16384 ? blocks : 16384)), dim3(256), 0, s, n, X, m, U);
  return hipGetLastError();
}
hipError_t assemble_surrogate(hipStream_t s, const GeomDev& g, const FreqDev& f, double kh, const uint32_t* pairs,
                              long long np, SurDev m, float tol, double* A, double* S, long long ld,
                              uint32_t* fb_eigene, uint32_t* fb_eigen, unsigned long long* fb_count) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_assemble_sur, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, g, f, kh, pairs, np, m, tol,
                     A, S, ld, fb_eigene, fb_eigen, fb_count);
  return hipGetLastError();
}

hipError_t sur_eval_points(hipStream_t s, int n, const double* rh, const double* zs, const double* zd, const double* kh,
                           const double* wvn, double g_disp, SurDev m, double* tn, double* tp, double* bound) {
  (void)g_disp;
  double* d = nullptr;
  const size_t per = 4 + kNk + 3 + 3 + 1;
  hipError_t e = hipMalloc(&d, sizeof(double) * per * (size_t)n);
  if (e != hipSuccess) return e;
  double *drh = d, *dzs = d + n, *dzd = d + 2 * (size_t)n, *dkh = d + 3 * (size_t)n, *dw = d + 4 * (size_t)n;
  double *dtn = dw + (size_t)kNk * n, *dtp = dtn + 3 * (size_t)n, *db = dtp + 3 * (size_t)n;
  const size_t nb = sizeof(double) * n;
  if ((e = hipMemcpyAsync(drh, rh, nb, hipMemcpyHostToDevice, s)) == hipSuccess &&
      (e = hipMemcpyAsync(dzs, zs, nb, hipMemcpyHostToDevice, s)) == hipSuccess &&
      (e = hipMemcpyAsync(dzd, zd, nb, hipMemcpyHostToDevice, s)) == hipSuccess &&
      (e = hipMemcpyAsync(dkh, kh, nb, hipMemcpyHostToDevice, s)) == hipSuccess &&
      (e = hipMemcpyAsync(dw, wvn, nb * kNk, hipMemcpyHostToDevice, s)) == hipSuccess) {
    hipLaunchKernelGGL(k_eval_points, dim3((n + 255) / 256), dim3(256), 0, s, n, drh, dzs, dzd, dkh, dw, m, dtn, dtp, db);
    if ((e = hipGetLastError()) == hipSuccess &&
        (e = hipMemcpyAsync(tn, dtn, 3 * nb, hipMemcpyDeviceToHost, s)) == hipSuccess &&
        (e = hipMemcpyAsync(tp, dtp, 3 * nb, hipMemcpyDeviceToHost, s)) == hipSuccess &&
        (e = hipMemcpyAsync(bound, db, nb, hipMemcpyDeviceToHost, s)) == hipSuccess)
      e = hipStreamSynchronize(s);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return e;
}

}  // namespace hm
