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

// ---- inference on MFMA ----------------------------------------------------------
// The three 64 x 64 hidden layers of a wave's 64 pairs are 16 x 16 x 16 BF16
// MFMA tiles with FP32 accumulation, computed transposed: out[o][pair] =
// sum_k W[o][k] act[pair][k] (weights as the A operand, 4 output tiles x 4 pair
// tiles x 4 input steps per layer), so that each lane ends up with 4
// consecutive output features of one pair and stores them, with bias, SiLU and
// BF16 rounding applied, as one 8-byte LDS write in the layout the next layer
// reads. The first layer (FP32 weights) and the output layer run per lane.
// Operand layout of v_mfma_f32_16x16x16_bf16 (lane l, slot i < 4, k = 4 (l / 16)
// + i): A[l % 16][k], B[k][l % 16]; result slot v is D[4 (l / 16) + v][l % 16].
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// element (row, col) of a 64 x 64 BF16 tile: the 8-byte chunk index is XORed
// with the row (mod 16), so the 16 rows of a fragment access fall on distinct
// LDS banks
__device__ __forceinline__ int tile_at(int row, int col) {
  return row * kSurH + ((((col >> 2) ^ (row & 15))) << 2) + (col & 3);
}

// the network as the inference kernels hold it in LDS (also its image in
// device memory, built once per parameter change by k_build_img)
struct alignas(16) SurImg {
  unsigned short w[3][kSurH * kSurH];        // hidden weights [out][in], BF16 bits, tile_at layout
  float b[3][kSurH];
  float w0[kSurH][kSurIn], b0[kSurH];
  float w4[kSurOut][kSurH], b4[kSurOut], pad;
};
static_assert(sizeof(SurImg) == kSurImgBytes, "surrogate image size");
static_assert(kSurImgBytes % 16 == 0, "surrogate image copies in 16-byte units");
struct alignas(16) SurLds {
  SurImg m;
  unsigned short act[4][kSurH * kSurH];      // per wave: 64 pairs x 64 activations, BF16 bits, tile_at layout
};

__device__ __forceinline__ unsigned short bf16_bits(float x) { return (unsigned short)(__float_as_uint(x) >> 16); }
__device__ __forceinline__ float bf16_val(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
// two floats to BF16 (round to nearest even, as bf16r) packed low, high
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, bf16x2v));
}
// z sigmoid(z) with the hardware exp2 and reciprocal (relative error ~1e-7,
// far below the BF16 rounding that follows)
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// image of the BF16-rounded parameters q (one workgroup)
__global__ __launch_bounds__(256) void k_build_img(const float* q, SurImg* img) {
  for (int e = threadIdx.x; e < 3 * kSurH * kSurH; e += blockDim.x) {
    const int l = e / (kSurH * kSurH), r = e % (kSurH * kSurH);
    img->w[l][tile_at(r / kSurH, r % kSurH)] = bf16_bits(q[kSurOffW1 + l * kSurHiddenStride + r]);
  }
  for (int e = threadIdx.x; e < 3 * kSurH; e += blockDim.x)
    img->b[e / kSurH][e % kSurH] = q[kSurOffW1 + (e / kSurH) * kSurHiddenStride + kSurH * kSurH + e % kSurH];
  for (int e = threadIdx.x; e < kSurH * kSurIn; e += blockDim.x) img->w0[e / kSurIn][e % kSurIn] = q[kSurOffW0 + e];
  for (int e = threadIdx.x; e < kSurH; e += blockDim.x) img->b0[e] = q[kSurOffB0 + e];
  for (int e = threadIdx.x; e < kSurOut * kSurH; e += blockDim.x) img->w4[e / kSurH][e % kSurH] = q[kSurOffW4 + e];
  for (int e = threadIdx.x; e < kSurOut; e += blockDim.x) img->b4[e] = q[kSurOffB4 + e];
  if (threadIdx.x == 0) img->pad = 0.0f;
}

__device__ __forceinline__ void load_sur_lds(SurImg& L, const void* img) {
  const uint4* src = static_cast<const uint4*>(img);
  uint4* dst = reinterpret_cast<uint4*>(&L);
  for (int e = threadIdx.x; e < (int)(kSurImgBytes / 16); e += blockDim.x) dst[e] = src[e];
}

// y = network(x) for this lane's pair; every thread of the (256-thread)
// workgroup calls it, after L.m is loaded and a barrier
__device__ void mlp_block_mfma(SurLds& L, const float* x, float* y) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned short* act = L.act[w];
  for (int c = 0; c < kSurH / 4; ++c) {
    float a[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 4 * c + v;
      float z = L.m.b0[i];
#pragma unroll
      for (int j = 0; j < kSurIn; ++j) z = fmaf(L.m.w0[i][j], x[j], z);
      a[v] = silu_fast(z);
    }
    *reinterpret_cast<uint2*>(&act[tile_at(lane, 4 * c)]) = uint2{pack_bf16(a[0], a[1]), pack_bf16(a[2], a[3])};
  }
  const int r16 = lane & 15, kq = 4 * (lane >> 4);
  for (int l = 0; l < 3; ++l) {
    __syncthreads();
    f32x4 acc[4][4];                         // [output tile][pair tile]
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) acc[ot][pt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bf16x4 wf[4], af[4];
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) wf[ot] = *reinterpret_cast<const bf16x4*>(&L.m.w[l][tile_at(16 * ot + r16, 16 * s + kq)]);
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) af[pt] = *reinterpret_cast<const bf16x4*>(&act[tile_at(16 * pt + r16, 16 * s + kq)]);
#pragma unroll
      for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int pt = 0; pt < 4; ++pt)
          acc[ot][pt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf[ot], af[pt], acc[ot][pt], 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      const float4 bb = *reinterpret_cast<const float4*>(&L.m.b[l][16 * ot + kq]);
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) {
        const float a0 = silu_fast(acc[ot][pt][0] + bb.x), a1 = silu_fast(acc[ot][pt][1] + bb.y);
        const float a2 = silu_fast(acc[ot][pt][2] + bb.z), a3 = silu_fast(acc[ot][pt][3] + bb.w);
        *reinterpret_cast<uint2*>(&act[tile_at(16 * pt + r16, 16 * ot + kq)]) = uint2{pack_bf16(a0, a1), pack_bf16(a2, a3)};
      }
    }
  }
  __syncthreads();
  float z0 = L.m.b4[0], z1 = L.m.b4[1], z2 = L.m.b4[2];
  for (int c = 0; c < kSurH / 4; ++c) {
    const uint2 u = *reinterpret_cast<const uint2*>(&act[tile_at(lane, 4 * c)]);
    const float a[4] = {bf16_val((unsigned short)u.x), bf16_val((unsigned short)(u.x >> 16)),
                        bf16_val((unsigned short)u.y), bf16_val((unsigned short)(u.y >> 16))};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      z0 = fmaf(L.m.w4[0][4 * c + v], a[v], z0);
      z1 = fmaf(L.m.w4[1][4 * c + v], a[v], z1);
      z2 = fmaf(L.m.w4[2][4 * c + v], a[v], z2);
    }
  }
  y[0] = z0;
  y[1] = z1;
  y[2] = z2;
}

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

// Training runs layer by layer with one thread per (sample, feature): the
// feature is uniform over a workgroup (blockIdx.y), so the weight row is read
// through scalar loads, and each sum runs over its inputs in the same order as
// mlp_eval. Z, A: [layer 0..3][feature][sample] with sample stride 1.
//
// forward of layer l: 0 (inputs X [sample][4]), 1..3 (hidden), 4 (output -> Yh [sample][3])
__global__ __launch_bounds__(256) void k_fwd_layer(int B, int l, const float* __restrict__ X, const float* __restrict__ q,
                                                   float* __restrict__ Z, float* __restrict__ A, float* __restrict__ Yh,
                                                   int quant) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (s >= B) return;
  const long long Bl = B;
  float z;
  if (l == 0) {
    z = q[kSurOffB0 + j];
    for (int i = 0; i < kSurIn; ++i) z = fmaf(q[kSurOffW0 + j * kSurIn + i], X[4 * (long long)s + i], z);
  } else {
    const float* W = l < 4 ? q + kSurOffW1 + (l - 1) * kSurHiddenStride : q + kSurOffW4;
    const float* in = A + (long long)(l - 1) * kSurH * Bl + s;
    z = l < 4 ? W[kSurH * kSurH + j] : q[kSurOffB4 + j];
    for (int i = 0; i < kSurH; ++i) z = fmaf(W[j * kSurH + i], in[i * Bl], z);
  }
  if (l == 4) {
    Yh[3 * (long long)s + j] = z;
    return;
  }
  const long long o = ((long long)l * kSurH + j) * Bl + s;
  Z[o] = z;
  A[o] = quant ? bf16r(z * sigm(z)) : z * sigm(z);
}

// output deltas D[4 * 64 * B + 3 s + o] = d loss / d y; loss accumulates the
// sum of squared normalized errors
__global__ __launch_bounds__(256) void k_out_delta(int B, const float* __restrict__ Y, const float* __restrict__ yscale,
                                                   const float* __restrict__ Yh, float* __restrict__ D, double* loss) {
  __shared__ double red[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (s < B) {
    const float inv = 1.0f / (3.0f * (float)B);
    for (int o = 0; o < kSurOut; ++o) {
      const float e = Yh[3 * (long long)s + o] - Y[3 * (long long)s + o] / yscale[o];
      sq += (double)e * e;
      D[4 * (long long)kSurH * B + 3 * (long long)s + o] = 2.0f * e * inv;
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

// deltas of layer l (3..0): D[l][j][s] = (sum_i Wnext[i][j] Dnext[i][s]) silu'(Z[l][j][s])
__global__ __launch_bounds__(256) void k_bwd_layer(int B, int l, const float* __restrict__ q, const float* __restrict__ Z,
                                                   float* __restrict__ D) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (s >= B) return;
  const long long Bl = B;
  float t = 0.0f;
  if (l == 3) {
    const float* dy = D + 4 * (long long)kSurH * Bl + 3 * (long long)s;
    for (int o = 0; o < kSurOut; ++o) t = fmaf(q[kSurOffW4 + o * kSurH + j], dy[o], t);
  } else {
    const float* W = q + kSurOffW1 + l * kSurHiddenStride;
    const float* dn = D + (long long)(l + 1) * kSurH * Bl + s;
    for (int i = 0; i < kSurH; ++i) t = fmaf(W[i * kSurH + j], dn[i * Bl], t);
  }
  const long long o = ((long long)l * kSurH + j) * Bl + s;
  const float z = Z[o];
  const float sg = sigm(z);
  D[o] = t * (sg + z * sg * (1.0f - sg));
}

// weight gradients in two passes: part[c][pi] = sum over samples of chunk c
// (kSurChunk samples) of delta x input, in FP64; then grad[pi] = sum over
// chunks. The hidden weight matrices (12288 of the parameters) are 64 x 64 x
// chunk products, tiled through LDS (k_wgrad_hidden); k_wgrad takes the rest.
constexpr int kSurChunk = 64;
__device__ __forceinline__ bool hidden_weight(int pi) {
  return pi >= kSurOffW1 && pi < kSurOffW4 && (pi - kSurOffW1) % kSurHiddenStride < kSurH * kSurH;
}

// dW_l[i][j] over one chunk = sum_s D[l + 1][i][s] A[l][j][s]; grid (chunk, l).
// Operands transposed into LDS ([s][feature], row stride 68 floats) so that a
// thread's 4 consecutive features of one sample are one 16-byte read; each
// thread accumulates a 4 x 4 block in FP64.
__global__ __launch_bounds__(256) void k_wgrad_hidden(int B, const float* __restrict__ A, const float* __restrict__ D,
                                                      double* __restrict__ part) {
  constexpr int kLd = kSurH + 4;
  __shared__ alignas(16) float dt[kSurChunk][kLd];
  __shared__ alignas(16) float at[kSurChunk][kLd];
  const int c = blockIdx.x, l = blockIdx.y;
  const int s0 = c * kSurChunk, ns = B - s0 < kSurChunk ? B - s0 : kSurChunk;
  const long long Bl = B;
  const float* Dl = D + (long long)(l + 1) * kSurH * Bl + s0;
  const float* Al = A + (long long)l * kSurH * Bl + s0;
  for (int e = threadIdx.x; e < kSurH * kSurChunk; e += blockDim.x) {
    const int f = e / kSurChunk, sm = e % kSurChunk;
    dt[sm][f] = sm < ns ? Dl[f * Bl + sm] : 0.0f;
    at[sm][f] = sm < ns ? Al[f * Bl + sm] : 0.0f;
  }
  __syncthreads();
  const int i0 = 4 * (threadIdx.x >> 4), j0 = 4 * (threadIdx.x & 15);
  double acc[4][4] = {};
  for (int sm = 0; sm < ns; ++sm) {
    const float4 dv = *reinterpret_cast<const float4*>(&dt[sm][i0]);
    const float4 av = *reinterpret_cast<const float4*>(&at[sm][j0]);
    const double d4[4] = {dv.x, dv.y, dv.z, dv.w}, a4[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(d4[a], a4[b], acc[a][b]);
  }
  double* out = part + (long long)c * kSurParams + kSurOffW1 + l * kSurHiddenStride;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[(i0 + a) * kSurH + j0 + b] = acc[a][b];
}

__global__ __launch_bounds__(256) void k_wgrad(int B, const float* X, const float* A, const float* D, double* part) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (pi >= kSurParams || hidden_weight(pi)) return;
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
  const int s0 = c * kSurChunk, s1 = s0 + kSurChunk < B ? s0 + kSurChunk : B;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int sm = s0;
  if (iv) {
    for (; sm + 4 <= s1; sm += 4) {
      a0 += (double)dv[sm * ds] * (double)iv[sm * is_];
      a1 += (double)dv[(sm + 1) * ds] * (double)iv[(sm + 1) * is_];
      a2 += (double)dv[(sm + 2) * ds] * (double)iv[(sm + 2) * is_];
      a3 += (double)dv[(sm + 3) * ds] * (double)iv[(sm + 3) * is_];
    }
    for (; sm < s1; ++sm) a0 += (double)dv[sm * ds] * (double)iv[sm * is_];
  } else {
    for (; sm + 4 <= s1; sm += 4) {
      a0 += (double)dv[sm * ds];
      a1 += (double)dv[(sm + 1) * ds];
      a2 += (double)dv[(sm + 2) * ds];
      a3 += (double)dv[(sm + 3) * ds];
    }
    for (; sm < s1; ++sm) a0 += (double)dv[sm * ds];
  }
  part[(long long)c * kSurParams + pi] = (a0 + a1) + (a2 + a3);
}

__global__ __launch_bounds__(256) void k_wgrad_sum(int nchunk, const double* part, float* grad) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= kSurParams) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += part[(long long)c * kSurParams + pi];
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

// one pair per thread; the BF16 network runs on MFMA (the FP32 network, quant
// = 0, per lane from global memory)
__global__ __launch_bounds__(256) void k_predict(long long n, const float* X, SurDev m, float* U) {
  __shared__ SurLds L;
  if (m.quant) load_sur_lds(L.m, m.img);
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ic = i < n ? i : n - 1;
  float y[kSurOut];
  if (m.quant) mlp_block_mfma(L, X + 4 * ic, y);
  else mlp_eval(m.p, X + 4 * ic, y, nullptr, nullptr, 0, false);
  if (i >= n) return;
  for (int o = 0; o < kSurOut; ++o) U[3 * i + o] = y[o] * m.yscale[o];
}

// ---- assembly -----------------------------------------------------------------------
// One pair of the EIGENE segment per thread. A pair inside the trained range
// whose error-map cell bound is <= tol takes the network: done[t] = 1 and
// vreg[3 t ..] = its regular part V (value, R- and z-derivative) from the
// network; every other pair gets done[t] = 0. blk[b] = pairs taken by workgroup
// b. The MFMA network needs the whole workgroup, so a workgroup runs it when
// any of its pairs takes it (the segment is in key order, i.e. by R/h, so whole
// workgroups mostly agree). SYM: the pair (i, j) stands for its mirror (j, i)
// too (depths swapped); it is taken only if both directions are, and vreg holds
// 6 floats a pair (the pair's V, then the mirror's).
template <bool QUANT, bool SYM>
__global__ __launch_bounds__(256) void k_sur_regular(GeomDev g, FreqDev f, double kh, const uint32_t* __restrict__ pairs,
                                                     long long np, SurDev m, float tol, uint8_t* __restrict__ done,
                                                     float* __restrict__ vreg, unsigned* __restrict__ blk,
                                                     unsigned long long* count) {
  __shared__ SurLds L;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < np;
  const uint32_t p = valid ? pairs[t] : pairs[np - 1];
  const int i = (int)(p >> 16), j = (int)(p & 0xffffu);
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double h = f.h;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double zf = xq[2], zp = xp[2], rh = rr / h;
  float x[4], xm[4];
  sur_inputs(rh > kSurRhMin ? rh : kSurRhMin, (zf + zp) / h, (zf - zp) / h, kh, x);
  if (SYM) sur_inputs(rh > kSurRhMin ? rh : kSurRhMin, (zp + zf) / h, (zp - zf) / h, kh, xm);
  const bool inside = rh >= kSurRhMin && rh < 0.5 && zf >= -kSurDepth * h && zp >= -kSurDepth * h;
  const bool net = valid && inside && m.map[sur_cell(x)] <= tol && (!SYM || m.map[sur_cell(xm)] <= tol);
  if (valid) done[t] = net ? 1 : 0;
  const int nb = __syncthreads_count(net);   // uniform over the workgroup
  if (threadIdx.x == 0) {
    blk[blockIdx.x] = (unsigned)nb;
    if (nb) atomicAdd(count, (unsigned long long)nb * (SYM ? 2u : 1u));
  }
  constexpr int nv = SYM ? 6 : 3;
  float y[kSurOut], ym[kSurOut];
  if (QUANT) {
    if (nb == 0) return;
    load_sur_lds(L.m, m.img);
    __syncthreads();
    mlp_block_mfma(L, x, y);
    if (SYM) {
      __syncthreads();
      mlp_block_mfma(L, xm, ym);
    }
    if (!net) return;
  } else {
    if (!net) return;
    mlp_eval(m.p, x, y, nullptr, nullptr, 0, false);
    if (SYM) mlp_eval(m.p, xm, ym, nullptr, nullptr, 0, false);
  }
  // U = Y R^p e^{k1 R} is dimensionless for any depth (R, k1 dimensional)
  const double u[3] = {(double)(y[0] * m.yscale[0]), (double)(y[1] * m.yscale[1]), (double)(y[2] * m.yscale[2])};
  double v[3];
  sur_regular_from(rr, zf, zp, f.wvn[1], u, v);
  vreg[nv * t] = (float)v[0];
  vreg[nv * t + 1] = (float)v[1];
  vreg[nv * t + 2] = (float)v[2];
  if (SYM) {
    const double um[3] = {(double)(ym[0] * m.yscale[0]), (double)(ym[1] * m.yscale[1]),
                          (double)(ym[2] * m.yscale[2])};
    sur_regular_from(rr, zp, zf, f.wvn[1], um, v);
    vreg[nv * t + 3] = (float)v[0];
    vreg[nv * t + 4] = (float)v[1];
    vreg[nv * t + 5] = (float)v[2];
  }
}

// exclusive prefix sums of the per-workgroup counts (one workgroup of 1024
// threads, each over a contiguous run); off[nblk] = total
__global__ __launch_bounds__(1024) void k_scan_counts(const unsigned* __restrict__ cnt, long long nblk,
                                                      unsigned long long* __restrict__ off) {
  __shared__ unsigned long long part[1024];
  const long long per = (nblk + 1023) / 1024;
  const long long b0 = threadIdx.x * per, b1 = b0 + per < nblk ? b0 + per : nblk;
  unsigned long long sum = 0;
  for (long long b = b0; b < b1; ++b) sum += cnt[b];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {         // inclusive scan of the 1024 run sums
    const unsigned long long v = threadIdx.x >= d ? part[threadIdx.x - d] : 0ull;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned long long run = part[threadIdx.x] - sum;
  for (long long b = b0; b < b1; ++b) {
    off[b] = run;
    run += cnt[b];
  }
  if (threadIdx.x == 1023) off[nblk] = part[1023];
}

// Stable partition of the segment by done: the series pairs first, then the
// network pairs, each in segment (key) order, so that every wave of the masked
// series kernel is (up to the one at the boundary) either all series or all
// given. Writes out (pair codes), mask and the regular parts of the given pairs
// in the partitioned order. Same workgroup decomposition as k_sur_regular.
__global__ __launch_bounds__(256) void k_partition_done(const uint32_t* __restrict__ pairs, long long np,
                                                        const uint8_t* __restrict__ done, const float* __restrict__ vreg,
                                                        const unsigned long long* __restrict__ off,
                                                        long long nblk, int nv, uint32_t* __restrict__ out,
                                                        uint8_t* __restrict__ mask, float* __restrict__ vout) {
  __shared__ unsigned wsum[4];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < np;
  const bool d = valid && done[t];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long bd = __ballot(d);
  if (lane == 0) wsum[w] = (unsigned)__popcll(bd);
  __syncthreads();
  unsigned before = (unsigned)__popcll(bd & ((1ull << lane) - 1ull));
  for (int k = 0; k < w; ++k) before += wsum[k];
  if (!valid) return;
  const unsigned long long ndone = off[nblk];
  const unsigned long long pd = off[blockIdx.x] + before;     // given pairs before t
  const unsigned long long dst = d ? (unsigned long long)np - ndone + pd : (unsigned long long)t - pd;
  out[dst] = pairs[t];
  mask[dst] = d ? 1 : 0;
  if (d)
    for (int c = 0; c < nv; ++c) vout[nv * dst + c] = vreg[nv * t + c];
}

// network and plain evanescent sums (unscaled, h = 1) at given points, with the
// error-map bound of each point's cell
__global__ __launch_bounds__(256) void k_eval_points(int n, const double* rh, const double* zs, const double* zd,
                                                     const double* kh, const double* wvn, SurDev m, double* tn,
                                                     double* tp, double* bound) {
  __shared__ SurLds L;
  if (m.quant) load_sur_lds(L.m, m.img);
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = i0 < n ? i0 : n - 1;
  float x[4], y[kSurOut];
  sur_inputs(rh[i], zs[i], zd[i], kh[i], x);
  if (m.quant) mlp_block_mfma(L, x, y);
  else mlp_eval(m.p, x, y, nullptr, nullptr, 0, false);
  if (i0 >= n) return;
  const double* wv = wvn + (long long)kNk * i;
  const double zf = 0.5 * (zs[i] + zd[i]), zp = 0.5 * (zs[i] - zd[i]);
  double tg, tr, tz, u[3], v[3];
  if (fin_regime(rh[i], 1.0) == kEigene) fin_eigene_evan(rh[i], zf, zp, wv, 1.0, tg, tr, tz);
  else fin_eigen_evan(rh[i], zf, zp, wv, 1.0, tg, tr, tz);
  sur_target(rh[i], zf, zp, wv[1], tg, tr, tz, u);
  sur_regular_from(rh[i], zf, zp, wv[1], u, v);
  for (int o = 0; o < kSurOut; ++o) tp[3 * i + o] = v[o];
  for (int o = 0; o < kSurOut; ++o) u[o] = (double)(y[o] * m.yscale[o]);
  sur_regular_from(rh[i], zf, zp, wv[1], u, v);
  for (int o = 0; o < kSurOut; ++o) tn[3 * i + o] = v[o];
  bound[i] = m.map[sur_cell(x)];
}

// ---- infinite-depth branch --------------------------------------------------------
// inputs: log KR, the phase of the wave terms (sin KR, cos KR: the Struve and
// Bessel parts oscillate in KR with period 2 pi, a dozen times over the upper
// decade of the domain, which a smooth function of log KR alone does not
// follow) and K(z+zeta)
__device__ __forceinline__ void sur_inputs_deep(double kr, double kv, float* x) {
  double sk, ck;
  sincos(kr, &sk, &ck);
  x[0] = (float)(2.0 * (log(kr) - log(kSurKrMin)) / (log(kSurKrMax) - log(kSurKrMin)) - 1.0);
  x[1] = (float)sk;
  x[2] = (float)ck;
  x[3] = (float)(2.0 * kv / -kSurKvMin + 1.0);
}
// targets: Re GF and Re GFh less their singular terms (third output unused,
// trained to 0); mag: the dimensionless sizes of the full components
__device__ __forceinline__ void sur_target_deep(double kr, double kv, double* u, double* mag) {
  double gr, gi, ghr, ghi, s0, s1;
  havelock(kr, kv, gr, gi, ghr, ghi);
  havelock_singular(kr, kv, s0, s1);
  u[0] = gr - s0;
  u[1] = ghr - s1;
  u[2] = 0.0;
  // sizes of the full dimensionless components: the value |GF| + 1/d, and the
  // gradient vector (the horizontal part GFh, the vertical part GF - 2/d of
  // INFGREEN3D, and the Rankine image's 1/d^2)
  const double d = sqrt(kr * kr + kv * kv);
  mag[0] = hypot(gr, gi) + 1.0 / d;
  mag[1] = hypot(ghr, ghi) + hypot(gr, gi) + 2.0 / d + 1.0 / (d * d);
  mag[2] = 0.0;
}

__global__ __launch_bounds__(256) void k_generate_deep(long long n, uint64_t seed, float* X, float* Y, float* M,
                                                       int* cell) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double kr = exp(log(kSurKrMin) + unit(seed, i, 0) * (log(kSurKrMax) - log(kSurKrMin)));
    // K (z + zeta) in [-20, 0], denser near the free surface
    const double u1 = unit(seed, i, 1);
    const double kv = kSurKvMin * u1 * u1;
    double u[3], mag[3];
    sur_target_deep(kr, kv, u, mag);
    for (int k = 0; k < 3; ++k) Y[3 * i + k] = (float)u[k];
    float x[4];
    sur_inputs_deep(kr, kv, x);
    for (int k = 0; k < 4; ++k) X[4 * i + k] = x[k];
    if (M) {
      for (int k = 0; k < 3; ++k) M[3 * i + k] = (float)mag[k];
      cell[i] = sur_cell(x);
    }
  }
}

// the 2-D grid of all (field point, panel) pairs in deep water, as
// k_assemble_grid<-1>, each pair in a cell whose bound is <= tol taking the
// Havelock function from the network (its imaginary parts and the singular
// terms of its real parts in closed form). The MFMA network needs the whole
// workgroup, so a workgroup runs it when any of its pairs takes it.
template <bool QUANT>
__global__ __launch_bounds__(256) void k_sur_assemble_deep(GeomDev g, FreqDev f, SurDev m, float tol, double* A,
                                                           double* S, long long ld, unsigned long long* count) {
  __shared__ SurLds L;
  if (QUANT) load_sur_lds(L.m, m.img);
  __syncthreads();
  const int i0 = blockIdx.y * 16 + (threadIdx.x >> 4), j0 = blockIdx.x * 16 + (threadIdx.x & 15);
  const bool valid = i0 < g.nt && j0 < g.n;
  const int i = valid ? i0 : 0, j = valid ? j0 : 0;
  double xp[3];
  field_point(g, i, j, xp);
  const double* xq = g.xyz_p + 3 * (long long)j;
  const double rr = horiz_dist(xq[0], xp[0], xq[1], xp[1]);
  const double kr = f.v * rr, kv = f.v * (xp[2] + xq[2]);
  const bool limit = fabs(f.v) < F32(1.e-8) || fabs(f.v + 1.0) < F32(1.e-8);
  const bool inside = !limit && kr >= kSurKrMin && kr <= kSurKrMax && kv >= kSurKvMin && kv <= 0.0;
  float x[4], y[kSurOut];
  sur_inputs_deep(fmin(fmax(kr, kSurKrMin), kSurKrMax), fmin(fmax(kv, kSurKvMin), 0.0), x);
  const bool take = valid && inside && m.map[sur_cell(x)] <= tol;
  const int ntake = __syncthreads_count(take);
  if (ntake > 0) {
    if (QUANT) mlp_block_mfma(L, x, y);
    else mlp_eval(m.p, x, y, nullptr, nullptr, 0, false);
  }
  if (threadIdx.x == 0 && ntake > 0) atomicAdd(count, (unsigned long long)ntake);
  if (!valid) return;
  const int flag = near_flag(xp, xq, g.pnsz[j]);
  Grn4 gr;
  if (take) {
    double s0, s1, gi, ghi;
    havelock_singular(kr, kv, s0, s1);
    havelock_imag(kr, kv, gi, ghi);
    const double gfr = (double)(y[0] * m.yscale[0]) + s0, ghr = (double)(y[1] * m.yscale[1]) + s1;
    infgreen3d_given(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gfr, gi, ghr, ghi, gr);
  } else {
    infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], f.v, flag, gr);
  }
  write_pair_g(g, i, j, xp, flag, gr, A, S, ld);
}

__global__ __launch_bounds__(256) void k_eval_points_deep(int n, const double* kr, const double* kv, SurDev m,
                                                          double* net, double* plain, double* bound) {
  __shared__ SurLds L;
  if (m.quant) load_sur_lds(L.m, m.img);
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = i0 < n ? i0 : n - 1;
  float x[4], y[kSurOut];
  sur_inputs_deep(kr[i], kv[i], x);
  if (m.quant) mlp_block_mfma(L, x, y);
  else mlp_eval(m.p, x, y, nullptr, nullptr, 0, false);
  if (i0 >= n) return;
  double gr, gi, ghr, ghi, s0, s1;
  havelock(kr[i], kv[i], gr, gi, ghr, ghi);
  havelock_singular(kr[i], kv[i], s0, s1);
  plain[2 * i] = gr;
  plain[2 * i + 1] = ghr;
  net[2 * i] = (double)(y[0] * m.yscale[0]) + s0;
  net[2 * i + 1] = (double)(y[1] * m.yscale[1]) + s1;
  bound[i] = m.map[sur_cell(x)];
}

}  // namespace

hipError_t sur_generate_deep(hipStream_t s, long long n, uint64_t seed, float* X, float* Y, float* M, int* cell) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_generate_deep, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, n, seed, X,
                     Y, M, cell);
  return hipGetLastError();
}

hipError_t sur_assemble_deep(hipStream_t s, const GeomDev& g, const FreqDev& f, SurDev m, float tol, double* A,
                             double* S, long long ld, unsigned long long* count) {
  if (m.quant && !m.img) return hipErrorInvalidValue;
  const dim3 grid((g.n + 15) / 16, (g.nt + 15) / 16);
  if (m.quant) hipLaunchKernelGGL((k_sur_assemble_deep<true>), grid, dim3(256), 0, s, g, f, m, tol, A, S, ld, count);
  else hipLaunchKernelGGL((k_sur_assemble_deep<false>), grid, dim3(256), 0, s, g, f, m, tol, A, S, ld, count);
  return hipGetLastError();
}

hipError_t sur_eval_points_deep(hipStream_t s, int n, const double* kr, const double* kv, SurDev m, double* net,
                                double* plain, double* bound) {
  if (n <= 0) return hipSuccess;
  if (m.quant && !m.img) return hipErrorInvalidValue;
  double* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(double) * 7 * (size_t)n);
  if (e != hipSuccess) return e;
  double *dkr = d, *dkv = d + n, *dn = d + 2 * (size_t)n, *dp = d + 4 * (size_t)n, *db = d + 6 * (size_t)n;
  const size_t nb = sizeof(double) * n;
  if ((e = hipMemcpyAsync(dkr, kr, nb, hipMemcpyHostToDevice, s)) == hipSuccess &&
      (e = hipMemcpyAsync(dkv, kv, nb, hipMemcpyHostToDevice, s)) == hipSuccess) {
    hipLaunchKernelGGL(k_eval_points_deep, dim3((n + 255) / 256), dim3(256), 0, s, n, dkr, dkv, m, dn, dp, db);
    if ((e = hipGetLastError()) == hipSuccess && (e = hipMemcpyAsync(net, dn, 2 * nb, hipMemcpyDeviceToHost, s)) == hipSuccess &&
        (e = hipMemcpyAsync(plain, dp, 2 * nb, hipMemcpyDeviceToHost, s)) == hipSuccess &&
        (e = hipMemcpyAsync(bound, db, nb, hipMemcpyDeviceToHost, s)) == hipSuccess)
      e = hipStreamSynchronize(s);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return e;
}

hipError_t sur_generate(hipStream_t s, long long n, uint64_t seed, int nkh, const double* kh, const double* wvn,
                        float* X, float* Y, float* M, int* cell) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, n, seed, nkh, kh,
                     wvn, X, Y, M, cell);
  return hipGetLastError();
}
hipError_t sur_round(hipStream_t s, const float* p, float* q, void* img) {
  hipLaunchKernelGGL(k_round, dim3((kSurParams + 255) / 256), dim3(256), 0, s, p, q);
  if (img) hipLaunchKernelGGL(k_build_img, dim3(1), dim3(256), 0, s, q, static_cast<SurImg*>(img));
  return hipGetLastError();
}
hipError_t sur_forward(hipStream_t s, int B, const float* X, const float* q, float* Z, float* A, float* Yh, bool quant) {
  if (B <= 0) return hipSuccess;
  const unsigned gx = (unsigned)((B + 255) / 256);
  for (int l = 0; l < 5; ++l)
    hipLaunchKernelGGL(k_fwd_layer, dim3(gx, l < 4 ? kSurH : kSurOut), dim3(256), 0, s, B, l, X, q, Z, A, Yh,
                       quant ? 1 : 0);
  return hipGetLastError();
}
size_t sur_grad_work_bytes(int B) { return sizeof(double) * (size_t)kSurParams * (size_t)((B + kSurChunk - 1) / kSurChunk); }
hipError_t sur_loss_grad(hipStream_t s, int B, const float* X, const float* Y, const float* yscale, const float* q,
                         const float* Z, const float* A, const float* Yh, float* D, float* grad, double* loss,
                         double* work) {
  if (B <= 0) return hipSuccess;
  const unsigned gx = (unsigned)((B + 255) / 256);
  const int nchunk = (B + kSurChunk - 1) / kSurChunk;
  hipLaunchKernelGGL(k_out_delta, dim3(gx), dim3(256), 0, s, B, Y, yscale, Yh, D, loss);
  for (int l = 3; l >= 0; --l) hipLaunchKernelGGL(k_bwd_layer, dim3(gx, kSurH), dim3(256), 0, s, B, l, q, Z, D);
  hipLaunchKernelGGL(k_wgrad_hidden, dim3(nchunk, 3), dim3(256), 0, s, B, A, D, work);
  hipLaunchKernelGGL(k_wgrad, dim3((kSurParams + 255) / 256, nchunk), dim3(256), 0, s, B, X, A, D, work);
  hipLaunchKernelGGL(k_wgrad_sum, dim3((kSurParams + 255) / 256), dim3(256), 0, s, nchunk, work, grad);
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
  if (n <= 0) return hipSuccess;
  if (m.quant && !m.img) return hipErrorInvalidValue;   // the MFMA kernels copy the image into LDS
  hipLaunchKernelGGL(k_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, X, m, U);
  return hipGetLastError();
}
size_t sur_regular_bytes(long long np, bool sym) {
  const long long nblk = (np + 255) / 256;
  const size_t nv = sym ? 6 : 3;
  return (8 * nv + 6) * (size_t)np + 4 * (size_t)nblk + 8 * (size_t)(nblk + 1) + 64;
}

hipError_t sur_regular_parts(hipStream_t s, const GeomDev& g, const FreqDev& f, double kh, const uint32_t* pairs,
                             long long np, SurDev m, float tol, void* work, unsigned long long* count,
                             SurPartition& part, bool sym) {
  part = SurPartition{pairs, nullptr, nullptr};
  if (np <= 0) return hipSuccess;
  if (m.quant && !m.img) return hipErrorInvalidValue;   // the MFMA kernels copy the image into LDS
  const long long nblk = (np + 255) / 256;
  const int nv = sym ? 6 : 3;
  // work: vout (4 nv np) | vreg (4 nv np) | out (4 np) | done (np) | mask (np) | off (8 (nblk + 1)) | cnt (4 nblk)
  char* w = static_cast<char*>(work);
  float* vout = reinterpret_cast<float*>(w);
  float* vreg = vout + (long long)nv * np;
  uint32_t* out = reinterpret_cast<uint32_t*>(vreg + (long long)nv * np);
  uint8_t* done = reinterpret_cast<uint8_t*>(out + np);
  uint8_t* mask = done + np;
  unsigned long long* off = reinterpret_cast<unsigned long long*>(
      (reinterpret_cast<uintptr_t>(mask + np) + 7) & ~static_cast<uintptr_t>(7));
  unsigned* cnt = reinterpret_cast<unsigned*>(off + nblk + 1);
  const dim3 gd((unsigned)nblk), bd(256);
  if (m.quant && sym)
    hipLaunchKernelGGL((k_sur_regular<true, true>), gd, bd, 0, s, g, f, kh, pairs, np, m, tol, done, vreg, cnt, count);
  else if (m.quant)
    hipLaunchKernelGGL((k_sur_regular<true, false>), gd, bd, 0, s, g, f, kh, pairs, np, m, tol, done, vreg, cnt, count);
  else if (sym)
    hipLaunchKernelGGL((k_sur_regular<false, true>), gd, bd, 0, s, g, f, kh, pairs, np, m, tol, done, vreg, cnt, count);
  else
    hipLaunchKernelGGL((k_sur_regular<false, false>), gd, bd, 0, s, g, f, kh, pairs, np, m, tol, done, vreg, cnt, count);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cnt, nblk, off);
  hipLaunchKernelGGL(k_partition_done, gd, bd, 0, s, pairs, np, done, vreg, off, nblk, nv, out, mask, vout);
  part = SurPartition{out, mask, vout};
  return hipGetLastError();
}

hipError_t sur_eval_points(hipStream_t s, int n, const double* rh, const double* zs, const double* zd, const double* kh,
                           const double* wvn, double g_disp, SurDev m, double* tn, double* tp, double* bound) {
  (void)g_disp;
  if (n <= 0) return hipSuccess;
  if (m.quant && !m.img) return hipErrorInvalidValue;   // the MFMA kernels copy the image into LDS
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
