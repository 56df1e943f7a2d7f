// hm_surrogate.hpp -- learned surrogate of the finite-depth evanescent-mode sum.
//
// For R/h >= 0.05 the reference's FINGREEN3D (EIGENE / EIGEN,
// FinGreen3D.f90:281-506) is
//   wave part = propagating mode (J0, Y0, J1, Y1, cosh)  +  evanescent sum  -  (1/r + 1/r1)
// The propagating mode carries all of the imaginary part and is cheap and
// exact in FP64, as is the Rankine strip. The evanescent sum (up to 50 K0/K1
// terms and three Wynn epsilon limits) is the expensive part; it is a smooth
// function of four nondimensional inputs
//   R/h, (z+zeta)/h, (z-zeta)/h, K h          (h = depth, K = omega^2 / g)
// (the eigenfunctions cos k(z+h) cos k(zeta+h) depend on z+zeta and z-zeta).
// The evanescent sums T carry the 1/r singularity that the strip S = -(1/r +
// 1/r1) (and its R-, z-derivatives) removes. The network predicts, per
// component, Y = T + S e^{-k1 R} (k1 = first evanescent wave number): near
// R = 0 this is the smooth regular part T + S, for large R both terms decay
// like e^{-k1 R}. A 4-64-64-64-64-3 SiLU network predicts the dimensionless
// U = Y R^p e^{k1 R} (p = 1, 2, 2), O(1) over the domain; the assembly adds
// V = U e^{-k1 R} / R^p + S (1 - e^{-k1 R}) to the exact propagating mode.
//
// Precision: the first layer (4 -> 64) runs in FP32 (BF16 inputs would quantize
// log(R/h) to 1 %); the hidden and output layers take BF16 weights and BF16
// activations with FP32 accumulation, as the inference kernel computes them.
// Training is quantization-aware (forward with the BF16-rounded copy, FP32
// master weights, Adam).
//
// Error map: after training, the network is compared with the plain series on
// an independent validation set; the largest relative error of the Green
// function value and of its gradient vector (each against the size of the full
// component: propagating + evanescent + Rankine) is recorded per cell of a 24 x 8 x 8 x 24 grid over the four inputs. The
// assembly uses the network only in cells whose bound is below the tolerance
// and falls back to the plain series elsewhere. The domain is the EIGENE
// regime, R/h in [0.05, 0.5], with z/h, zeta/h in [-0.25, 0] (the draft range
// of surface-piercing hulls) and Kh in [0.01, 50]; R/h < 0.05 (PIDCOCK,
// LINTON), R/h >= 0.5 (EIGEN, at most 9 terms, cheaper than a network
// evaluation) and deeper points always take the plain series.
//
// Infinite-depth branch: the wave part of INFGREEN3D (InfGreen_Appr.f90:104-183)
// is the Havelock function GF(KR, K(z+zeta)) and its KR-derivative GFh
// (HavelockGF :185). Their imaginary parts are closed forms (-2 pi e^v J0,
// 2 pi e^v J1); a second network of the same architecture predicts the real
// parts less their singular terms (havelock_singular: the log(d - v) of L0 and
// the (beta + h)/(d - v) of Ls), from (log KR, sin KR, cos KR, K(z+zeta)) over
// KR in [1e-3, 50], K(z+zeta) in [-20, 0] (the phase inputs let the network
// follow the oscillation of the Struve and Bessel terms in KR; the error map's
// cells split on all four inputs as the finite-depth map does). Its error map is
// against |GF| + 1/d and |GFh| + |GF| + 2/d + 1/d^2 (d = K R1), the
// dimensionless sizes of the full value and gradient vector. The deep-water assembly (sur_assemble_deep) takes the
// network for the pairs in cells whose bound is <= the tolerance.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hm_assembly.hpp"

namespace hm {

constexpr int kSurIn = 4, kSurH = 64, kSurOut = 3;
constexpr int kSurOffW0 = 0;
constexpr int kSurOffB0 = kSurOffW0 + kSurH * kSurIn;
constexpr int kSurOffW1 = kSurOffB0 + kSurH;                 // hidden layer l = 1..3: W (64x64), b (64)
constexpr int kSurHiddenStride = kSurH * kSurH + kSurH;
constexpr int kSurOffW4 = kSurOffW1 + 3 * kSurHiddenStride;
constexpr int kSurOffB4 = kSurOffW4 + kSurOut * kSurH;
constexpr int kSurParams = kSurOffB4 + kSurOut;              // 12995
constexpr int kMapR = 24, kMapS = 8, kMapD = 8, kMapK = 24;
constexpr int kSurMap = kMapR * kMapS * kMapD * kMapK;       // 36864 cells
constexpr int kSurBlob = kSurParams + kSurOut + kSurMap;     // params, output scales, error map
constexpr double kSurRhMin = 0.05, kSurRhMax = 0.5, kSurKhMin = 0.01, kSurKhMax = 50.0;
constexpr double kSurDepth = 0.25;   // z/h and zeta/h in [-0.25, 0]: the draft range of surface-piercing hulls
constexpr double kSurKrMin = 1e-3, kSurKrMax = 50.0, kSurKvMin = -20.0;   // infinite-depth branch

// LDS image of the BF16 network for the MFMA inference kernels (hidden weights
// as BF16 bits, the FP32 first and output layers and the biases), built on the
// device whenever the parameters change, copied into LDS by each workgroup
constexpr size_t kSurImgBytes = 27408;
// byte offset of the image in the surrogate buffer (after p, q, yscale, map)
constexpr size_t kSurImgOffset = (((size_t)(2 * kSurParams + kSurOut + kSurMap) * 4) + 255) & ~(size_t)255;
constexpr size_t kSurBufBytes = kSurImgOffset + kSurImgBytes;

// a trained model on the device: FP32 parameters, the BF16-rounded copy the
// forward pass uses (as floats), output scales, the error map and the image
struct SurDev {
  const float* p;     // kSurParams
  const float* q;     // kSurParams, BF16-rounded
  const float* yscale;
  const float* map;
  int quant;          // 1: BF16 weights and activations; 0: FP32 network
  const void* img;    // kSurImgBytes
};

struct SurTrainStats {
  double train_loss, val_rmse, val_max_err, seconds;
  long long samples;
  int epochs, steps;
};

// training and validation data from the plain kernels (h = 1)
hipError_t sur_generate(hipStream_t s, long long n, uint64_t seed, int nkh, const double* kh, const double* wvn,
                        float* X, float* Y, float* M, int* cell);
// the infinite-depth branch's data from the plain Havelock function
hipError_t sur_generate_deep(hipStream_t s, long long n, uint64_t seed, float* X, float* Y, float* M, int* cell);
// deep-water assembly of all pairs (as assemble_grid with deep = true), the
// network taking the Havelock function of pairs in cells whose bound is <= tol
// (count += pairs taken)
hipError_t sur_assemble_deep(hipStream_t s, const GeomDev& g, const FreqDev& f, SurDev m, float tol, double* A,
                             double* S, long long ld, unsigned long long* count);
// network and plain Re GF, Re GFh at host points (kr, kv), with each point's cell bound
hipError_t sur_eval_points_deep(hipStream_t s, int n, const double* kr, const double* kv, SurDev m, double* net,
                                double* plain, double* bound);
// q = BF16-rounded p; img (may be null) = the LDS image of q
hipError_t sur_round(hipStream_t s, const float* p, float* q, void* img = nullptr);
hipError_t sur_forward(hipStream_t s, int B, const float* X, const float* q, float* Z, float* A, float* Yh,
                       bool quant = true);
// loss (accumulated into *loss) and parameter gradient of one batch; work:
// sur_grad_work_bytes(B) bytes of per-chunk partial sums
size_t sur_grad_work_bytes(int B);
hipError_t sur_loss_grad(hipStream_t s, int B, const float* X, const float* Y, const float* yscale, const float* q,
                         const float* Z, const float* A, const float* Yh, float* D, float* grad, double* loss,
                         double* work);
hipError_t sur_adam(hipStream_t s, float* p, const float* grad, float* m, float* v, float lr, int t);
hipError_t sur_error_map(hipStream_t s, long long n, const float* Y, const float* M, const int* cell,
                         const float* Yh, unsigned* map_bits, double* sq_err);
// network evaluation of scaled outputs U for normalized inputs (n x 4)
hipError_t sur_predict(hipStream_t s, long long n, const float* X, SurDev m, float* U);

// network and plain evanescent sums T (unscaled, h = 1) at host points, with
// the cell bound of each point (test and ablation entry)
hipError_t sur_eval_points(hipStream_t s, int n, const double* rh, const double* zs, const double* zd, const double* kh,
                           const double* wvn, double g_disp, SurDev m, double* tn, double* tp, double* bound);

// the network over the pairs of the EIGENE list segment: a pair inside the
// trained range whose cell bound is <= tol is taken by the network, which gives
// its regular part V = T - (1/r + 1/r1) (value, R- and z-derivative; T the
// evanescent sum). The segment comes back partitioned (stable, by key order
// within each part): part.pairs = the pairs left to the series, then those
// taken; part.mask[t] = 1 and part.vreg[nv t .. nv t + 2] = V for a taken pair
// (nv = 3). sym: each pair also stands for its mirror (depths swapped, for
// assemble_sym): taken only if both directions are, nv = 6 with the mirror's V
// in vreg[nv t + 3 ..]. The masked EIGENE assembly (assemble_regime /
// assemble_sym with mask and vreg) writes A and S. count += pairs taken
// (mirrors counted). work: sur_regular_bytes(np, sym) bytes.
struct SurPartition {
  const uint32_t* pairs;
  const uint8_t* mask;
  const float* vreg;
};
size_t sur_regular_bytes(long long np, bool sym = false);
hipError_t sur_regular_parts(hipStream_t s, const GeomDev& g, const FreqDev& f, double kh, const uint32_t* pairs,
                             long long np, SurDev m, float tol, void* work, unsigned long long* count,
                             SurPartition& part, bool sym = false);

}  // namespace hm
