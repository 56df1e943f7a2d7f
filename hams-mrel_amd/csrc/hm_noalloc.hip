// hm_noalloc.hip -- the finite-depth Green function batch of hm_fingreen3d_batch
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// hm_fingreen3d_batch stages its inputs (six coordinate arrays, the flags and
// the 200 wave numbers) and its output (4 complex values per pair) in a device
// buffer it allocates per call. This variant stages them in `scratch`, a
// file-scope device array of fixed size that every GPU the library runs on
// holds until the process ends: at most kNoAllocMaxPoints pairs per call. The
// host never uses the array's name as an address: its device address comes
// from hipGetSymbolAddress, once per device (at context creation, with the
// context's device current), and is kept per device. Every call on one GPU
// shares that GPU's array, so calls that overlap on one GPU (two contexts, two
// host threads) race; a call returns only after its output is back on the host.
// Nothing needs clearing between calls: each call writes every output it reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hm_assembly.hpp"
#include "hm_noalloc.hpp"

namespace {

// fixed layout for the largest call: 6 coordinate arrays, flags, output, roots
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kCoordBytes = align_up(sizeof(double) * (size_t)hm::kNoAllocMaxPoints);
constexpr size_t kFlagOff = 6 * kCoordBytes;
constexpr size_t kOutOff = kFlagOff + align_up(sizeof(int) * (size_t)hm::kNoAllocMaxPoints);
constexpr size_t kWvnOff = kOutOff + align_up(8 * sizeof(double) * (size_t)hm::kNoAllocMaxPoints);
constexpr size_t kScratchBytes = kWvnOff + align_up(sizeof(double) * (size_t)hm::kNk);
static_assert(kScratchBytes == hm::kNoAllocBytes, "scratch layout and the stated size differ");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {nullptr};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace hm {

hipError_t noalloc_setup() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[dev] = p;
  return hipSuccess;
}

hipError_t noalloc_green(hipStream_t s, int n, bool deep, const double* const* in, double v, const double* wvn,
                         double h, const int* flag, double* grn) {
  if (n < 0 || n > kNoAllocMaxPoints) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  void* base_v = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base_v = g_ptr[dev];
  }
  if (!base_v) {
    // a device that no context was created on: take its address now
    if ((e = noalloc_setup()) != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    base_v = g_ptr[dev];
  }
  char* base = static_cast<char*>(base_v);
  const size_t sz = sizeof(double) * (size_t)n;
  double* d[6];
  for (int k = 0; k < 6; ++k) {
    d[k] = reinterpret_cast<double*>(base + k * kCoordBytes);
    if ((e = hipMemcpyAsync(d[k], in[k], sz, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  }
  int* df = reinterpret_cast<int*>(base + kFlagOff);
  double* dout = reinterpret_cast<double*>(base + kOutOff);
  double* dw = reinterpret_cast<double*>(base + kWvnOff);
  if ((e = hipMemcpyAsync(df, flag, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if (!deep && (e = hipMemcpyAsync(dw, wvn, sizeof(double) * kNk, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = green_batch(s, n, deep, d[0], d[1], d[2], d[3], d[4], d[5], v, dw, h, df, dout)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(grn, dout, 8 * sz, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

}  // namespace hm
