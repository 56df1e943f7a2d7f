// hm_noalloc.hpp -- the Green function batch without device allocation
// (hm_noalloc.hip): inputs and outputs staged in a fixed device array.
#pragma once
#include <hip/hip_runtime.h>

namespace hm {

// largest batch of one call, and the device memory each GPU holds for it
// (6 coordinate arrays, flags, 4 complex outputs per pair, the 200 roots)
constexpr int kNoAllocMaxPoints = 262144;
constexpr size_t kNoAllocBytes = 30410496;

// the device address of the array for the current device (once per device)
hipError_t noalloc_setup();
// hm_fingreen3d_batch / hm_infgreen3d_batch's work on that array: in = xf, xp,
// yf, yp, zf, zp (host, n each), wvn (host, 200, finite depth), flag (host),
// grn (host, 8 n doubles); returns after grn is written
hipError_t noalloc_green(hipStream_t s, int n, bool deep, const double* const* in, double v, const double* wvn,
                         double h, const int* flag, double* grn);

}  // namespace hm
