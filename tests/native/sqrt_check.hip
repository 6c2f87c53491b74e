// Exhaustive check of sqrt_rn (pbf_device.h) against sqrtf, over every
// non-negative finite float, on the device; exits 1 on any difference.  The
// count of differences from the double-precision root rounded to float is
// printed beside it (that root is the correctly rounded one).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "pbf_device.h"

__global__ void k_check(unsigned lo, unsigned hi, unsigned long long* bad_f,
                        unsigned long long* bad_d, unsigned* first) {
  const unsigned stride = gridDim.x * blockDim.x;
  unsigned long long nf = 0, nd = 0;
  // hi + stride stays below 2^32: b never wraps
  for (unsigned b = lo + blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float(b);
    const float r = fluidhip::sqrt_rn(x);
    const float f = sqrtf(x);
    const float d = static_cast<float>(sqrt(static_cast<double>(x)));
    if (__float_as_uint(r) != __float_as_uint(f)) {
      ++nf;
      atomicMin(first, b);
    }
    if (__float_as_uint(r) != __float_as_uint(d)) ++nd;
  }
  if (nf) atomicAdd(bad_f, nf);
  if (nd) atomicAdd(bad_d, nd);
}

int main() {
  unsigned long long* bad = nullptr;
  unsigned* first = nullptr;
  if (hipMalloc(&bad, 2 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&first, sizeof(unsigned)) != hipSuccess) {
    std::printf("hipMalloc failed\n");
    return 2;
  }
  hipMemset(bad, 0, 2 * sizeof(unsigned long long));
  const unsigned none = 0xffffffffu;
  hipMemcpy(first, &none, sizeof(unsigned), hipMemcpyHostToDevice);
  // every non-negative float below +inf: bit patterns [0, 0x7f800000)
  k_check<<<4096, 256>>>(0u, 0x7f800000u, bad, bad + 1, first);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::printf("kernel failed\n");
    return 2;
  }
  unsigned long long h[2];
  unsigned f = 0;
  hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost);
  hipMemcpy(&f, first, sizeof(f), hipMemcpyDeviceToHost);
  std::printf("checked %u floats: %llu differ from sqrtf, %llu from the double root; first 0x%08x\n",
              0x7f800000u, h[0], h[1], f);
  return h[0] ? 1 : 0;
}
