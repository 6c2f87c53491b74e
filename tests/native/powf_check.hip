// Accuracy of the device powf(x, n) that s_corr uses for exponents outside
// {2, 3, 4} (pow_ratio, pbf_device.h): over every float x in [0, 1.5] (the
// ratio W(r)/W(dq) is at most W(0)/W(dq) = 1/(1 - q^2)^3, 1.13 at q = 0.2),
// the distance in ulps from the double-precision power rounded to float.
// Prints the largest distance and how many x are not the correctly rounded
// value; exits 1 when the largest distance exceeds 1 ulp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

__global__ void k_check(unsigned hi, float n, unsigned* max_ulp, unsigned long long* inexact) {
  const unsigned stride = gridDim.x * blockDim.x;
  unsigned worst = 0;
  unsigned long long off = 0;
  for (unsigned b = blockIdx.x * blockDim.x + threadIdx.x; b <= hi; b += stride) {
    const float x = __uint_as_float(b);
    const float p = powf(x, n);
    const float d = static_cast<float>(pow(static_cast<double>(x), static_cast<double>(n)));
    const int u = static_cast<int>(__float_as_uint(p)) - static_cast<int>(__float_as_uint(d));
    const unsigned a = static_cast<unsigned>(u < 0 ? -u : u);
    worst = a > worst ? a : worst;
    off += a != 0;
  }
  atomicMax(max_ulp, worst);
  if (off) atomicAdd(inexact, off);
}

int main(int argc, char** argv) {
  const float n = argc > 1 ? static_cast<float>(std::atof(argv[1])) : 5.0f;
  unsigned* mx = nullptr;
  unsigned long long* cnt = nullptr;
  if (hipMalloc(&mx, sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&cnt, sizeof(unsigned long long)) != hipSuccess) {
    std::printf("hipMalloc failed\n");
    return 2;
  }
  hipMemset(mx, 0, sizeof(unsigned));
  hipMemset(cnt, 0, sizeof(unsigned long long));
  const unsigned hi = 0x3fc00000u;  // 1.5f
  k_check<<<4096, 256>>>(hi, n, mx, cnt);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::printf("kernel failed\n");
    return 2;
  }
  unsigned m = 0;
  unsigned long long c = 0;
  hipMemcpy(&m, mx, sizeof(m), hipMemcpyDeviceToHost);
  hipMemcpy(&c, cnt, sizeof(c), hipMemcpyDeviceToHost);
  std::printf("device powf(x, %g), x in [0, 1.5]: max %u ulp, %llu of %u not correctly rounded\n",
              n, m, c, hi + 1);
  return m > 1 ? 1 : 0;
}
