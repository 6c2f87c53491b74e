// CPU check of fluidhip::add_repeat against k plain float additions, bit for
// bit, on random and constructed cases (ties at every binade, both signs,
// zero crossings, subnormals, values near overflow, long runs).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../fluidsimulator_amd/csrc/exact_sum.h"

static float naive(float x, float w, int k) {
  volatile float s = x;
  for (int i = 0; i < k; ++i) s = s + w;
  return s;
}

static uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 2000000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long fails = 0;
  auto rand_float = [&](int emin, int emax) {
    const double m = 1.0 + U(rng);
    const int e = emin + static_cast<int>(U(rng) * (emax - emin + 1));
    float f = static_cast<float>(std::ldexp(m, e));
    return U(rng) < 0.5 ? -f : f;
  };
  for (long c = 0; c < cases; ++c) {
    float x, w;
    int k;
    const int kind = static_cast<int>(c % 8);
    if (kind == 0) {  // accumulation from zero, as in a neighbor sum
      x = 0.0f;
      w = rand_float(-20, 20);
      k = 1 + static_cast<int>(U(rng) * 100000);
    } else if (kind == 1) {  // ties: w an odd multiple of half the spacing of x
      x = std::fabs(rand_float(-10, 10));
      int e;
      (void)std::frexp(x, &e);
      const double u = std::ldexp(1.0, e - 24);
      const int q = static_cast<int>(U(rng) * 64);
      w = static_cast<float>((q + 0.5) * u);
      if (U(rng) < 0.5) { x = -x; }
      if (U(rng) < 0.3) { w = -w; }
      k = 1 + static_cast<int>(U(rng) * 5000);
    } else if (kind == 2) {  // mixed signs, crossing zero
      x = rand_float(-5, 5);
      w = -x / (1.0f + static_cast<float>(U(rng) * 300));
      k = 1 + static_cast<int>(U(rng) * 2000);
    } else if (kind == 3) {  // tiny and subnormal values
      x = rand_float(-149, -120);
      w = rand_float(-149, -118);
      k = 1 + static_cast<int>(U(rng) * 3000);
    } else if (kind == 4) {  // near overflow
      x = rand_float(120, 127);
      w = std::fabs(rand_float(100, 126));
      if (x < 0) w = -w;
      k = 1 + static_cast<int>(U(rng) * 3000);
    } else if (kind == 5) {  // w far below the spacing of x (absorbed) and near half of it
      x = rand_float(-10, 10);
      int e;
      (void)std::frexp(std::fabs(x), &e);
      w = static_cast<float>(std::ldexp(U(rng), e - 24 - static_cast<int>(U(rng) * 3)));
      if (U(rng) < 0.5) w = -w;
      k = 1 + static_cast<int>(U(rng) * 100000);
    } else {  // anything
      x = rand_float(-30, 30);
      w = rand_float(-30, 30);
      k = 1 + static_cast<int>(U(rng) * 20000);
    }
    const float a = fluidhip::add_repeat(x, w, k);
    const float b = naive(x, w, k);
    if (bits(a) != bits(b)) {
      if (fails < 10)
        std::printf("MISMATCH x=%a w=%a k=%d fast=%a plain=%a\n", x, w, k, a, b);
      ++fails;
    }
    // the in-place wrapper, on both sides of its cut-over
    const int ks = 1 + (k % (2 * fluidhip::kAddInPlace));
    if (bits(fluidhip::add_k(x, w, ks)) != bits(naive(x, w, ks)) ||
        bits(fluidhip::add_k(x, w, k)) != bits(b)) {
      if (fails < 10) std::printf("MISMATCH add_k x=%a w=%a k=%d\n", x, w, k);
      ++fails;
    }
  }
  // special values
  const float sp[] = {0.0f, -0.0f, INFINITY, -INFINITY, NAN, 1.0f, -1.0f, 0x1p-149f, 0x1.fffffep127f};
  for (float x : sp)
    for (float w : sp)
      for (int k : {0, 1, 2, 7, 1000}) {
        const float a = fluidhip::add_repeat(x, w, k);
        const float b = naive(x, w, k);
        const bool same = bits(a) == bits(b) || (std::isnan(a) && std::isnan(b));
        if (!same) {
          if (fails < 20) std::printf("MISMATCH special x=%a w=%a k=%d\n", x, w, k);
          ++fails;
        }
      }
  std::printf("%ld cases, %ld mismatches\n", cases, fails);
  return fails == 0 ? 0 : 1;
}
