// Exact shortcut for a run of identical float additions.
//
// The PBF passes add, for every neighbor, one term to a float accumulator, in
// list order.  Where many consecutive neighbors contribute the same term
// (particles piled up at one point), the result of those k additions,
//   s_k = fl(fl(...fl(s_0 + w) + w...) + w)      (round to nearest even),
// is computed here without doing k dependent additions: while the sum stays
// inside one binade [2^(e-1), 2^e) its grid spacing u is fixed, and each
// addition moves it by the same whole number of u (w/u rounded to nearest;
// when w/u is exactly halfway, the first addition lands on an even multiple
// and every later one by q + (q odd), q = floor(w/u)).  Every binade costs a
// few double-precision operations, all exact (integers below 2^26 times a
// power of two).  Wherever those premises do not hold (the next addition
// leaves the binade, tiny or non-finite values) one ordinary addition is
// made, so the result equals the k plain additions bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace fluidhip {

// Not inlined: callers flush runs rarely, and inlining its double-precision
// temporaries into long gather loops costs registers there.
__host__ __device__ inline __attribute__((noinline)) float add_repeat(float x, float w, int k) {
  if (k <= 0) return x;
  // a zero, infinite or NaN term: one addition fixes the result for good
  if (w == 0.0f || !std::isfinite(w) || !std::isfinite(x)) return x + w;
  // fl(-a) = -fl(a) for every nonzero result: accumulate with w > 0.  A
  // single addition is made in the caller's signs, since an exact
  // cancellation gives +0 whatever the signs of its operands.
  const bool neg = w < 0.0f;
  if (neg) {
    x = -x;
    w = -w;
  }
  auto plain = [neg, w](float v) { return neg ? -((-v) + (-w)) : v + w; };
  // parity of an integer-valued double (exact below 2^53)
  auto odd = [](double v) { return v - 2.0 * std::floor(0.5 * v) != 0.0; };
  const double kTop = 16777216.0;    // 2^24: one past the largest multiple of u in the binade
  const double kBottom = 8388608.0;  // 2^23: the binade's lowest multiple of u
  while (k > 0) {
    if (!std::isfinite(x)) break;  // overflowed: later additions keep it
    const float ax = std::fabs(x);
    if (ax < 0x1p-125f) {  // zero, subnormal or first binade: plain addition
      x = plain(x);
      --k;
      continue;
    }
    // ax is normal: ax in [2^(e-1), 2^e) with e from its exponent bits
    uint32_t ab;
    std::memcpy(&ab, &ax, sizeof(ab));
    const int e = static_cast<int>((ab >> 23) & 0xffu) - 126;
    // in units of u = 2^(e-24), the binade's spacing (power-of-two scalings
    // are exact)
    const double M = std::ldexp(static_cast<double>(ax), 24 - e);  // integer in [2^23, 2^24)
    const double W = std::ldexp(static_cast<double>(w), 24 - e);
    const bool up = x > 0.0f;  // |x| grows; otherwise it shrinks toward zero
    if (up ? !(M + W < kTop) : !(M - W >= kBottom)) {
      x = plain(x);  // this addition leaves the binade
      --k;
      continue;
    }
    const double q = std::floor(W);
    const double f = W - q;
    double Mc = M;
    double d;
    if (f == 0.5) {
      // halfway: the nearest even multiple, then q + (q odd) per addition
      const double a = up ? M + q : M - q;
      const double b = up ? a + 1.0 : a - 1.0;
      Mc = (odd(a) ? b : a);
      --k;
      d = q + (odd(q) ? 1.0 : 0.0);
    } else {
      d = (f < 0.5) ? q : q + 1.0;
    }
    if (k > 0) {
      if (d == 0.0) {
        k = 0;  // w is below half a spacing: the sum no longer moves
      } else {
        // additions whose exact sum stays inside the binade
        double t;
        if (up) {
          const double room = kTop - Mc - W;
          t = room > 0.0 ? std::ceil(room / d) : 0.0;
        } else {
          const double room = Mc - W - kBottom;
          t = room >= 0.0 ? std::floor(room / d) + 1.0 : 0.0;
        }
        if (t > static_cast<double>(k)) t = static_cast<double>(k);
        Mc = up ? Mc + t * d : Mc - t * d;
        k -= static_cast<int>(t);
      }
    }
    const float mag = static_cast<float>(std::ldexp(Mc, e - 24));
    x = up ? mag : -mag;
  }
  return neg ? -x : x;
}

// Runs of up to this many terms are added one by one in place (a measured
// cut-over); longer ones in O(binades) by add_repeat.  On the piled-up
// million-particle dam break (substeps/s): 24: 40.6, 32: 41.3, 40: 41.8,
// 48: 42.2, 56: 42.4, 64: 42.6, 96: 42.5, 128: 42.4.
constexpr int kAddInPlace = 64;

// add_repeat for hot loops, with the in-place cut-over.
__host__ __device__ inline float add_k(float x, float w, int k) {
  if (k <= kAddInPlace) {
    for (int i = 0; i < k; ++i) x += w;
    return x;
  }
  return add_repeat(x, w, k);
}

}  // namespace fluidhip
