// Exact closed form of the Siddon alpha accumulators.
//
// The reference advances each crossing parameter by repeated float addition,
// alpha += alpha_u (recon/src/Siddon.cpp:322,330,338 and :626,637), so the value of
// the n-th crossing is n rounded additions, not alpha0 + n*alpha_u.  A voxel-driven
// backprojector needs those exact values without walking the ray.  Within one binade
// [2^e, 2^(e+1)) every float is a multiple of q = 2^(e-23), so fl(a + u) =
// a + round_q(u) as long as the exact sum stays below 2^(e+1): the increment in units
// of q is constant (round-half-even ties settle after one step), and a run of k
// additions is one multiply.  Binade crossings and values outside the normal
// positive range are taken one real float addition at a time, so the result is
// bit-identical to the loop.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define APIRL_HD __host__ __device__ __forceinline__
#else
#define APIRL_HD inline
#endif

namespace apirl {

APIRL_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  union { float f; uint32_t u; } c;
  c.f = f;
  return c.u;
#endif
}

APIRL_HD float u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  union { uint32_t u; float f; } c;
  c.u = u;
  return c.f;
#endif
}

// 2^k as a float for k in [-126, 127].
APIRL_HD float pow2f(int k) { return u2f((uint32_t)(k + 127) << 23); }

// a after n repeated float additions of u (a_{t+1} = fl(a_t + u)).
APIRL_HD float alpha_after(float a, float u, int n) {
  while (n > 0) {
    const uint32_t ab = f2u(a);
    const int E = (int)((ab >> 23) & 0xffu);  // biased exponent
    // fall back to one real addition outside the positive normal range, for a
    // non-finite or non-positive increment, or near the top of the exponent range
    if ((ab >> 31) != 0u || E < 24 || E >= 250 || !(u > 0.f) || !(u < 3.0e38f)) {
      a = a + u;
      n--;
      continue;
    }
    const int e2 = E - 150;  // q = 2^e2
    const uint32_t N = (ab & 0x7fffffu) | 0x800000u;  // a = N * q, N in [2^23, 2^24)
    // U = u / q, exact (power-of-two scaling) while it stays a normal float
    const int eu = (int)((f2u(u) >> 23) & 0xffu);
    const int eU = eu - e2;  // biased exponent of U
    if (eu == 0 || eU <= 0 || eU >= 127 + 24) {  // denormal u, or u >= 2^24 q = 2a
      a = a + u;
      n--;
      continue;
    }
    const float U = u2f((f2u(u) & 0x807fffffu) | ((uint32_t)eU << 23));
    const float M = floorf(U);
    const float fr = U - M;  // exact
    const uint32_t Mi = (uint32_t)M;
    uint32_t d;
    if (fr == 0.5f) {
      // tie: round half to even.  From an even N the step is M + (M & 1) and keeps N
      // even; from an odd N take one real addition first.
      if (N & 1u) {
        a = a + u;
        n--;
        continue;
      }
      d = Mi + (Mi & 1u);
    } else {
      d = Mi + (fr > 0.5f ? 1u : 0u);
    }
    // steps k whose exact sum N + (k-1) d + U stays below 2^24 round on the q grid:
    // k <= kmax = (L0 - 1) / d + 1 with L0 = 2^24 - N - M (all below 2^25).
    const int32_t L0 = (int32_t)(1u << 24) - (int32_t)N - (int32_t)Mi;
    int k;
    if (L0 < 1) {
      k = 0;
    } else if (d == 0u || (uint64_t)(uint32_t)(n - 1) * d <= (uint64_t)(L0 - 1)) {
      k = n;  // the whole run stays in this binade (common case: no division)
    } else {
      // kmax - 1 = floor((L0 - 1) / d): float estimate, corrected to the exact quotient
      uint32_t q = (uint32_t)((float)(L0 - 1) / (float)d);
      while (q * d > (uint32_t)(L0 - 1)) q--;
      while ((q + 1) * d <= (uint32_t)(L0 - 1)) q++;
      k = (int)q + 1;  // < n here
    }
    if (k == 0) {
      a = a + u;  // this addition crosses into the next binade
      n--;
      continue;
    }
    // The last of the k sums was below 2^24 before rounding, so N + k d <= 2^24: the
    // conversion to float and the scaling by q are exact.
    const uint32_t Nn = N + (uint32_t)k * d;
    a = (float)Nn * pow2f(e2);
    n -= k;
  }
  return a;
}

}  // namespace apirl
