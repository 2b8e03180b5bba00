// Brute-force check of apirl::alpha_after against the loop it replaces
// (a += u, n times), over random and adversarial inputs: tie increments, binade
// crossings, negative and tiny starting values.  Prints the number of mismatches.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "alpha_exact.h"

static float loop_add(float a, float u, int n) {
  volatile float x = a;
  for (int i = 0; i < n; i++) x = x + u;
  return x;
}

int main(int argc, char** argv) {
  long trials = argc > 1 ? atol(argv[1]) : 2000000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<float> ua(-0.2f, 1.2f), uu(1e-4f, 0.05f);
  std::uniform_int_distribution<int> un(0, 600);
  long bad = 0, ties = 0;
  for (long t = 0; t < trials; t++) {
    float a = ua(rng), u = uu(rng);
    int n = un(rng);
    int kind = (int)(t % 5);
    if (kind == 1) {
      // force a tie: clear the bits of u below half an ulp of a's binade and set that bit
      uint32_t ab = apirl::f2u(a > 0 ? a : -a);
      int E = (int)((ab >> 23) & 0xff);
      int eu = (int)((apirl::f2u(u) >> 23) & 0xff);
      int shift = (E - 23 - 1) - (eu - 23);  // position of q/2 within u's mantissa
      if (shift > 0 && shift < 23) {
        uint32_t ub = apirl::f2u(u);
        ub &= ~((1u << shift) - 1u);
        ub |= (1u << (shift - 1));
        u = apirl::u2f(ub);
        ties++;
      }
    } else if (kind == 2) {
      a = 0.25f - u * (float)(n / 3);  // cross the 0.25 binade edge
    } else if (kind == 3) {
      a = apirl::u2f(apirl::f2u(0.5f) - (uint32_t)(t % 7));  // just below 0.5
    } else if (kind == 4) {
      a = ua(rng) * 1e-3f;  // tiny or negative starts
    }
    float x = loop_add(a, u, n), y = apirl::alpha_after(a, u, n);
    if (apirl::f2u(x) != apirl::f2u(y)) {
      if (bad < 10) printf("mismatch a=%.9g u=%.9g n=%d loop=%.9g closed=%.9g\n", a, u, n, x, y);
      bad++;
    }
  }
  printf("trials=%ld ties=%ld mismatches=%ld\n", trials, ties, bad);
  return bad ? 1 : 0;
}
