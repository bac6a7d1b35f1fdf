/* Random and edge-case check that qdiv (spruce_amd/csrc/qdiv.hpp), the reciprocal-corrected
 * quotient used by the GPU stencils, equals the IEEE quotient a/b bit for bit: ordinary exponents,
 * significands near all ones, exact quotients, zeros, infinities, NaN, and exponents on both sides
 * of the thresholds where qdiv switches to the IEEE division. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qdiv.hpp"

static uint64_t s = 0x1234567;
static uint64_t nx(void) { s += 0x9E3779B97F4A7C15ull; uint64_t z = s; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
static double rnd(int emin, int emax) { uint64_t m = nx() & ((1ull << 52) - 1); int e = emin + (int)(nx() % (uint64_t)(emax - emin + 1));
  uint64_t b = ((uint64_t)(e + 1023) << 52) | m; double d; memcpy(&d, &b, 8); if (nx() & 1) d = -d; return d; }
int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000L, bad = 0, fast = 0;
  for (long k = 0; k < n; k++) {
    double a, b; int mode = k % 7;
    if (mode == 0) { a = rnd(-60, 60); b = fabs(rnd(-40, 40)); }
    else if (mode == 1) { a = rnd(-1, 1); uint64_t bb = 0x3FFFFFFFFFFFFFFFull - (nx() % 1024); memcpy(&b, &bb, 8); }
    else if (mode == 2) { a = rnd(-1000, 1000); b = fabs(rnd(-100, 100)); }
    else if (mode == 3) { a = (double)(int64_t)(nx() % 2000001) - 1000000.0; b = (double)(1 + nx() % 1000); }
    else if (mode == 4) { a = (k & 8) ? 0.0 : -0.0; if (k & 16) a = (k & 32) ? INFINITY : NAN; b = fabs(rnd(-40, 40)); }
    else if (mode == 5) { a = rnd(-710, -690); if (k & 8) a = rnd(690, 710); b = fabs(rnd(-160, 160)); }
    else { a = rnd(-1022, 1023); b = fabs(rnd(-1022, 1023)); if (k & 8) { uint64_t u = nx() & ((1ull << 52) - 1); memcpy(&a, &u, 8); } }
    const double rb = 1.0 / b;
    if (!(rb > 0x1p-1022) || isinf(rb)) continue;   /* qdiv's contract: rb a normal number */
    double x = spruce::qdiv(a, b, rb), y = a / b;
    fast += spruce::qdiv_safe(a, a * rb);
    if (memcmp(&x, &y, 8) != 0 && !(isnan(x) && isnan(y))) { if (bad < 5) printf("mismatch a=%a b=%a %a %a\n", a, b, x, y); bad++; }
  }
  printf("checked %ld (corrected-reciprocal path %ld) mismatches %ld\n", n, fast, bad);
  return bad != 0;
}
