// Host check of the correctly rounded powers in spruce_amd/csrc/crpow.hpp against quad-precision powq.
#include <quadmath.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "../spruce_amd/csrc/crpow.hpp"
using namespace spruce;
static uint64_t s = 0x51ED;
static uint64_t nx() { s += 0x9E3779B97F4A7C15ull; uint64_t z = s; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 1000000;
  struct { const char* name; double y; double (*f)(double); } fs[] = {
    {"1.5", 1.5, pow_1p5}, {"2.5", 2.5, pow_2p5}, {"-1.5", -1.5, pow_m1p5}, {"1/3", 1.0 / 3.0, pow_third}};
  int fail = 0;
  for (auto& f : fs) {
    long bad = 0;
    for (long i = 0; i < n; ++i) {
      double u = (nx() >> 11) * 0x1p-53;
      double x = (i % 3 == 0) ? exp2(-60.0 + 160.0 * u) : exp2(3.0 + 30.0 * u);  // wide range, and 8..8e9
      double c = (double)powq((__float128)x, (__float128)f.y);
      double g = f.f(x);
      if (memcmp(&c, &g, 8)) { if (bad < 3) printf("  %s x=%a got %a want %a\n", f.name, x, g, c); ++bad; }
    }
    printf("x^%s: %ld / %ld differ from the correctly rounded value\n", f.name, bad, n);
    fail |= bad != 0;
  }
  return fail;
}
