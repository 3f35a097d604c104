// Checks chq::div6 (numerics.hpp) against the IEEE division x / 6 on random doubles of every
// exponent, on dense samples around 1, on integers and on the edge values of its fast range.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "common.hpp"
#include "numerics.hpp"

static uint64_t s = 0x9e3779b97f4a7c15ull;
static uint64_t next() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static bool same(double a, double b) { uint64_t x, y; std::memcpy(&x, &a, 8); std::memcpy(&y, &b, 8); return x == y; }

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 50000000L;
  long bad = 0, checked = 0;
  for (long t = 0; t < n; ++t) {
    const uint64_t b = next();
    double x;
    std::memcpy(&x, &b, 8);
    if (t % 4 == 1) x = std::ldexp(double(b >> 11) / 9007199254740992.0 + 1.0, int(b % 64) - 32) * ((b >> 3) & 1 ? -1 : 1);
    if (t % 4 == 2) x = double(int64_t(b % 200001)) - 100000.0;
    if (!std::isfinite(x)) continue;
    ++checked;
    if (!same(chq::div6(x), x / 6.0)) { if (bad < 5) std::printf("mismatch %a\n", x); ++bad; }
  }
  const double edge[] = {0.0, -0.0, 6.0, -6.0, 1.0, 0x1p-900, -0x1p-900, 0x1.fffffffffffffp-901, 0x1p1000,
                         0x1.0000000000001p1000, 1e-310, -1e-310, 1e308, -1e308, INFINITY, -INFINITY};
  for (double x : edge) {
    ++checked;
    if (!same(chq::div6(x), x / 6.0)) { std::printf("edge mismatch %a\n", x); ++bad; }
  }
  const double nan = std::nan("");
  if (!std::isnan(chq::div6(nan))) { std::printf("NaN not propagated\n"); ++bad; }
  std::printf("checked %ld mismatches %ld\n", checked, bad);
  return bad ? 1 : 0;
}
