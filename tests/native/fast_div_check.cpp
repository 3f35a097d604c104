// Host check of fast_div / fdiv (kernels.hpp), the multiply-and-shift division the face-flux
// kernels use for their cell and thread indices: fdiv(n, fast_div(d)) == n / d for every
// divisor 1..5000 and a spread of large ones (powers of two, their neighbours, the largest
// allowed), at numerators around every kind of multiple of d up to 2^31 - 1: q d - 1, q d,
// q d + 1 and q d + d - 1 for q from 0 to the top, plus random numerators.
#include <cstdio>
#include <random>
#include <vector>

#include "kernels.hpp"

int main() {
  std::vector<long long> ds;
  for (long long d = 1; d <= 5000; ++d) ds.push_back(d);
  for (int b = 13; b <= 30; ++b)
    for (long long o : {-3LL, -1LL, 0LL, 1LL, 3LL}) ds.push_back((1LL << b) + o);
  for (long long d : {65536LL + 262 * 262, 262LL * 262 * 262, 1000003LL, 2147483646LL, 2147483647LL}) ds.push_back(d);
  const long long top = (1LL << 31) - 1;
  std::mt19937_64 rng(12345);
  long bad = 0, checked = 0;
  for (long long d : ds) {
    const chq::FastDiv f = chq::fast_div(int(d));
    auto chk = [&](long long n) {
      if (n < 0 || n > top) return;
      ++checked;
      if ((long long)chq::fdiv(unsigned(n), f) != n / d) {
        if (bad < 10) std::printf("d %lld n %lld: got %u want %lld\n", d, n, chq::fdiv(unsigned(n), f), n / d);
        ++bad;
      }
    };
    const long long qmax = top / d;
    const long long step = qmax / 4000 + 1;
    for (long long q = 0; q <= qmax + 1; q += (q < 64 || q > qmax - 64) ? 1 : step)
      for (long long r : {-1LL, 0LL, 1LL, d - 1}) chk(q * d + r);
    for (long long n = top - 70; n <= top; ++n) chk(n);
    for (int t = 0; t < 2000; ++t) chk((long long)(rng() & (unsigned long long)top));
  }
  std::printf("divisors %zu checked %ld mismatches %ld\n", ds.size(), checked, bad);
  return bad ? 1 : 0;
}
