// Host check of xcd_tile (kernels.hpp), the XCD-grouped tile order of con2prim, the corner EMF
// and the float update: for every tile count n and group size it is a bijection on [0, n), and
// inside the whole groups the block that follows block b on the same XCD (b + 8) takes the
// next tile, except across the end of a run.
#include <cstdio>
#include <vector>

#include "kernels.hpp"

int main() {
  long bad = 0, checked = 0;
  const int groups[] = {0, 1, 2, 3, 4, 8, 16, 32, 64};
  std::vector<int> ns;
  for (int n = 0; n <= 700; ++n) ns.push_back(n);
  for (int n : {1023, 1024, 1025, 4096, 16383, 66049, 70384, 163840}) ns.push_back(n);
  for (int grp : groups) {
    for (int n : ns) {
      std::vector<char> seen(size_t(n), 0);
      for (int b = 0; b < n; ++b) {
        const int t = chq::xcd_tile(b, n, grp);
        if (t < 0 || t >= n || seen[size_t(t)]++) ++bad;
      }
      if (grp > 1) {
        const int span = 8 * grp, full = (n / span) * span;
        for (int b = 0; b + 8 < full; ++b) {
          const int l = b >> 3;
          if (l % grp == grp - 1) continue;   // the last block of a run
          if (chq::xcd_tile(b + 8, n, grp) != chq::xcd_tile(b, n, grp) + 1) ++bad;
        }
        for (int b = full; b < n; ++b)
          if (chq::xcd_tile(b, n, grp) != b) ++bad;   // the tail keeps the launch order
      } else {
        for (int b = 0; b < n; ++b)
          if (chq::xcd_tile(b, n, grp) != b) ++bad;
      }
      ++checked;
    }
  }
  std::printf("cases %ld mismatches %ld\n", checked, bad);
  return bad ? 1 : 0;
}
