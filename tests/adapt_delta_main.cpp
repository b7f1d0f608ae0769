// Reads lines "d d0 move wide n" and prints sph::adapt_delta of each (host
// build of the device header; used by tests/test_adapt_delta.py).
#include <cstdio>

#include "sph_device.hpp"

int main() {
  double d, d0;
  unsigned move, wide;
  int n;
  while (std::scanf("%lf %lf %u %u %d", &d, &d0, &move, &wide, &n) == 5)
    std::printf("%.17g\n", sph::adapt_delta(d, d0, move, wide, n));
  return 0;
}
