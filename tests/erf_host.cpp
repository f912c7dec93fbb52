// Host build of the kernels' table erf (xpg_amd/csrc/xpg_erf.h): reads numbers from stdin, prints erf_tab of each
// with 17 significant digits (an exact round trip for float64).
#include <cstdio>

#include "xpg_erf.h"

int main() {
  double x;
  while (std::scanf("%lf", &x) == 1) std::printf("%.17g\n", xpg::erf_tab(x));
  return 0;
}
