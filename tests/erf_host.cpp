// Host build of the kernels' table erf (xpg_amd/csrc/xpg_erf.h): reads numbers from stdin, prints erf_tab of each
// and erfc_tab of its absolute value, with 17 significant digits (an exact round trip for float64).
#include <cmath>
#include <cstdio>

#include "xpg_erf.h"

int main() {
  double x;
  while (std::scanf("%lf", &x) == 1) std::printf("%.17g %.17g\n", xpg::erf_tab(x), xpg::erfc_tab(std::fabs(x)));
  return 0;
}
