// Host check of the tabulated t tail the kernels use (TTail, tail_build and
// p_from_t_tab in eqtl_amd/csrc/numerics.hpp): reads "df t" pairs on stdin
// and prints, per pair, the tabulated p and the continued-fraction p
// (p_from_t_dev, the restatement of p_from_t, util.cpp:209-264).
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>

#include "../eqtl_amd/csrc/numerics.hpp"

int main() {
  std::map<double, std::unique_ptr<eqtl_amd::TTail>> tabs;
  double df, t;
  while (std::scanf("%lf %lf", &df, &t) == 2) {
    const double lbeta = std::lgamma(0.5 * df) + std::lgamma(0.5) - std::lgamma(0.5 * df + 0.5);
    auto& T = tabs[df];
    if (!T) {
      T.reset(new eqtl_amd::TTail());
      eqtl_amd::tail_build(*T, df, lbeta);
    }
    std::printf("%.17g %.17g\n", eqtl_amd::p_from_t_tab(*T, t), eqtl_amd::p_from_t_dev(t, df, lbeta));
  }
  return 0;
}
