// Prints normal_quantile(p) for p read from stdin, then the inverse normal
// transform of a vector read from the next line (values separated by spaces,
// "nan" for missing), for the test to compare with scipy.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../eqtl_amd/host/pheno_norm.hpp"

int main() {
  std::string line;
  std::getline(std::cin, line);
  std::istringstream ps(line);
  double p;
  while (ps >> p) std::printf("%.17g\n", eqtl::normal_quantile(p));
  std::printf("--\n");
  std::getline(std::cin, line);
  std::istringstream ys(line);
  std::vector<double> y;
  std::string tok;
  while (ys >> tok) y.push_back(std::stod(tok));
  eqtl::inverse_normal_transform(y.data(), y.size());
  for (double v : y) std::printf("%.17g\n", v);
  return 0;
}
