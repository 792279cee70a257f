// Host check of the x87 long-double ACAT emulation shared with the kernels:
// compares acat_term_ld(p) with the platform's own tanl((0.5L - p) * kPi)
// (stats_extra.cpp:20) and acat_p_final with AcatAcc::p's formula.
#include <cmath>
#include <cstdio>
#include <random>

#include "../eqtl_amd/csrc/numerics.hpp"

int main() {
  const long double kPi = 3.14159265358979323846L;
  std::mt19937_64 r(1);
  std::uniform_real_distribution<double> u(0, 1);
  double worst = 0;
  long n = 0;
  auto chk = [&](double p) {
    const long double t = tanl((0.5L - static_cast<long double>(p)) * kPi);
    const double e = eqtl_amd::acat_term_ld(p);
    const double rel = t == 0 ? std::fabs(e) : std::fabs(static_cast<double>(t) - e) / std::fabs(static_cast<double>(t));
    if (rel > worst) worst = rel;
    ++n;
  };
  const double fixed[] = {1e-300, 1e-20, 1.3e-20, 1.355e-20, 0x1p-66, 0x1.0000000000001p-66, 1.4e-20, 1e-19,
                          1e-15, 1e-10, 0.1, 0.2499999, 0.25, 0.3, 0.5, 0.7, 0.75, 0.7500001, 0.99, 1 - 1e-11};
  for (double p : fixed) chk(p);
  for (int i = 0; i < 400000; ++i) {
    const double p = std::pow(10.0, -300 * u(r));
    if (p < 1 - 1e-12) chk(p);
    chk(u(r) * (1 - 1e-12));
  }
  // finals: AcatAcc::p on random sums
  double worst_f = 0;
  for (int i = 0; i < 200000; ++i) {
    const int used = 1 + static_cast<int>(u(r) * 1000);
    const double T = std::tan((u(r) - 0.5) * 3.14159) * std::pow(10.0, 20 * u(r));
    const long double Tl = static_cast<long double>(T) * used / used;
    double ref = 0.5 - static_cast<double>(atanl(Tl) / kPi);
    ref = ref < 0 ? 0 : (ref > 1 ? 1 : ref);
    const double got = eqtl_amd::acat_p_final(eqtl_amd::dd{T * used, 0.0}, used);
    // the reference rounds atanl(T)/kPi (x87) to double next to 0.5, so its
    // result is only defined to one step of that grid (2^-54): compare to it
    const double err = std::fabs(got - ref) / (1.2e-16 + 1e-12 * ref);
    if (err > worst_f) worst_f = err;
  }
  // the dispatching term (double-precision form on [1e-6, 1 - 1e-6]) against tanl
  double worst_d = 0;
  for (int i = 0; i < 400000; ++i) {
    const double p = i % 2 ? std::pow(10.0, -8 * u(r)) : u(r) * (1 - 1e-12);
    const long double t = tanl((0.5L - static_cast<long double>(p)) * kPi);
    const double e = eqtl_amd::acat_term(p);
    const double rel = std::fabs(static_cast<double>(t) - e) / std::fabs(static_cast<double>(t));
    if (rel > worst_d) worst_d = rel;
  }
  std::printf("terms=%ld worst_term_rel=%.3e worst_dispatch_rel=%.3e worst_final_scaled=%.3e\n", n, worst,
              worst_d, worst_f);
  return (worst < 2e-15 && worst_d < 1e-13 && worst_f <= 1.0) ? 0 : 1;
}
