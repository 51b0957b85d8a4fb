// Host-side check that KastaunT::mu_plus_fast() returns mu_plus() bit for
// bit (the reference's 50-step bisection, c2p.py:162-184) over random
// conserved states spanning the admissible range and beyond: D from 1e-12 to
// 1e6, W - 1 from 1e-8 to 1e4, h - 1 from 1e-10 to 1e4, |B| / sqrt(D) from
// 1e-8 to 1e4, with clipped tau and over-bound S now and then.
// Usage: mu_plus_check N SEED   -> prints "checked N mismatches M fallback F"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "srmhd_c2p.h"

using namespace srmhd;

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 100000;
  std::mt19937_64 rng(argc > 2 ? atol(argv[2]) : 1);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto logu = [&](double lo, double hi) {
    return std::exp(std::log(lo) + U(rng) * (std::log(hi) - std::log(lo)));
  };
  srmhd_eos e{};
  e.K = 0.0;
  e.gamma = 5.0 / 3.0;
  e.gamma_th = 5.0 / 3.0;
  long bad = 0, big = 0;
  for (long q = 0; q < n; ++q) {
    const double D = logu(1e-12, 1e6), W = 1.0 + logu(1e-8, 1e4), h = 1.0 + logu(1e-10, 1e4);
    const double v = std::sqrt(1.0 - 1.0 / (W * W)), th = U(rng) * M_PI, ph = 2 * M_PI * U(rng);
    const double vv[3] = {v * std::sin(th) * std::cos(ph), v * std::sin(th) * std::sin(ph),
                          v * std::cos(th)};
    const double Bm = logu(1e-8, 1e4) * std::sqrt(D);
    double B[3];
    for (int i = 0; i < 3; ++i) B[i] = Bm * (2 * U(rng) - 1);
    const double rho = D / W, p = 0.4 * rho * (h - 1.0), rhoh = rho * h;
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const double vB = vv[0] * B[0] + vv[1] * B[1] + vv[2] * B[2];
    double c[8];
    c[CD] = D;
    for (int i = 0; i < 3; ++i) c[CS + i] = (rhoh * W * W + B2) * vv[i] - vB * B[i];
    c[CT] = rhoh * W * W - p - D + 0.5 * (B2 + B2 * v * v - vB * vB);
    for (int i = 0; i < 3; ++i) c[CB + i] = B[i];
    if (U(rng) < 0.05) c[CT] *= 2 * U(rng) - 0.5;
    if (U(rng) < 0.02) c[CS] *= 3.0;
    Kastaun k;
    k.init(c, e);
    const double a = k.mu_plus(), b = k.mu_plus_fast();
    if (std::memcmp(&a, &b, sizeof a) != 0) ++bad;
    if (!(k.r2 <= 1e8)) ++big;
  }
  std::printf("checked %ld mismatches %ld fallback %ld\n", n, bad, big);
  return bad != 0;
}
