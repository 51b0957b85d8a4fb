// Host-side check that the reverse-mode HLLD flux VJP (flux_vjp_hlld) equals
// the forward-mode one (flux_vjp_dual, one dual pass per input) on random
// face states of both normal directions, at the root p* the forward solver
// finds (faces it sends to the HLLE fallback included), for three EOS
// (ideal gas, hybrid with a cold polytrope, cold polytrope alone).
// Usage: flux_adj_check N SEED -> prints "faces N fallback F worst W"
//        (W: the largest |lq_rev - lq_dual| / max_m |lq_dual|); exit 1 if
//        W > 1e-7.  The two forms round differently; where the implicit
//        derivative of p* is ill-conditioned the difference reaches ~2e-8
//        (one face in 20 000 with the hybrid EOS, the same for a form whose
//        side states are differentiated in forward mode); a dropped or
//        wrong term gives 1e-4 and more.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "srmhd_flux_adj.h"

using namespace srmhd;

template <int N>
static double one(std::mt19937_64& rng, const srmhd_eos& e, bool* fellback) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto logu = [&](double lo, double hi) {
    return std::exp(std::log(lo) + U(rng) * (std::log(hi) - std::log(lo)));
  };
  double ql[7], qr[7];
  for (double* q : {ql, qr}) {
    q[0] = logu(0.1, 10.0);
    q[1] = logu(0.01, 10.0);
    for (int m = 2; m < 5; ++m) q[m] = 3.0 * (2.0 * U(rng) - 1.0);
    for (int m = 5; m < 7; ++m) q[m] = 2.0 * (2.0 * U(rng) - 1.0);
  }
  if (U(rng) < 0.2) {   // nearby states (a weak jump)
    for (int m = 0; m < 7; ++m) qr[m] = ql[m] * (1.0 + 0.05 * (2.0 * U(rng) - 1.0));
  }
  const double Bn = U(rng) < 0.05 ? 0.0 : 2.0 * U(rng) - 1.0;
  Prim L, R;
  side_prims_q<N, double>(ql, Bn, e, -1, L);
  side_prims_q<N, double>(qr, Bn, e, -1, R);
  HlldPrep<N> h;
  h.run(L, R, e);
  Kastaun k;
  k.init(h.uH, e);
  Prim ph;
  c2p_single(k, 1e-10, 1e-10, ph);
  double F[8], ps;
  int it;
  hlld_finish<N>(h, L, R, p_hll_from_prim(ph, e), F, nullptr, ps, it);
  *fellback = !(ps >= 0.0);
  double lf[8];
  for (int m = 0; m < 8; ++m) lf[m] = 2.0 * U(rng) - 1.0;
  lf[CB + N] = 0.0;
  double a[15], b[15];
  flux_vjp_dual<N, SOLVER_HLLD>(ql, qr, Bn, e, ps, lf, a);
  flux_vjp_hlld<N>(ql, qr, Bn, e, ps, lf, b);
  double mx = 0.0, err = 0.0;
  for (int m = 0; m < 15; ++m) mx = std::fmax(mx, std::fabs(a[m]));
  for (int m = 0; m < 15; ++m) err = std::fmax(err, std::fabs(a[m] - b[m]));
  if (!(mx > 0.0)) return 0.0;
  return err / mx;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 2000;
  std::mt19937_64 rng(argc > 2 ? atol(argv[2]) : 1);
  // the ideal gas (K = 0, gamma = 4/3), a hybrid EOS with a cold polytrope
  // (K = 0.5, gamma = 2, gamma_th = 5/3) and the cold polytrope alone
  srmhd_eos eos[3] = {};
  eos[0].K = 0.0;
  eos[0].gamma = eos[0].gamma_th = 4.0 / 3.0;
  eos[1].K = 0.5;
  eos[1].gamma = 2.0;
  eos[1].gamma_th = 5.0 / 3.0;
  eos[2].K = 0.5;
  eos[2].gamma = 2.0;
  eos[2].gamma_th = 1.0;
  eos[2].is_cold = 1;
  double worst = 0.0;
  long fb = 0, bad = 0;
  for (long q = 0; q < n; ++q) {
    bool f;
    const srmhd_eos& e = eos[(q / 2) % 3];
    const double w = (q % 2) ? one<1>(rng, e, &f) : one<0>(rng, e, &f);
    fb += f;
    if (!(w <= 1e-7)) {
      ++bad;
      if (bad <= 5)
        std::printf("face %ld (eos %ld): rel err %.3e (fallback %d)\n", q, (q / 2) % 3, w, (int)f);
    }
    if (w > worst || w != w) worst = w;
  }
  std::printf("faces %ld fallback %ld bad %ld worst %.3e\n", n, fb, bad, worst);
  return bad != 0;
}
