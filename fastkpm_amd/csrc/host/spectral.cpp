// Spectral math of the KPM (host, FP64): kernels, Chebyshev arrays, expansion
// coefficients, moment transforms, Fermi functions and energies.  Formulas
// follow src/fastkpm.cpp of the reference (cited per function).  The root
// finder of filling_to_mu is a bracketing Illinois (modified regula falsi)
// solver of this library; the reference uses Boost's TOMS748 with the same
// bracket and a 30-bit relative tolerance.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <stdexcept>

#include "fastkpm.h"

namespace fkpm {

// src/fastkpm.cpp:13-20
Vec<double> jackson_kernel(int M) {
  Vec<double> g(M);
  const double Mp = M + 1.0;
  for (int m = 0; m < M; m++)
    g[m] = (1.0 / Mp) * ((Mp - m) * std::cos(Pi * m / Mp) + std::sin(Pi * m / Mp) / std::tan(Pi / Mp));
  return g;
}

// src/fastkpm.cpp:22-28
Vec<double> lorentz_kernel(int M, double lambda) {
  Vec<double> g(M);
  for (int m = 0; m < M; m++) g[m] = std::sinh(lambda * (1.0 - (double)m / M)) / std::sinh(lambda);
  return g;
}

// src/fastkpm.cpp:30-39
void chebyshev_fill_array(double x, Vec<double>& ret, int kind) {
  assert(kind == 1 || kind == 2);
  if (!ret.empty()) ret[0] = 1.0;
  if (ret.size() > 1) ret[1] = kind * x;
  for (size_t m = 2; m < ret.size(); m++) ret[m] = 2 * x * ret[m - 1] - ret[m - 2];
}

// src/fastkpm.cpp:41-78 (direct-sum branch)
Vec<double> expansion_coefficients(int M, int Mq, std::function<double(double)> f, EnergyScale es) {
  if (Mq < M) throw std::invalid_argument("expansion_coefficients: Mq must be >= M");
  const Vec<double> kernel = jackson_kernel(M);
  Vec<double> fp(M, 0.0), T(M), ret(M);
  for (int i = 0; i < Mq; i++) {
    const double x_i = std::cos(Pi * (i + 0.5) / Mq);
    const double f_i = f(es.unscale(x_i));
    chebyshev_fill_array(x_i, T);
    for (int m = 0; m < M; m++) fp[m] += f_i * T[m];
  }
  for (int m = 0; m < M; m++) ret[m] = (m == 0 ? 1.0 : 2.0) * kernel[m] * fp[m] / Mq;
  return ret;
}

// src/fastkpm.cpp:291-298
double moment_product(Vec<double> const& c, Vec<double> const& mu) {
  double r = 0.0;
  for (size_t i = 0; i < c.size(); i++) r += c[i] * mu[i];
  return r;
}

// src/fastkpm.cpp:236-256
Vec<double> moment_transform(Vec<double> const& moments, int Mq) {
  const int M = (int)moments.size();
  const Vec<double> kernel = jackson_kernel(M);
  Vec<double> T(M), mup(M), gamma(Mq);
  for (int m = 0; m < M; m++) mup[m] = moments[m] * kernel[m];
  for (int i = 0; i < Mq; i++) {
    const double x_i = std::cos(Pi * (i + 0.5) / Mq);
    chebyshev_fill_array(x_i, T);
    double g = 0.0;
    for (int m = 0; m < M; m++) g += (m == 0 ? 1 : 2) * mup[m] * T[m];
    gamma[i] = g;
  }
  return gamma;
}

// src/fastkpm.cpp:314-322
double density_product(Vec<double> const& gamma, std::function<double(double)> f, EnergyScale es) {
  const int Mq = (int)gamma.size();
  double r = 0.0;
  for (int i = 0; i < Mq; i++) r += gamma[i] * f(es.unscale(std::cos(Pi * (i + 0.5) / Mq)));
  return r / Mq;
}

// src/fastkpm.cpp:324-333
void density_function(Vec<double> const& gamma, EnergyScale es, Vec<double>& x, Vec<double>& rho) {
  const int Mq = (int)gamma.size();
  x.resize(Mq);
  rho.resize(Mq);
  for (int i = Mq - 1; i >= 0; i--) {
    const double x_i = std::cos(Pi * (i + 0.5) / Mq);
    x[Mq - 1 - i] = es.unscale(x_i);
    rho[Mq - 1 - i] = gamma[i] / (Pi * std::sqrt(1 - x_i * x_i) * es.mag());
  }
}

// src/fastkpm.cpp:362-373
void integrated_density_function(Vec<double> const& gamma, EnergyScale es, Vec<double>& x,
                                 Vec<double>& irho) {
  const int Mq = (int)gamma.size();
  x.resize(Mq);
  irho.resize(Mq);
  double acc = 0.0;
  for (int i = Mq - 1; i >= 0; i--) {
    x[Mq - 1 - i] = es.unscale(std::cos(Pi * (i + 0.5) / Mq));
    irho[Mq - 1 - i] = (acc + 0.5 * gamma[i]) / Mq;
    acc += gamma[i];
  }
}

// src/fastkpm.cpp:375-383
double fermi_energy(double x, double kT, double mu) {
  const double alpha = (x - mu) / std::abs(kT);
  if (kT < 1e-15 || std::abs(alpha) > 20) return (x < mu) ? (x - mu) : 0.0;
  return -kT * std::log(1 + std::exp(-alpha));
}

// src/fastkpm.cpp:385-393
double fermi_density(double x, double kT, double mu) {
  const double alpha = (x - mu) / std::abs(kT);
  if (kT < 1e-15 || std::abs(alpha) > 20) return (x < mu) ? 1.0 : 0.0;
  return 1.0 / (std::exp(alpha) + 1.0);
}

// src/fastkpm.cpp:395-400
double mu_to_filling(Vec<double> const& gamma, EnergyScale const& es, double kT, double mu) {
  const double n_occ = density_product(gamma, [&](double x) { return fermi_density(x, kT, mu); }, es);
  const double n_tot = density_product(gamma, [](double) { return 1.0; }, es);
  return n_occ / n_tot;
}

// src/fastkpm.cpp:401-408
double mu_to_filling(Vec<double> const& evals, double kT, double mu) {
  double n_occ = 0;
  for (double x : evals) n_occ += fermi_density(x, kT, mu);
  return n_occ / (double)evals.size();
}

// Root of f in [lo, hi] (f(lo), f(hi) of opposite sign or one of them zero):
// Illinois-modified regula falsi, stopping when the bracket is below a 2^-30
// relative width or after 50 evaluations; returns the bracket midpoint (the
// contract of the reference's root_solver, src/fastkpm.cpp:410-416).
static double root_solver(std::function<double(double)> f, double lo, double hi) {
  double fl = f(lo), fh = f(hi);
  if (fl == 0) return lo;
  if (fh == 0) return hi;
  if ((fl > 0) == (fh > 0)) throw std::domain_error("root_solver: root is not bracketed");
  const double tol = std::ldexp(1.0, -30);
  int side = 0;
  for (int it = 0; it < 50; it++) {
    if (std::abs(hi - lo) <= tol * std::max(std::abs(lo), std::abs(hi))) break;
    double x = (lo * fh - hi * fl) / (fh - fl);
    if (!(x > std::min(lo, hi) && x < std::max(lo, hi))) x = 0.5 * (lo + hi);
    const double fx = f(x);
    if (fx == 0) return x;
    if ((fx > 0) == (fh > 0)) {
      hi = x;
      fh = fx;
      if (side == -1) fl /= 2;
      side = -1;
    } else {
      lo = x;
      fl = fx;
      if (side == 1) fh /= 2;
      side = 1;
    }
  }
  return 0.5 * (lo + hi);
}

// src/fastkpm.cpp:418-432
double filling_to_mu(Vec<double> const& gamma, EnergyScale const& es, double kT, double filling,
                     double delta_filling) {
  kT = std::max(kT, 0.1 * es.mag() / gamma.size());
  const double c = kT * std::log(1.0 / filling - 1.0);
  auto f1 = [&](double x) { return mu_to_filling(gamma, es, kT, x) - (filling + delta_filling); };
  auto f2 = [&](double x) { return mu_to_filling(gamma, es, kT, x) - (filling - delta_filling); };
  const double a = es.lo - std::max(c, 0.0), b = es.hi + std::max(-c, 0.0);
  if (delta_filling == 0) return root_solver(f1, a, b);
  return 0.5 * (root_solver(f1, a, b) + root_solver(f2, a, b));
}

// src/fastkpm.cpp:433-438
double filling_to_mu(Vec<double> const& evals, double kT, double filling) {
  if (!(kT > 0)) throw std::invalid_argument("filling_to_mu() requires thermal smearing!");
  auto f = [&](double x) { return mu_to_filling(evals, kT, x) - filling; };
  const auto mm = std::minmax_element(evals.begin(), evals.end());
  return root_solver(f, *mm.first, *mm.second);
}

// src/fastkpm.cpp:440-443
double electronic_grand_energy(Vec<double> const& gamma, EnergyScale const& es, double kT,
                               double mu) {
  return density_product(gamma, [&](double x) { return fermi_energy(x, kT, mu); }, es);
}

// src/fastkpm.cpp:444-450
double electronic_grand_energy(Vec<double> const& evals, double kT, double mu) {
  double acc = 0;
  for (double x : evals) acc += fermi_energy(x, kT, mu);
  return acc;
}

// src/fastkpm.cpp:452-456
double electronic_energy(Vec<double> const& gamma, EnergyScale const& es, double kT,
                         double filling, double mu) {
  const double n_tot = density_product(gamma, [](double) { return 1.0; }, es);
  return electronic_grand_energy(gamma, es, kT, mu) + mu * filling * n_tot;
}

// src/fastkpm.cpp:457-476
double electronic_energy(Vec<double> const& evals, double kT, double filling) {
  if (kT == 0) {
    Vec<double> s = evals;
    std::sort(s.begin(), s.end());
    const int n_occ = int(filling * evals.size() + 0.5);
    assert(n_occ >= 0 && n_occ <= (int)evals.size());
    double acc = 0;
    for (int i = 0; i < n_occ; i++) acc += s[i];
    return acc;
  }
  const double n_occ = filling * evals.size();
  const double mu = filling_to_mu(evals, kT, filling);
  return electronic_grand_energy(evals, kT, mu) + mu * n_occ;
}

}  // namespace fkpm
