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

// src/fastkpm.cpp:80-233, the direct-sum branch (no FFTW in this build):
// static (omega = 0) Kubo-Bastin coefficients from the Chebyshev nodes with
// 1 - x^2 >= 1e-4, or the optical ones from the nodes i >= i_start; both
// times (2 - delta_m1,0)(2 - delta_m2,0) g_m1 g_m2 * 2 pi / (Mq mag^2).
Vec<Vec<cx_double>> electrical_conductivity_coefficients(int M, int Mq, double kT, double mu,
                                                         double omega, EnergyScale es,
                                                         Vec<double> const& kernel) {
  if ((int)kernel.size() != M) throw std::invalid_argument("conductivity: kernel size != M");
  if (Mq < 2 * M) throw std::invalid_argument("conductivity: Mq must be >= 2 M");
  if (omega < 0.0) throw std::invalid_argument("conductivity: omega must be >= 0");
  constexpr double cutoff = 1e-4;
  const double ws = omega / es.mag();
  Vec<Vec<cx_double>> ret(M, Vec<cx_double>(M, cx_double(0.0, 0.0)));
  if (ws >= 2.0) return ret;
  Vec<double> Ti(M), Tj(M);
  if (ws < 1.0e-10) {
    for (int i = 0; i < Mq; i++) {
      const double x_i = std::cos(Pi * (i + 0.5) / Mq);
      if (1.0 - x_i * x_i < cutoff) continue;
      chebyshev_fill_array(x_i, Ti);     // T_m(x) = cos(m t)
      chebyshev_fill_array(x_i, Tj, 2);  // U_m(x)
      const double s1 = std::sqrt(1.0 - x_i * x_i);
      for (int m = M - 1; m > 0; m--) Tj[m] = Tj[m - 1] * s1;  // sin(m t)
      Tj[0] = 0.0;
      const double f_i = fermi_density(es.unscale(x_i), kT, mu) / std::pow(1.0 - x_i * x_i, 1.5);
      for (int m1 = 0; m1 < M; m1++)
        for (int m2 = 0; m2 < M; m2++)
          ret[m1][m2] += (Ti[m1] * cx_double(Ti[m2], -Tj[m2]) * cx_double(x_i, m2 * s1) +
                          Ti[m2] * cx_double(Ti[m1], Tj[m1]) * cx_double(x_i, -m1 * s1)) *
                         f_i;
    }
  } else {
    const int i_start = (int)std::ceil(std::acos(1.0 - ws) / Pi * Mq - 0.5);
    if (Mq - i_start < 20)
      throw std::invalid_argument("conductivity: fewer than 20 nodes below 1 - omega");
    for (int i = i_start; i < Mq; i++) {
      const double x_i = std::cos(Pi * (i + 0.5) / Mq);
      const double xw = x_i + ws;
      if (x_i < -1.0 + cutoff || xw > 1.0 - cutoff) continue;
      chebyshev_fill_array(xw, Ti);   // T_m(x + omega)
      chebyshev_fill_array(x_i, Tj);  // T_m(x)
      const double s2 = std::sqrt(1.0 - xw * xw);
      const double f_i =
          (fermi_density(es.unscale(x_i), kT, mu) - fermi_density(es.unscale(x_i) + omega, kT, mu)) /
          (ws * s2);
      for (int m1 = 0; m1 < M; m1++)
        for (int m2 = 0; m2 < M; m2++) ret[m1][m2] += Ti[m1] * Tj[m2] * f_i;
    }
  }
  for (int m1 = 0; m1 < M; m1++) {
    // the factor 2 pi changes the unit from e^2/hbar to e^2/h
    const double t1 = 2.0 * Pi * (m1 == 0 ? 1.0 : 2.0) * kernel[m1] / (Mq * es.mag() * es.mag());
    for (int m2 = 0; m2 < M; m2++) ret[m1][m2] *= t1 * (m2 == 0 ? 1.0 : 2.0) * kernel[m2];
  }
  return ret;
}

// src/fastkpm.cpp:300-312
cx_double moment_product(Vec<Vec<cx_double>> const& c, Vec<Vec<cx_double>> const& mu) {
  if (c.size() != mu.size() || (c.size() > 0 && c[0].size() != mu[0].size()))
    throw std::invalid_argument("moment_product: shapes differ");
  cx_double r(0.0, 0.0);
  for (size_t a = 0; a < c.size(); a++)
    for (size_t b = 0; b < c[a].size(); b++) r += c[a][b] * mu[a][b];
  return r;
}

// src/fastkpm.cpp:258-289: gamma[i][j] = sum T_m1(x_i) T_m2(x_j) mup[m1][m2],
// evaluated as two matrix products over the node table T[i][m] = T_m(x_i).
Vec<Vec<cx_double>> moment_transform(Vec<Vec<cx_double>> const& moments, int Mq,
                                     Vec<double> const& kernel) {
  const int M = (int)moments.size();
  if ((int)kernel.size() != M) throw std::invalid_argument("moment_transform: kernel size != M");
  Vec<Vec<cx_double>> mup(M, Vec<cx_double>(M));
  for (int a = 0; a < M; a++)
    for (int b = 0; b < M; b++)
      mup[a][b] = (a == 0 ? 1.0 : 2.0) * (b == 0 ? 1.0 : 2.0) * kernel[a] * kernel[b] * moments[a][b];
  Vec<Vec<double>> T(Mq, Vec<double>(M));
  for (int i = 0; i < Mq; i++) chebyshev_fill_array(std::cos(Pi * (i + 0.5) / Mq), T[i]);
  Vec<Vec<cx_double>> G(Mq, Vec<cx_double>(M, cx_double(0.0, 0.0)));  // G = T mup
  for (int i = 0; i < Mq; i++)
    for (int a = 0; a < M; a++)
      for (int b = 0; b < M; b++) G[i][b] += T[i][a] * mup[a][b];
  Vec<Vec<cx_double>> gamma(Mq, Vec<cx_double>(Mq, cx_double(0.0, 0.0)));  // gamma = G T^t
  for (int i = 0; i < Mq; i++)
    for (int j = 0; j < Mq; j++) {
      cx_double g(0.0, 0.0);
      for (int b = 0; b < M; b++) g += G[i][b] * T[j][b];
      gamma[i][j] = g;
    }
  return gamma;
}

// src/fastkpm.cpp:335-360
void density_function(Vec<Vec<cx_double>> const& gamma, EnergyScale es, Vec<double>& x,
                      Vec<double>& y, Vec<Vec<cx_double>>& rho) {
  const int Mq1 = (int)gamma.size();
  const int Mq2 = Mq1 > 0 ? (int)gamma[0].size() : 0;
  x.resize(Mq1);
  y.resize(Mq2);
  rho.assign(Mq1, Vec<cx_double>(Mq2, cx_double(0.0, 0.0)));
  for (int i2 = Mq2 - 1; i2 >= 0; i2--) y[Mq2 - 1 - i2] = es.unscale(std::cos(Pi * (i2 + 0.5) / Mq2));
  for (int i1 = Mq1 - 1; i1 >= 0; i1--) {
    const double x1 = std::cos(Pi * (i1 + 0.5) / Mq1);
    x[Mq1 - 1 - i1] = es.unscale(x1);
    for (int i2 = Mq2 - 1; i2 >= 0; i2--) {
      const double x2 = std::cos(Pi * (i2 + 0.5) / Mq2);
      rho[Mq1 - 1 - i1][Mq2 - 1 - i2] =
          gamma[i1][i2] /
          (Pi * std::sqrt(1.0 - x1 * x1) * Pi * std::sqrt(1.0 - x2 * x2) * es.mag() * es.mag());
    }
  }
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
