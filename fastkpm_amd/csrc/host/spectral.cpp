// Spectral math of the KPM (host, FP64): kernels, Chebyshev arrays, expansion
// coefficients, moment transforms, Fermi functions and energies.  Formulas
// follow src/fastkpm.cpp of the reference (cited per function).  The root
// finder of filling_to_mu is a bracketing Illinois (modified regula falsi)
// solver of this library; the reference uses Boost's TOMS748 with the same
// bracket and a 30-bit relative tolerance.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

#include "fastkpm.h"

namespace fkpm {

namespace {

using cd = std::complex<double>;

// In-place DFT of a power-of-two length n, a_k <- sum_j a_j exp(sign 2 pi i j k / n),
// iterative radix 2; the bit-reversal permutation and the twiddle factors
// (each evaluated directly, not by recurrence) are built once per length.
class Pow2Fft {
 public:
  explicit Pow2Fft(size_t n) : n_(n), rev_(n), tw_(n / 2) {
    for (size_t i = 1, j = 0; i < n; i++) {
      size_t bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      rev_[i] = j;
    }
    for (size_t k = 0; k < n / 2; k++) tw_[k] = std::polar(1.0, -2.0 * Pi * (double)k / (double)n);
  }
  size_t size() const { return n_; }
  void run(std::vector<cd>& a, int sign) const {
    for (size_t i = 1; i < n_; i++)
      if (i < rev_[i]) std::swap(a[i], a[rev_[i]]);
    for (size_t len = 2; len <= n_; len <<= 1) {
      const size_t half = len / 2, step = n_ / len;
      for (size_t i = 0; i < n_; i += len)
        for (size_t k = 0; k < half; k++) {
          const cd w = sign < 0 ? tw_[k * step] : std::conj(tw_[k * step]);
          const cd u = a[i + k], v = a[i + k + half] * w;
          a[i + k] = u + v;
          a[i + k + half] = u - v;
        }
    }
  }

 private:
  size_t n_;
  std::vector<size_t> rev_;
  std::vector<cd> tw_;
};

// DFT of any length n > 0, X_k = sum_j x_j exp(sign 2 pi i j k / n): the
// radix-2 FFT itself for a power of two, otherwise Bluestein's chirp-z
// identity j k = (j^2 + k^2 - (k - j)^2) / 2, which makes it a cyclic
// convolution of length L >= 2n - 1 (a power of two) with the chirp, whose
// transform is kept in the plan (two FFTs of length L per call).
class Dft {
 public:
  Dft(size_t n, int sign) : n_(n), sign_(sign), fft_(pow2_for(n)) {
    if (pow2(n)) return;
    const size_t L = fft_.size();
    c_.resize(n);  // exp(sign pi i j^2 / n), j^2 reduced mod 2n
    for (size_t j = 0; j < n; j++) {
      const unsigned long long r = ((unsigned long long)j * j) % (2ull * n);
      c_[j] = std::polar(1.0, sign * Pi * (double)r / (double)n);
    }
    bhat_.assign(L, cd(0.0, 0.0));
    bhat_[0] = std::conj(c_[0]);
    for (size_t m = 1; m < n; m++) bhat_[m] = bhat_[L - m] = std::conj(c_[m]);
    fft_.run(bhat_, -1);
    for (cd& v : bhat_) v /= (double)L;  // the inverse transform's 1/L, folded in once
  }
  // x holds n values on entry and the transform on return.
  void run(std::vector<cd>& x) const {
    if (c_.empty()) {
      fft_.run(x, sign_);
      return;
    }
    const size_t L = fft_.size();
    std::vector<cd>& a = work_;
    a.assign(L, cd(0.0, 0.0));
    for (size_t j = 0; j < n_; j++) a[j] = x[j] * c_[j];
    fft_.run(a, -1);
    for (size_t k = 0; k < L; k++) a[k] *= bhat_[k];
    fft_.run(a, +1);
    for (size_t k = 0; k < n_; k++) x[k] = c_[k] * a[k];
  }

 private:
  static bool pow2(size_t n) { return (n & (n - 1)) == 0; }
  static size_t pow2_for(size_t n) {
    if (pow2(n)) return n;
    size_t L = 1;
    while (L < 2 * n - 1) L <<= 1;
    return L;
  }
  size_t n_;
  int sign_;
  Pow2Fft fft_;
  std::vector<cd> c_, bhat_;
  mutable std::vector<cd> work_;
};

// DCT-II and DST-II of n real values (the FFTW REDFT10 and RODFT10 conventions):
//   yc_k = 2 sum_j x_j cos(pi (j + 1/2) k / n),
//   ys_k = 2 sum_j x_j sin(pi (j + 1/2) (k + 1) / n),   k = 0 .. n-1,
// both from one DFT X of length 2n of x padded with zeros:
//   yc_k = 2 Re(exp(-i pi k / 2n) X_k),  ys_k = -2 Im(exp(-i pi (k+1) / 2n) X_{k+1}).
class Dct2Dst2 {
 public:
  explicit Dct2Dst2(size_t n) : n_(n), dft_(2 * n, -1), ph_(n + 1) {
    for (size_t k = 0; k <= n; k++) ph_[k] = std::polar(1.0, -Pi * (double)k / (2.0 * n));
  }
  void run(const std::vector<double>& x, std::vector<double>* yc, std::vector<double>* ys) const {
    std::vector<cd>& z = work_;
    z.assign(2 * n_, cd(0.0, 0.0));
    for (size_t j = 0; j < n_; j++) z[j] = x[j];
    dft_.run(z);
    if (yc) {
      yc->resize(n_);
      for (size_t k = 0; k < n_; k++) (*yc)[k] = 2.0 * (ph_[k] * z[k]).real();
    }
    if (ys) {
      ys->resize(n_);
      for (size_t k = 0; k < n_; k++) (*ys)[k] = -2.0 * (ph_[k + 1] * z[k + 1]).imag();
    }
  }

 private:
  size_t n_;
  Dft dft_;
  std::vector<cd> ph_;
  mutable std::vector<cd> work_;
};

// Chebyshev synthesis at the Mq Chebyshev nodes (a DCT-III) of M <= Mq complex
// coefficients:
//   g_i = sum_{m < M} a_m T_m(x_i) = sum_m a_m cos(pi m (i + 1/2) / Mq),  i < Mq.
// For real r, sum_m r_m cos(pi m (2i + 1) / 2Mq) = Re sum_m (r_m e^{i pi m / 2Mq}) e^{2 pi i m i / 2Mq},
// so the real and imaginary parts of a take one DFT of length 2Mq each.
class ChebSynthesis {
 public:
  ChebSynthesis(size_t M, size_t Mq) : M_(M), Mq_(Mq), dft_(2 * Mq, +1), ph_(M) {
    for (size_t m = 0; m < M; m++) ph_[m] = std::polar(1.0, Pi * (double)m / (2.0 * Mq));
  }
  void run(const cd* a, std::vector<cd>& out) const {
    out.assign(Mq_, cd(0.0, 0.0));
    std::vector<cd>& z = work_;
    for (int part = 0; part < 2; part++) {
      z.assign(2 * Mq_, cd(0.0, 0.0));
      bool any = false;
      for (size_t m = 0; m < M_; m++) {
        const double r = part == 0 ? a[m].real() : a[m].imag();
        any |= r != 0.0;
        z[m] = r * ph_[m];
      }
      if (!any) continue;
      dft_.run(z);
      for (size_t i = 0; i < Mq_; i++)
        out[i] += part == 0 ? cd(z[i].real(), 0.0) : cd(0.0, z[i].real());
    }
  }

 private:
  size_t M_, Mq_;
  Dft dft_;
  std::vector<cd> ph_;
  mutable std::vector<cd> work_;
};

}  // namespace

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

// src/fastkpm.cpp:80-233, the branch the reference takes with FFTW, on this
// library's own transforms (Dct2Dst2 above):
// - static (omega = 0): with f_i the Fermi density at node x_i, t_i = 1 - x_i^2
//   and nodes t_i < 1e-4 left out, DCT-II and DST-II of x1_i = f_i x_i / t_i^1.5
//   and x2_i = f_i / t_i give every C_{m1,m2} from the sums and differences of
//   m1, m2 (Mq >= 2M keeps m1 + m2 inside the transforms), times
//   pi (2 - delta_m1,0)(2 - delta_m2,0) g_m1 g_m2 / (2 Mq mag^2);
// - optical: for each m1 one DCT-II over the nodes i >= i_start of
//   T_m1(x_i + omega) (f(x_i) - f(x_i + omega)) / (2 omega sqrt(1 - (x_i + omega)^2)),
//   times 2 pi (2 - delta_m1,0)(2 - delta_m2,0) g_m1 g_m2 / (Mq mag^2).
// O(Mq log Mq + M^2) and O(M Mq log Mq) instead of the direct node sums'
// O(M^2 Mq); the unit is e^2/h (the factor 2 pi).
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
  const double mag2 = es.mag() * es.mag();
  if (ws < 1.0e-10) {
    std::vector<double> x1(Mq, 0.0), x2(Mq, 0.0);
    for (int i = 0; i < Mq; i++) {
      const double x_i = std::cos(Pi * (i + 0.5) / Mq);
      const double t = 1.0 - x_i * x_i;
      if (t < cutoff) continue;
      const double f_i = fermi_density(es.unscale(x_i), kT, mu);
      x1[i] = f_i * x_i / std::pow(t, 1.5);
      x2[i] = f_i / t;
    }
    std::vector<double> yc1, yc2, ys1, ys2;
    const Dct2Dst2 tr(Mq);
    tr.run(x1, &yc1, &ys1);
    tr.run(x2, &yc2, &ys2);
    auto ce = [](const std::vector<double>& y, int k) { return y[k >= 0 ? k : -k]; };
    auto se = [](const std::vector<double>& y, int k) {
      return k > 0 ? y[k - 1] : (k < 0 ? -y[-k - 1] : 0.0);
    };
    for (int m1 = 0; m1 < M; m1++) {
      const double t1 = Pi * (m1 == 0 ? 1.0 : 2.0) * kernel[m1] / (2.0 * Mq * mag2);
      for (int m2 = 0; m2 < M; m2++) {
        const double pref = t1 * (m2 == 0 ? 1.0 : 2.0) * kernel[m2];
        const int ms = m1 + m2, md = m1 - m2;
        const double y_re = 2.0 * yc1[ms] + 2.0 * ce(yc1, md) + (ms > 0 ? ms * ys2[ms - 1] : 0.0) +
                            md * se(ys2, md);
        const double y_im = 2.0 * se(ys1, md) - md * (yc2[ms] + ce(yc2, md));
        ret[m1][m2] = cx_double(pref * y_re, pref * y_im);
      }
    }
    return ret;
  }
  const int i_start = (int)std::ceil(std::acos(1.0 - ws) / Pi * Mq - 0.5);
  if (Mq - i_start < 20)
    throw std::invalid_argument("conductivity: fewer than 20 nodes below 1 - omega");
  std::vector<double> w(Mq, 0.0), acw(Mq, 0.0);  // node weights and arccos(x_i + omega)
  std::vector<char> use(Mq, 0);
  for (int i = i_start; i < Mq; i++) {
    const double x_i = std::cos(Pi * (i + 0.5) / Mq);
    const double xw = x_i + ws;
    if (x_i < -1.0 + cutoff || xw > 1.0 - cutoff) continue;
    use[i] = 1;
    acw[i] = std::acos(xw);
    w[i] = 0.5 * (fermi_density(es.unscale(x_i), kT, mu) -
                  fermi_density(es.unscale(x_i) + omega, kT, mu)) /
           (ws * std::sqrt(1.0 - xw * xw));
  }
  std::vector<double> X(Mq), Y;
  const Dct2Dst2 tr(Mq);
  for (int m1 = 0; m1 < M; m1++) {
    for (int i = 0; i < Mq; i++) X[i] = use[i] ? std::cos(m1 * acw[i]) * w[i] : 0.0;
    tr.run(X, &Y, nullptr);
    const double t1 = 2.0 * Pi * (m1 == 0 ? 1.0 : 2.0) * kernel[m1] / (Mq * mag2);
    for (int m2 = 0; m2 < M; m2++)
      ret[m1][m2] = cx_double(Y[m2] * t1 * (m2 == 0 ? 1.0 : 2.0) * kernel[m2], 0.0);
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
// evaluated as Chebyshev syntheses (DCT-III, ChebSynthesis above) along m2
// for every m1 and then along m1 for every node j: O(M Mq log Mq + Mq^2 log Mq)
// instead of the reference's O(Mq^2 M^2) quadruple loop.
Vec<Vec<cx_double>> moment_transform(Vec<Vec<cx_double>> const& moments, int Mq,
                                     Vec<double> const& kernel) {
  const int M = (int)moments.size();
  if ((int)kernel.size() != M) throw std::invalid_argument("moment_transform: kernel size != M");
  if (Mq < M) throw std::invalid_argument("moment_transform: Mq must be >= M");
  const ChebSynthesis syn(M, Mq);
  std::vector<cd> row(M);
  Vec<std::vector<cd>> G(M);  // G[m1][j] = sum_m2 mup[m1][m2] T_m2(x_j)
  for (int a = 0; a < M; a++) {
    for (int b = 0; b < M; b++)
      row[b] = (a == 0 ? 1.0 : 2.0) * (b == 0 ? 1.0 : 2.0) * kernel[a] * kernel[b] * moments[a][b];
    syn.run(row.data(), G[a]);
  }
  Vec<Vec<cx_double>> gamma(Mq, Vec<cx_double>(Mq, cx_double(0.0, 0.0)));
  std::vector<cd> col(M), g;
  for (int j = 0; j < Mq; j++) {
    for (int a = 0; a < M; a++) col[a] = G[a][j];
    syn.run(col.data(), g);
    for (int i = 0; i < Mq; i++) gamma[i][j] = g[i];
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
