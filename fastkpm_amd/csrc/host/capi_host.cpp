// C entry points of include/fastkpm_host.h over the C++ host code.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fastkpm.h"
#include "fastkpm_hip.h"
#include "fastkpm_host.h"

namespace fkpm {
// probing phases (defined with the Engine<T> base class, engine.cpp)
template <typename T>
T random_phase(RNG& rng);
template <>
cx_double random_phase<cx_double>(RNG& rng);
}  // namespace fkpm

namespace {

// An engine that only generates probing vectors: it runs the Engine<T> base
// class code of set_R_* and keeps R on the host.
template <typename T>
class RGen : public fkpm::Engine<T> {
 public:
  fkpm::EnergyScale energy_scale(fkpm::SpMatBsr<T> const&, double, int) override {
    throw std::logic_error("RGen");
  }
  void set_H(fkpm::SpMatBsr<T> const&, fkpm::EnergyScale const&) override {
    throw std::logic_error("RGen");
  }
  fkpm::Vec<double> moments(int) override { throw std::logic_error("RGen"); }
  fkpm::Vec<fkpm::Vec<fkpm::cx_double>> moments2_v1(int, fkpm::SpMatBsr<T> const&,
                                                    fkpm::SpMatBsr<T> const&, int, int) override {
    throw std::logic_error("RGen");
  }
  fkpm::Vec<fkpm::Vec<fkpm::cx_double>> moments2_v2(int, fkpm::SpMatBsr<T> const&,
                                                    fkpm::SpMatBsr<T> const&, int, int) override {
    throw std::logic_error("RGen");
  }
  void stoch_matrix(fkpm::Vec<double> const&, fkpm::SpMatBsr<T>&) override {
    throw std::logic_error("RGen");
  }
  void autodiff_matrix(fkpm::Vec<double> const&, fkpm::SpMatBsr<T>&) override {
    throw std::logic_error("RGen");
  }
};

template <typename T>
void copy_R(RGen<T> const& g, void* out) {
  std::memcpy(out, g.R.memptr(), g.R.size() * sizeof(T));
}

thread_local std::string g_host_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (std::exception const& e) {
    g_host_error = e.what();
    return FKPM_E_ARG;
  } catch (...) {
    g_host_error = "unknown exception";
    return FKPM_E_ARG;
  }
}

template <typename T>
int set_R_id(int n, int j0, int j1, void* out) {
  RGen<T> g;
  g.set_R_identity(n, j0, j1);
  copy_R(g, out);
  return 0;
}

template <typename T>
int set_R_unc(int n, int s, void* rng, int j0, int j1, void* out) {
  RGen<T> g;
  g.set_R_uncorrelated(n, s, *static_cast<fkpm::RNG*>(rng), j0, j1);
  copy_R(g, out);
  return 0;
}

template <typename T>
int set_R_cor(int n, const int* groups, void* rng, int j0, int j1, void* out, int* s_out) {
  RGen<T> g;
  fkpm::Vec<int> gr(groups, groups + n);
  if (j0 < 0)
    g.set_R_correlated(gr, *static_cast<fkpm::RNG*>(rng));
  else
    g.set_R_correlated(gr, *static_cast<fkpm::RNG*>(rng), j0, j1);
  copy_R(g, out);
  if (s_out) *s_out = (int)g.R.n_cols;
  return 0;
}

template <template <typename> class F, typename... A>
int by_dtype(int dtype, A... a) {
  switch (dtype) {
    case FKPM_F32: return F<float>::run(a...);
    case FKPM_F64: return F<double>::run(a...);
    case FKPM_C32: return F<fkpm::cx_float>::run(a...);
    case FKPM_C64: return F<fkpm::cx_double>::run(a...);
  }
  return FKPM_E_ARG;
}

template <typename T>
struct IdF {
  static int run(int n, int j0, int j1, void* out) { return set_R_id<T>(n, j0, j1, out); }
};
template <typename T>
struct UncF {
  static int run(int n, int s, void* rng, int j0, int j1, void* out) {
    return set_R_unc<T>(n, s, rng, j0, j1, out);
  }
};
template <typename T>
struct CorF {
  static int run(int n, const int* groups, void* rng, int j0, int j1, void* out, int* s_out) {
    return set_R_cor<T>(n, groups, rng, j0, j1, out, s_out);
  }
};

}  // namespace

extern "C" {

const char* fkpm_host_last_error(void) { return g_host_error.c_str(); }

void* fkpm_host_rng_new(uint32_t seed) { return new fkpm::RNG(seed); }
void fkpm_host_rng_free(void* rng) { delete static_cast<fkpm::RNG*>(rng); }
uint32_t fkpm_host_rng_next(void* rng) { return (uint32_t)(*static_cast<fkpm::RNG*>(rng))(); }
void fkpm_host_rng_discard(void* rng, unsigned long long k) {
  static_cast<fkpm::RNG*>(rng)->discard(k);
}

int fkpm_host_set_R_identity(int dtype, int n, int j0, int j1, void* R_out) {
  int rc = 0;
  int g = guarded([&] { rc = by_dtype<IdF>(dtype, n, j0, j1, R_out); });
  return g ? g : rc;
}

int fkpm_host_set_R_uncorrelated(int dtype, int n, int s, void* rng, int j0, int j1, void* R_out) {
  int rc = 0;
  int g = guarded([&] { rc = by_dtype<UncF>(dtype, n, s, rng, j0, j1, R_out); });
  return g ? g : rc;
}

int fkpm_host_set_R_correlated(int dtype, int n, const int* groups, void* rng, int j0, int j1,
                               void* R_out, int* s_out) {
  int rc = 0;
  int g = guarded([&] { rc = by_dtype<CorF>(dtype, n, groups, rng, j0, j1, R_out, s_out); });
  return g ? g : rc;
}

int fkpm_host_set_R_probing(int n, const int* groups, void* rng, int* rcol_out, void* rval_out,
                            int* s_out) {
  return guarded([&] {
    if (n <= 0 || !groups || !rng || !rcol_out || !rval_out)
      throw std::invalid_argument("set_R_probing: bad arguments");
    const auto mm = std::minmax_element(groups, groups + n);
    if (*mm.first != 0) throw std::invalid_argument("set_R_correlated: colours must start at 0");
    const int s = *mm.second + 1;
    fkpm::cx_double* rv = static_cast<fkpm::cx_double*>(rval_out);
    if (s >= n) {  // identity (Engine::set_R_correlated, src/engine.cpp:103-116)
      for (int i = 0; i < n; i++) {
        rcol_out[i] = i;
        rv[i] = fkpm::cx_double(1.0, 0.0);
      }
      if (s_out) *s_out = n;
      return;
    }
    // column j draws its phases from RNG(rng()) in row order over its rows
    // (src/engine.cpp:76-89), rows bucketed by colour once
    std::vector<int> cnt(s + 1, 0);
    for (int i = 0; i < n; i++) cnt[groups[i] + 1]++;
    for (int k = 1; k <= s; k++) cnt[k] += cnt[k - 1];
    std::vector<int> rows(n), pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; i++) rows[pos[groups[i]]++] = i;
    fkpm::RNG& master = *static_cast<fkpm::RNG*>(rng);
    for (int j = 0; j < s; j++) {
      fkpm::RNG rng_j(master());
      for (int k = cnt[j]; k < cnt[j + 1]; k++) {
        rcol_out[rows[k]] = j;
        rv[rows[k]] = fkpm::random_phase<fkpm::cx_double>(rng_j);
      }
    }
    if (s_out) *s_out = s;
  });
}

int fkpm_host_conductivity_coefficients(int M, int Mq, double kT, double mu, double omega,
                                        double lo, double hi, const double* kernel, double* out) {
  return guarded([&] {
    if (M <= 0 || !kernel || !out) throw std::invalid_argument("conductivity: bad arguments");
    fkpm::Vec<double> k(kernel, kernel + M);
    const auto c = fkpm::electrical_conductivity_coefficients(M, Mq, kT, mu, omega,
                                                              fkpm::EnergyScale{lo, hi}, k);
    for (int a = 0; a < M; a++)
      for (int b = 0; b < M; b++) {
        out[((size_t)a * M + b) * 2] = c[a][b].real();
        out[((size_t)a * M + b) * 2 + 1] = c[a][b].imag();
      }
  });
}

int fkpm_host_moment_transform2(int M, const double* mu, int Mq, const double* kernel,
                                double* gamma_out) {
  return guarded([&] {
    if (M <= 0 || Mq <= 0 || !mu || !kernel || !gamma_out)
      throw std::invalid_argument("moment_transform2: bad arguments");
    fkpm::Vec<fkpm::Vec<fkpm::cx_double>> m(M, fkpm::Vec<fkpm::cx_double>(M));
    for (int a = 0; a < M; a++)
      for (int b = 0; b < M; b++)
        m[a][b] = fkpm::cx_double(mu[((size_t)a * M + b) * 2], mu[((size_t)a * M + b) * 2 + 1]);
    const auto g = fkpm::moment_transform(m, Mq, fkpm::Vec<double>(kernel, kernel + M));
    for (int i = 0; i < Mq; i++)
      for (int j = 0; j < Mq; j++) {
        gamma_out[((size_t)i * Mq + j) * 2] = g[i][j].real();
        gamma_out[((size_t)i * Mq + j) * 2 + 1] = g[i][j].imag();
      }
  });
}

int fkpm_host_expansion_coefficients(int M, int Mq, int kind, double kT, double mu, double lo,
                                     double hi, double* c_out) {
  return guarded([&] {
    std::function<double(double)> f;
    if (kind == 0)
      f = [=](double x) { return fkpm::fermi_energy(x, kT, mu); };
    else if (kind == 1)
      f = [=](double x) { return fkpm::fermi_density(x, kT, mu); };
    else
      f = [](double) { return 1.0; };
    const auto c = fkpm::expansion_coefficients(M, Mq, f, fkpm::EnergyScale{lo, hi});
    std::memcpy(c_out, c.data(), sizeof(double) * M);
  });
}

void fkpm_host_jackson_kernel(int M, double* g_out) {
  const auto g = fkpm::jackson_kernel(M);
  std::memcpy(g_out, g.data(), sizeof(double) * M);
}

double fkpm_host_moment_product(int M, const double* c, const double* mu) {
  return fkpm::moment_product(fkpm::Vec<double>(c, c + M), fkpm::Vec<double>(mu, mu + M));
}

void fkpm_host_moment_transform(int M, const double* mu, int Mq, double* gamma_out) {
  const auto g = fkpm::moment_transform(fkpm::Vec<double>(mu, mu + M), Mq);
  std::memcpy(gamma_out, g.data(), sizeof(double) * Mq);
}

double fkpm_host_electronic_grand_energy(int Mq, const double* gamma, double lo, double hi,
                                         double kT, double mu) {
  return fkpm::electronic_grand_energy(fkpm::Vec<double>(gamma, gamma + Mq),
                                       fkpm::EnergyScale{lo, hi}, kT, mu);
}

double fkpm_host_filling_to_mu(int Mq, const double* gamma, double lo, double hi, double kT,
                               double filling, double delta_filling) {
  return fkpm::filling_to_mu(fkpm::Vec<double>(gamma, gamma + Mq), fkpm::EnergyScale{lo, hi}, kT,
                             filling, delta_filling);
}

}  // extern "C"
