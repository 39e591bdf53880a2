// C entry points of include/fastkpm_host.h over the C++ host code.
#include <cstring>
#include <stdexcept>

#include "fastkpm.h"
#include "fastkpm_hip.h"
#include "fastkpm_host.h"

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

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (...) {
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
