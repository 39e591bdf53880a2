// The reference-side binding of the C-ABI: an fkpm::Engine<T> whose device
// calls forward to include/fastkpm_hip.h, written the way a maintainer of the
// reference would add it beside Engine_cuSPARSE (src/engine_cusparse.cpp:
// 486-1040) and return it from mk_engine_cuSPARSE.  The reference's
// SpMatBsr::n_rows counts block rows (src/fastkpm.h:66-74, used so at
// src/engine_cusparse.cpp:705-727), so it is passed as the block-row count
// unchanged.  This file is compiled against this library's copy of the
// interface (include/fastkpm.h) and driven by tests/test_cpp_api.py.
//
//   ref_adapter TYPE DIR    TYPE = cx_double | cx_float | double | float;
//                           runs energy scale, correlated R, set_H, moments,
//                           autodiff_matrix and stoch_matrix on a 2-orbital
//                           lattice through the adapter and writes H, R, mu
//                           and both D's to DIR
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fastkpm.h"
#include "fastkpm_hip.h"

namespace fkpm {

template <typename T> struct HipDtype;
template <> struct HipDtype<float> { static const int v = FKPM_F32; };
template <> struct HipDtype<double> { static const int v = FKPM_F64; };
template <> struct HipDtype<cx_float> { static const int v = FKPM_C32; };
template <> struct HipDtype<cx_double> { static const int v = FKPM_C64; };

template <typename T>
class Engine_HipAbi : public Engine<T> {
  fkpm_handle* h = nullptr;
  EnergyScale es;
  static void ok(int rc) {
    if (rc) throw std::runtime_error(fkpm_last_error());
  }
  void set_D(SpMatBsr<T> const& D) {
    ok(fkpm_set_D_pattern(h, D.n_rows, (int)D.col_idx.size(), D.row_ptr.data(), D.row_idx.data(),
                          D.col_idx.data()));
  }

 public:
  explicit Engine_HipAbi(int device) {
    ok(fkpm_dev_open(device, HipDtype<T>::v, FKPM_PREC_FP64, &h));
  }
  ~Engine_HipAbi() override { fkpm_dev_close(h); }

  EnergyScale energy_scale(SpMatBsr<T> const& H, double extend, int iters) override {
    // the reference starts Lanczos from arma::randn; any normal start vector
    std::mt19937 g(0);
    std::normal_distribution<double> nd;
    std::vector<T> v((size_t)H.n_rows * H.b_len);
    for (auto& x : v) {
      if constexpr (std::is_same<T, cx_double>::value || std::is_same<T, cx_float>::value) {
        const double re = nd(g), im = nd(g);
        x = T(re, im);
      } else {
        x = T(nd(g));
      }
    }
    double lo = 0, hi = 0;
    ok(fkpm_energy_scale(h, H.n_rows, H.b_len, (int)H.col_idx.size(), H.row_ptr.data(),
                         H.col_idx.data(), H.val.data(), extend, iters, v.data(), &lo, &hi));
    return EnergyScale{lo, hi};
  }
  void transfer_R() override {
    ok(fkpm_set_R(h, (int)this->R.n_rows, (int)this->R.n_cols, this->R.memptr()));
  }
  void set_H(SpMatBsr<T> const& H, EnergyScale const& es_) override {
    es = es_;
    ok(fkpm_set_H(h, H.n_rows, H.b_len, (int)H.col_idx.size(), H.row_ptr.data(),
                  H.col_idx.data(), H.val.data(), es.avg(), es.mag()));
  }
  Vec<double> moments(int M) override {
    Vec<double> mu(M);
    ok(fkpm_moments(h, M, mu.data()));
    return mu;
  }
  void stoch_matrix(Vec<double> const& c, SpMatBsr<T>& D) override {
    set_D(D);
    ok(fkpm_stoch_matrix(h, (int)c.size(), c.data(), D.val.data()));
  }
  void autodiff_matrix(Vec<double> const& c, SpMatBsr<T>& D) override {
    // symmetrize and scale(1/mag) happen on the device
    set_D(D);
    ok(fkpm_autodiff_matrix(h, (int)c.size(), c.data(), D.val.data()));
  }
  Vec<Vec<cx_double>> moments2_v1(int M, SpMatBsr<T> const& j1, SpMatBsr<T> const& j2,
                                  int /*a_chunk_ncols*/, int R_chunk_ncols) override {
    Vec<double> mu(2 * (size_t)M * M);  // row-major (re, im)
    ok(fkpm_moments2(h, M, (int)j1.col_idx.size(), j1.row_ptr.data(), j1.col_idx.data(),
                     j1.val.data(), (int)j2.col_idx.size(), j2.row_ptr.data(),
                     j2.col_idx.data(), j2.val.data(), R_chunk_ncols, mu.data()));
    Vec<Vec<cx_double>> r(M, Vec<cx_double>(M));
    for (int a = 0; a < M; a++)
      for (int b = 0; b < M; b++) r[a][b] = {mu[2 * ((size_t)a * M + b)], mu[2 * ((size_t)a * M + b) + 1]};
    return r;
  }
  Vec<Vec<cx_double>> moments2_v2(int, SpMatBsr<T> const&, SpMatBsr<T> const&, int,
                                  int) override {
    std::fprintf(stderr, "moments2_v2 is not implemented on the GPU engine\n");
    std::exit(EXIT_FAILURE);
  }
};

}  // namespace fkpm

using namespace fkpm;

// 2-D periodic L x L lattice with 2 orbitals per site: hopping t*I2 to the
// four neighbours, on-site block [[e_i, g_i], [conj g_i, -e_i]] (g_i = 0 for a
// real element type), deterministic values.
template <typename T>
static SpMatBsr<T> two_orbital(int L, double t) {
  constexpr bool CX = std::is_same<T, cx_double>::value || std::is_same<T, cx_float>::value;
  SpMatElems<T> el(L * L, L * L, 2);
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++) {
      const int i = x + L * y;
      const double e = 1.5 * std::sin(0.7 * i + 0.2);
      const double gr = std::cos(1.1 * i), gi = CX ? std::sin(0.3 * i + 1.0) : 0.0;
      T on[4];
      if constexpr (CX) {
        on[0] = T(e, 0);   // (0,0)
        on[1] = T(gr, -gi);  // (1,0)
        on[2] = T(gr, gi);   // (0,1)
        on[3] = T(-e, 0);  // (1,1)
      } else {
        on[0] = T(e);
        on[1] = T(gr);
        on[2] = T(gr);
        on[3] = T(-e);
      }
      el.add(i, i, on);
      const T hop[4] = {T(t), T(0), T(0), T(t)};
      const int nb[4] = {((x + 1) % L) + L * y, ((x + L - 1) % L) + L * y, x + L * ((y + 1) % L),
                         x + L * ((y + L - 1) % L)};
      for (int k = 0; k < 4; k++) el.add(i, nb[k], hop);
    }
  return SpMatBsr<T>(el);
}

template <typename V>
static void dump(const std::string& path, const V* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(V)));
}

template <typename T>
static int run(const std::string& dir) {
  const int L = 12, M = 128;
  SpMatBsr<T> H = two_orbital<T>(L, -1.0);
  std::shared_ptr<Engine<T>> eng = std::make_shared<Engine_HipAbi<T>>(0);
  const EnergyScale es = eng->energy_scale(H, 0.05, 128);
  // colours: (x mod 2) + 2 (y mod 2), times 2 orbitals
  Vec<int> groups((size_t)L * L * 2);
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++)
      for (int o = 0; o < 2; o++) groups[2 * (x + L * y) + o] = o + 2 * ((x % 2) + 2 * (y % 2));
  RNG rng(2024);
  eng->set_R_correlated(groups, rng);
  eng->transfer_R();
  eng->set_H(H, es);
  const Vec<double> mu = eng->moments(M);
  const Vec<double> cE = expansion_coefficients(
      M, 2 * M, [](double x) { return fermi_energy(x, 0.05, 0.0); }, es);
  const Vec<double> cF = expansion_coefficients(
      M, 2 * M, [](double x) { return fermi_density(x, 0.05, 0.0); }, es);
  SpMatBsr<T> Dg = H, Ds = H;
  eng->autodiff_matrix(cE, Dg);
  eng->stoch_matrix(cF, Ds);
  std::ofstream meta(dir + "/meta.txt");
  meta.precision(17);
  meta << eng->R.n_rows << " " << eng->R.n_cols << " " << M << " " << H.n_rows << " "
       << H.col_idx.size() << " " << es.lo << " " << es.hi << "\n";
  dump(dir + "/H_row_idx.bin", H.row_idx.data(), H.row_idx.size());
  dump(dir + "/H_col_idx.bin", H.col_idx.data(), H.col_idx.size());
  dump(dir + "/H_row_ptr.bin", H.row_ptr.data(), H.row_ptr.size());
  dump(dir + "/H_val.bin", H.val.data(), H.val.size());
  dump(dir + "/R.bin", eng->R.memptr(), (size_t)eng->R.n_rows * eng->R.n_cols);
  dump(dir + "/mu.bin", mu.data(), mu.size());
  dump(dir + "/cE.bin", cE.data(), cE.size());
  dump(dir + "/cF.bin", cF.data(), cF.size());
  dump(dir + "/D_autodiff.bin", Dg.val.data(), Dg.val.size());
  dump(dir + "/D_stoch.bin", Ds.val.data(), Ds.val.size());
  std::printf("adapter OK (%zu blocks, S=%d)\n", H.col_idx.size(), (int)eng->R.n_cols);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s cx_double|cx_float|double|float DIR\n", argv[0]);
    return 2;
  }
  const std::string t = argv[1], dir = argv[2];
  try {
    if (t == "cx_double") return run<cx_double>(dir);
    if (t == "cx_float") return run<cx_float>(dir);
    if (t == "double") return run<double>(dir);
    if (t == "float") return run<float>(dir);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown element type %s\n", t.c_str());
  return 2;
}
