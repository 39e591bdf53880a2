// Drives the fkpm:: C++ API the way an application of the reference does
// (build H with SpMatElems, energy_scale, set_R_correlated, set_H, moments,
// expansion_coefficients, moment_product, autodiff_matrix, stoch_matrix,
// moments2_v1, electrical_conductivity_coefficients).
//
//   engine_flow host          host-only API checks (no device needed)
//   engine_flow device DIR    full flow on HIP device 0; writes H, R, mu and
//                             both D's to DIR for comparison with the oracle
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "fastkpm.h"

using namespace fkpm;

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "CHECK failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

// 2-D periodic Kondo lattice, 2 spin orbitals per site: hopping t*I2, on-site
// -J s_i.sigma with deterministic unit spins s_i.
static SpMatBsr<cx_double> kondo(int L, double t, double J) {
  SpMatElems<cx_double> el(L * L, L * L, 2);
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++) {
      const int i = x + L * y;
      const double th = 0.37 * i + 0.1, ph = 1.13 * i;
      const double sx = std::sin(th) * std::cos(ph), sy = std::sin(th) * std::sin(ph), sz = std::cos(th);
      // column-major 2x2: (0,0), (1,0), (0,1), (1,1)
      const cx_double on[4] = {-J * sz, cx_double(-J * sx, -J * sy), cx_double(-J * sx, J * sy), J * sz};
      el.add(i, i, on);
      const cx_double hop[4] = {t, 0, 0, t};
      const int nb[4] = {((x + 1) % L) + L * y, ((x + L - 1) % L) + L * y, x + L * ((y + 1) % L),
                         x + L * ((y + L - 1) % L)};
      for (int k = 3; k >= 0; k--) el.add(i, nb[k], hop);  // unsorted on purpose
    }
  return SpMatBsr<cx_double>(el);
}

static void host_checks() {
  // build(): sorting and duplicate summing
  SpMatElems<cx_double> el(3, 3, 1);
  const cx_double a = 1.0, b = 2.0, c = cx_double(0, 3);
  el.add(2, 1, &a);
  el.add(0, 2, &b);
  el.add(2, 1, &c);
  el.add(0, 0, &a);
  SpMatBsr<cx_double> M(el);
  CHECK(M.n_blocks() == 3);
  CHECK(M.row_ptr[0] == 0 && M.row_ptr[1] == 2 && M.row_ptr[2] == 2 && M.row_ptr[3] == 3);
  CHECK(M.col_idx[0] == 0 && M.col_idx[1] == 2 && M.col_idx[2] == 1);
  CHECK(*M(2, 1) == cx_double(1, 3));
  bool threw = false;
  try {
    M.find_index(1, 1);
  } catch (std::out_of_range&) {
    threw = true;
  }
  CHECK(threw);
  // symmetrize: Hermitian mean of (i,j), (j,i)
  SpMatElems<cx_double> e2(2, 2, 1);
  const cx_double u = cx_double(1, 1), v = cx_double(3, -5), d = cx_double(2, 7);
  e2.add(0, 1, &u);
  e2.add(1, 0, &v);
  e2.add(1, 1, &d);
  SpMatBsr<cx_double> S(e2);
  S.symmetrize();
  CHECK(std::abs(*S(1, 0) - (v + std::conj(u)) / 2.0) < 1e-15);
  CHECK(std::abs(*S(0, 1) - std::conj(*S(1, 0))) < 1e-15);
  CHECK(std::abs(*S(1, 1) - cx_double(2, 0)) < 1e-15);
  // spectral helpers
  const auto g = jackson_kernel(10);
  CHECK(std::abs(g[0] - 1.0) < 1e-15 && std::abs(g[1] - std::cos(Pi / 11)) < 1e-14);
  EnergyScale es{-2, 6};
  const auto one = expansion_coefficients(16, 32, [](double) { return 1.0; }, es);
  CHECK(std::abs(one[0] - 1) < 1e-14 && std::abs(one[5]) < 1e-14);
  CHECK(std::abs(fermi_energy(0.3, 0.0, 0.5) + 0.2) < 1e-15);
  // no CPU engine
  threw = false;
  try {
    mk_engine_cpu<cx_double>();
  } catch (std::logic_error&) {
    threw = true;
  }
  CHECK(threw);
  // the engine's RNG contract: one master draw per column (engine.cpp:56, :83)
  RNG r1(3), r2(3);
  r2.discard(5);
  (void)r1;
  std::printf("host OK\n");
}

template <typename T>
static void dump(const std::string& path, const T* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), n * sizeof(T));
}

static int device_flow(const std::string& dir) {
  auto eng = mk_engine<cx_double>();
  const int L = 12, M = 120;
  auto H = kondo(L, -1.0, 2.0);
  const int n = H.n_rows * H.b_len;
  const EnergyScale es = eng->energy_scale(H, 0.02);
  // bounds of the exact spectrum are within [-6, 6]; Lanczos + slack covers it
  CHECK(es.lo < -3.0 && es.hi > 3.0 && es.lo > -8.0 && es.hi < 8.0);
  Vec<int> groups(n);
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++)
      for (int o = 0; o < 2; o++) groups[2 * (x + L * y) + o] = o + 2 * ((x % 3) + 3 * (y % 3));
  RNG rng(2024);
  eng->set_R_correlated(groups, rng);
  CHECK((int)eng->R.n_rows == n && (int)eng->R.n_cols == 18);
  eng->set_H(H, es);
  // autodiff before moments violates the precondition and throws
  SpMatBsr<cx_double> D = H;
  bool threw = false;
  try {
    eng->autodiff_matrix(Vec<double>(M, 0.0), D);
  } catch (std::runtime_error&) {
    threw = true;
  }
  CHECK(threw);
  const auto mu = eng->moments(M);
  const auto cE = expansion_coefficients(M, 2 * M, [](double x) { return fermi_energy(x, 0.05, 0.0); }, es);
  const auto cF = expansion_coefficients(M, 2 * M, [](double x) { return fermi_density(x, 0.05, 0.0); }, es);
  const double E = moment_product(cE, mu);
  eng->autodiff_matrix(cE, D);
  SpMatBsr<cx_double> D2 = H;
  eng->stoch_matrix(cF, D2);
  // two-current moments with the x current j_x = i [H, X] (hopping bonds
  // along x only), probe columns in chunks of 5, then the static
  // conductivity sigma_xx = sum c[m1][m2] mu2[m1][m2]
  SpMatElems<cx_double> jel(L * L, L * L, 2);
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++) {
      const int i = x + L * y;
      const cx_double fw[4] = {cx_double(0, -1.0), 0, 0, cx_double(0, -1.0)};  // i t (+1), t = -1
      const cx_double bw[4] = {cx_double(0, 1.0), 0, 0, cx_double(0, 1.0)};    // i t (-1)
      jel.add(i, ((x + 1) % L) + L * y, fw);
      jel.add(i, ((x + L - 1) % L) + L * y, bw);
    }
  const SpMatBsr<cx_double> jx(jel);
  // stoch_matrix consumed the moments state; take it again before moments2
  const auto mu_again = eng->moments(M);
  CHECK(mu_again == mu);
  const int M2 = 16;
  const auto mu2 = eng->moments2_v1(M2, jx, jx, 10, 5);
  CHECK((int)mu2.size() == M2 && (int)mu2[0].size() == M2);
  const auto c2 = electrical_conductivity_coefficients(M2, 2 * M2, 0.05, 0.0, 0.0, es, jackson_kernel(M2));
  const cx_double sigma = moment_product(c2, mu2);
  // moments2 leaves the moments state alone: autodiff with the same M still
  // works and repeats the first result bit for bit
  SpMatBsr<cx_double> D3 = H;
  eng->autodiff_matrix(cE, D3);
  CHECK(D3.val == D.val);
  Vec<cx_double> mu2_flat;
  for (auto const& row : mu2) mu2_flat.insert(mu2_flat.end(), row.begin(), row.end());
  std::ofstream meta(dir + "/meta.txt");
  meta.precision(17);
  meta << n << " " << eng->R.n_cols << " " << M << " " << H.n_rows << " " << H.n_blocks() << " "
       << es.lo << " " << es.hi << " " << E << " " << M2 << " " << sigma.real() << " "
       << sigma.imag() << " " << jx.n_blocks() << "\n";
  dump(dir + "/H_row_ptr.bin", H.row_ptr.data(), H.row_ptr.size());
  dump(dir + "/H_row_idx.bin", H.row_idx.data(), H.row_idx.size());
  dump(dir + "/H_col_idx.bin", H.col_idx.data(), H.col_idx.size());
  dump(dir + "/H_val.bin", H.val.data(), H.val.size());
  dump(dir + "/R.bin", eng->R.memptr(), eng->R.size());
  dump(dir + "/mu.bin", mu.data(), mu.size());
  dump(dir + "/cE.bin", cE.data(), cE.size());
  dump(dir + "/cF.bin", cF.data(), cF.size());
  dump(dir + "/D_autodiff.bin", D.val.data(), D.val.size());
  dump(dir + "/D_stoch.bin", D2.val.data(), D2.val.size());
  dump(dir + "/mu2.bin", mu2_flat.data(), mu2_flat.size());
  dump(dir + "/j_row_ptr.bin", jx.row_ptr.data(), jx.row_ptr.size());
  dump(dir + "/j_row_idx.bin", jx.row_idx.data(), jx.row_idx.size());
  dump(dir + "/j_col_idx.bin", jx.col_idx.data(), jx.col_idx.size());
  dump(dir + "/j_val.bin", jx.val.data(), jx.val.size());
  std::printf("device OK E=%.12g\n", E);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "host") {
    host_checks();
    return 0;
  }
  if (argc >= 3 && std::string(argv[1]) == "device") return device_flow(argv[2]);
  std::fprintf(stderr, "usage: engine_flow host | engine_flow device DIR\n");
  return 2;
}
