// Engine<T>: probing-vector generation (the base class) and the HIP engine
// over the C-ABI of fastkpm_hip.h, with the factories of the reference API.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fastkpm.h"
#include "fastkpm_hip.h"

namespace fkpm {

// -- probing vectors (src/engine.cpp:7-116) -----------------------------------

template <typename T>
T random_phase(RNG& rng);

// {+1, -1}
template <>
double random_phase(RNG& rng) {
  std::uniform_int_distribution<int> dist2(0, 1);
  return 2 * dist2(rng) - 1;
}
template <>
float random_phase(RNG& rng) {
  return float(random_phase<double>(rng));
}

// {+1, +i, -1, -i}
template <>
cx_double random_phase(RNG& rng) {
  std::uniform_int_distribution<int> dist4(0, 3);
  switch (dist4(rng)) {
    case 0: return {+1, 0};
    case 1: return {0, +1};
    case 2: return {-1, 0};
    default: return {0, -1};
  }
}
template <>
cx_float random_phase(RNG& rng) {
  return cx_float(random_phase<cx_double>(rng));
}

template <typename T>
void Engine<T>::set_R_identity(int n, int j_start, int j_end) {
  if (!(0 <= j_start && j_start <= j_end && j_end <= n))
    throw std::invalid_argument("set_R_identity: bad column range");
  R.set_size(n, j_end - j_start);
  R.fill(T(0));
  for (int j = j_start; j < j_end; j++) R(j, j - j_start) = T(1);
  R2 = R;
  transfer_R();
}

template <typename T>
void Engine<T>::set_R_identity(int n) {
  set_R_identity(n, 0, n);
}

template <typename T>
void Engine<T>::set_R_uncorrelated(int n, int s, RNG& rng, int j_start, int j_end) {
  if (!(0 <= j_start && j_start <= j_end && j_end <= s))
    throw std::invalid_argument("set_R_uncorrelated: bad column range");
  R.set_size(n, j_end - j_start);
  const T x = T(1.0 / std::sqrt((double)s));
  for (int j = j_start; j < j_end; j++) {
    RNG rng_j(rng());  // one draw of the caller's engine per column
    for (int i = 0; i < n; i++) R(i, j - j_start) = random_phase<T>(rng_j) * x;
  }
  R2 = R;
  transfer_R();
}

template <typename T>
void Engine<T>::set_R_uncorrelated(int n, int s, RNG& rng) {
  set_R_uncorrelated(n, s, rng, 0, s);
}

template <typename T>
void Engine<T>::set_R_correlated(Vec<int> const& groups, RNG& rng, int j_start, int j_end) {
  const int n = (int)groups.size();
  if (!(0 <= j_start && j_start <= j_end && j_end <= n))
    throw std::invalid_argument("set_R_correlated: bad column range");
  R.set_size(n, j_end - j_start);
  R.fill(T(0));
  // rows grouped by colour once, then column j draws its phases in row order
  Vec<int> cnt(j_end - j_start + 1, 0);
  for (int i = 0; i < n; i++)
    if (groups[i] >= j_start && groups[i] < j_end) cnt[groups[i] - j_start + 1]++;
  for (size_t k = 1; k < cnt.size(); k++) cnt[k] += cnt[k - 1];
  Vec<int> rows(cnt.back());
  {
    Vec<int> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; i++)
      if (groups[i] >= j_start && groups[i] < j_end) rows[pos[groups[i] - j_start]++] = i;
  }
  for (int j = j_start; j < j_end; j++) {
    RNG rng_j(rng());
    for (int k = cnt[j - j_start]; k < cnt[j - j_start + 1]; k++)
      R(rows[k], j - j_start) = random_phase<T>(rng_j);
  }
  R2 = R;
  transfer_R();
}

template <typename T>
void Engine<T>::set_R_correlated(Vec<int> const& groups, RNG& rng) {
  if (groups.empty()) throw std::invalid_argument("set_R_correlated: empty groups");
  const auto mm = std::minmax_element(groups.begin(), groups.end());
  if (*mm.first != 0) throw std::invalid_argument("set_R_correlated: colours must start at 0");
  const int n = (int)groups.size();
  const int s = *mm.second + 1;
  if (s < n)
    set_R_correlated(groups, rng, 0, s);
  else
    set_R_identity(n, 0, n);
}

template class Engine<float>;
template class Engine<double>;
template class Engine<cx_float>;
template class Engine<cx_double>;

// -- HIP engine ----------------------------------------------------------------

template <typename T>
constexpr int dtype_code() {
  if (std::is_same<T, float>::value) return FKPM_F32;
  if (std::is_same<T, double>::value) return FKPM_F64;
  if (std::is_same<T, cx_float>::value) return FKPM_C32;
  return FKPM_C64;
}

static void check_rc(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + fkpm_last_error());
}

template <typename T>
class Engine_HIP : public Engine<T> {
 public:
  Engine_HIP(int device, Precision prec) {
    check_rc(fkpm_dev_open(device, dtype_code<T>(), prec == Precision::FP32 ? FKPM_PREC_FP32 : FKPM_PREC_FP64, &h_),
             "fkpm_dev_open");
  }
  ~Engine_HIP() override { fkpm_dev_close(h_); }

  // Lanczos on the device (src/engine_cusparse.cpp:598-673).  The start
  // vector is drawn from a fixed-seed std::mt19937 with normal variates (the
  // reference draws it with arma::randn).
  EnergyScale energy_scale(SpMatBsr<T> const& H, double extend, int iters) override {
    const size_t n = (size_t)H.n_rows * H.b_len;
    std::mt19937 g(0);
    std::normal_distribution<double> nd(0.0, 1.0);
    Vec<T> v(n);
    for (auto& x : v) x = make_normal(nd, g);
    EnergyScale es{0, 0};
    check_rc(fkpm_energy_scale(h_, H.n_rows, H.b_len, H.n_blocks(), H.row_ptr.data(),
                               H.col_idx.data(), H.val.data(), extend, iters, v.data(), &es.lo,
                               &es.hi),
             "energy_scale");
    return es;
  }

  void transfer_R() override {
    if (this->R.n_cols == 0) throw std::invalid_argument("cannot handle an empty R matrix");
    check_rc(fkpm_set_R(h_, (int)this->R.n_rows, (int)this->R.n_cols, this->R.memptr()),
             "transfer_R");
  }

  void set_H(SpMatBsr<T> const& H, EnergyScale const& es) override {
    if (H.n_rows != H.n_cols) throw std::invalid_argument("set_H: H must be square");
    es_ = es;
    check_rc(fkpm_set_H(h_, H.n_rows, H.b_len, H.n_blocks(), H.row_ptr.data(), H.col_idx.data(),
                        H.val.data(), es.avg(), es.mag()),
             "set_H");
  }

  Vec<double> moments(int M) override {
    Vec<double> mu(M);
    check_rc(fkpm_moments(h_, M, mu.data()), "moments");
    return mu;
  }

  // Two-current moments on the device (src/engine_cusparse.cpp:793-902).
  Vec<Vec<cx_double>> moments2_v1(int M, SpMatBsr<T> const& j1op, SpMatBsr<T> const& j2op,
                                  int a_chunk_ncols, int R_chunk_ncols) override {
    (void)a_chunk_ncols;  // every Chebyshev order stays in HBM
    for (const SpMatBsr<T>* j : {&j1op, &j2op})
      if (j->n_rows != j->n_cols || j->b_len <= 0)
        throw std::invalid_argument("moments2_v1: current operators must be square BSR matrices");
    Vec<double> flat((size_t)M * M * 2);
    check_rc(fkpm_moments2(h_, M, j1op.n_blocks(), j1op.row_ptr.data(), j1op.col_idx.data(),
                           j1op.val.data(), j2op.n_blocks(), j2op.row_ptr.data(),
                           j2op.col_idx.data(), j2op.val.data(), R_chunk_ncols, flat.data()),
             "moments2_v1");
    Vec<Vec<cx_double>> mu(M, Vec<cx_double>(M));
    for (int a = 0; a < M; a++)
      for (int b = 0; b < M; b++)
        mu[a][b] = cx_double(flat[((size_t)a * M + b) * 2], flat[((size_t)a * M + b) * 2 + 1]);
    return mu;
  }
  Vec<Vec<cx_double>> moments2_v2(int, SpMatBsr<T> const&, SpMatBsr<T> const&, int, int) override {
    std::cerr << "moments2_v2 is not provided by the HIP engine\n";
    std::exit(EXIT_FAILURE);
  }

  void stoch_matrix(Vec<double> const& c, SpMatBsr<T>& D) override {
    set_pattern(D);
    check_rc(fkpm_stoch_matrix(h_, (int)c.size(), c.data(), D.val.data()), "stoch_matrix");
  }

  void autodiff_matrix(Vec<double> const& c, SpMatBsr<T>& D) override {
    set_pattern(D);
    check_rc(fkpm_autodiff_matrix(h_, (int)c.size(), c.data(), D.val.data()), "autodiff_matrix");
  }

 private:
  fkpm_handle* h_ = nullptr;
  EnergyScale es_{0, 0};
  int d_rows_ = -1;
  Vec<int> d_row_ptr_, d_col_idx_;

  template <class G>
  static T make_normal(std::normal_distribution<double>& nd, G& g) {
    if constexpr (std::is_same<T, cx_float>::value || std::is_same<T, cx_double>::value) {
      const double re = nd(g);
      const double im = nd(g);
      return T(re, im);
    } else {
      return T(nd(g));
    }
  }

  // Upload D's pattern when it differs from the last one.
  void set_pattern(SpMatBsr<T>& D) {
    D.val.resize((size_t)D.n_blocks() * D.b_len * D.b_len);
    if (D.n_rows == d_rows_ && D.row_ptr == d_row_ptr_ && D.col_idx == d_col_idx_) return;
    check_rc(fkpm_set_D_pattern(h_, D.n_rows, D.n_blocks(), D.row_ptr.data(), D.row_idx.data(),
                                D.col_idx.data()),
             "D pattern");
    d_rows_ = D.n_rows;
    d_row_ptr_ = D.row_ptr;
    d_col_idx_ = D.col_idx;
  }
};

template <typename T>
std::shared_ptr<Engine<T>> mk_engine_hip(int device, Precision prec) {
  return std::make_shared<Engine_HIP<T>>(device, prec);
}

template <typename T>
std::shared_ptr<Engine<T>> mk_engine_cpu() {
  throw std::logic_error(
      "fastkpm-amd provides no CPU engine: use mk_engine<T>() (the HIP engine)");
}

static bool printed_devices = false;

template <typename T>
std::shared_ptr<Engine<T>> mk_engine_cuSPARSE(Vec<int> devices) {
  const int count = fkpm_device_count();
  if (count == 0) {
    if (!printed_devices) std::cout << "No HIP devices are available!\n";
    printed_devices = true;
    return nullptr;
  }
  if (devices.size() > 1)
    throw std::invalid_argument("one HIP device per engine: S-column sharding over several devices is not built");
  const int d = devices.empty() ? 0 : devices[0];
  if (d < 0 || d >= count) {
    std::cerr << "Invalid device #" << d << " (" << count << " available)!\n";
    std::exit(EXIT_FAILURE);
  }
  printed_devices = true;
  return mk_engine_hip<T>(d, Precision::FP64);
}

template <typename T>
std::shared_ptr<Engine<T>> mk_engine() {
  auto e = mk_engine_cuSPARSE<T>();
  if (!e) throw std::runtime_error("mk_engine: no HIP device is available (there is no CPU fallback)");
  return e;
}

template <typename T>
std::shared_ptr<Engine<T>> mk_engine_mpi() {
  return mk_engine<T>();
}

#define FKPM_INSTANTIATE(T)                                                  \
  template std::shared_ptr<Engine<T>> mk_engine_hip<T>(int, Precision);      \
  template std::shared_ptr<Engine<T>> mk_engine_cpu<T>();                    \
  template std::shared_ptr<Engine<T>> mk_engine_cuSPARSE<T>(Vec<int>);       \
  template std::shared_ptr<Engine<T>> mk_engine<T>();                        \
  template std::shared_ptr<Engine<T>> mk_engine_mpi<T>();
FKPM_INSTANTIATE(float)
FKPM_INSTANTIATE(double)
FKPM_INSTANTIATE(cx_float)
FKPM_INSTANTIATE(cx_double)
#undef FKPM_INSTANTIATE

// -- timer (src/timer.cpp) ----------------------------------------------------
Timer timer[10];
Timer::Timer() { reset(); }
void Timer::reset() { t0 = std::chrono::system_clock::now(); }
double Timer::measure() {
  const std::chrono::duration<double> d = std::chrono::system_clock::now() - t0;
  return d.count();
}

}  // namespace fkpm
