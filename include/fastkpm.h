// fastkpm.h -- C++ API of the MI355X stochastic Chebyshev engine.
//
// Source-compatible replacement for the reference header src/fastkpm.h (same
// namespace, class and function names, argument meaning and call-order
// rules), made independent of Armadillo:
//   - Engine<T>::R / R2 are fkpm::DenseMat<T> (column-major, the subset of the
//     arma::Mat API the reference uses: n_rows, n_cols, size(), memptr(),
//     colptr(), operator()(i,j), set_size, fill, zeros);
//   - SpMatBsr::to_arma()/to_arma_dense() become to_dense();
//   - the arma::vec overloads of the spectral functions take Vec<double>.
// The engine returned by mk_engine<T>() runs on one HIP device through the
// C-ABI of fastkpm_hip.h.  There is no CPU engine behind it.
#ifndef FASTKPM_AMD_FASTKPM_H
#define FASTKPM_AMD_FASTKPM_H

#include <chrono>
#include <complex>
#include <functional>
#include <memory>
#include <ostream>
#include <random>
#include <string>
#include <vector>

namespace fkpm {
typedef std::mt19937 RNG;  // src/fastkpm.h:17

template <typename T>
using Vec = std::vector<T>;

template <typename S, typename T>
void copy_vec(Vec<S> const& src_vec, Vec<T>& dst_vec) {
  dst_vec.resize(src_vec.size());
  for (size_t k = 0; k < src_vec.size(); k++) dst_vec[k] = T(src_vec[k]);
}

constexpr double Pi = 3.141592653589793238463;

typedef std::complex<float> cx_float;
typedef std::complex<double> cx_double;

// complex conjugation that preserves real values
inline float conj(float x) { return x; }
inline double conj(double x) { return x; }
inline cx_float conj(cx_float x) { return std::conj(x); }
inline cx_double conj(cx_double x) { return std::conj(x); }

// Serial parallel_for (src/fastkpm.h:39-47 without TBB).
inline void parallel_for(size_t start, size_t end, std::function<void(size_t)> fn) {
  for (size_t i = start; i < end; i++) fn(i);
}

// -- dense column-major matrix (stands in for arma::Mat in Engine::R) ------
template <typename T>
class DenseMat {
 public:
  size_t n_rows = 0, n_cols = 0;
  DenseMat() = default;
  DenseMat(size_t r, size_t c) { set_size(r, c); }
  void set_size(size_t r, size_t c) {
    n_rows = r;
    n_cols = c;
    data_.assign(r * c, T(0));
  }
  size_t size() const { return data_.size(); }
  size_t n_elem() const { return data_.size(); }
  T* memptr() { return data_.data(); }
  T const* memptr() const { return data_.data(); }
  T* colptr(size_t j) { return data_.data() + j * n_rows; }
  T const* colptr(size_t j) const { return data_.data() + j * n_rows; }
  T& operator()(size_t i, size_t j) { return data_[j * n_rows + i]; }
  T const& operator()(size_t i, size_t j) const { return data_[j * n_rows + i]; }
  void fill(T v) { std::fill(data_.begin(), data_.end(), v); }
  void zeros() { fill(T(0)); }
  void reset() { set_size(0, 0); }

 private:
  std::vector<T> data_;
};

// -- spmat.cpp -------------------------------------------------------------

// Blocks of matrix elements in unsorted order (src/fastkpm.h:53-63)
template <typename T>
class SpMatElems {
 public:
  int n_rows, n_cols, b_len;
  Vec<int> row_idx, col_idx;
  Vec<T> val;
  SpMatElems(int n_rows, int n_cols, int b_len);
  int n_blocks() const;
  void clear();
  void add(int i, int j, T const* v);
};

// Block compressed sparse row matrix; each dense block column-major
// (src/fastkpm.h:64-106).
template <typename T>
class SpMatBsr {
 public:
  int n_rows = 0, n_cols = 0, b_len = 0;
  Vec<int> row_idx, col_idx, row_ptr;
  Vec<T> val;
  SpMatBsr();
  SpMatBsr(SpMatElems<T> const& elems);
  void build(SpMatElems<T> const& elems);
  int n_blocks() const;
  void clear();
  int find_index(int i, int j) const;
  T* operator()(int i, int j);
  T const* operator()(int i, int j) const;
  void zeros();
  void symmetrize();
  void scale(T alpha);
  // dense (b*n_rows) x (b*n_cols) copy, column-major (replaces to_arma_dense)
  DenseMat<T> to_dense() const;

  template <typename S>
  SpMatBsr(SpMatBsr<S> const& that) : SpMatBsr() {
    *this = that;
  }
  template <typename S>
  SpMatBsr<T>& operator=(SpMatBsr<S> const& that) {
    if ((void*)this != (void*)&that) {
      n_rows = that.n_rows;
      n_cols = that.n_cols;
      b_len = that.b_len;
      copy_vec(that.row_idx, row_idx);
      copy_vec(that.col_idx, col_idx);
      copy_vec(that.row_ptr, row_ptr);
      copy_vec(that.val, val);
    }
    return *this;
  }
};

// -- spectral functions (src/fastkpm.h:109-181) ------------------------------

// Scale eigenvalues within range (-1, +1)
class EnergyScale {
 public:
  double lo, hi;
  double avg() const { return (hi + lo) / 2.0; }
  double mag() const { return (hi - lo) / 2.0; }
  double scale(double x) const { return (x - avg()) / mag(); }
  double unscale(double x) const { return x * mag() + avg(); }
  friend std::ostream& operator<<(std::ostream& stream, EnergyScale const& es) {
    return stream << "< lo = " << es.lo << " hi = " << es.hi << " >\n";
  }
};

Vec<double> jackson_kernel(int M);
Vec<double> lorentz_kernel(int M, double lambda);
void chebyshev_fill_array(double x, Vec<double>& ret, int kind = 1);
Vec<double> expansion_coefficients(int M, int Mq, std::function<double(double)> f, EnergyScale es);
// C_{m1,m2} for the optical (omega > 0) or static (omega = 0) conductivity,
// unit e^2/h (src/fastkpm.cpp:80-233; the direct-sum form, no FFTW here)
Vec<Vec<cx_double>> electrical_conductivity_coefficients(int M, int Mq, double kT, double mu,
                                                         double omega, EnergyScale es,
                                                         Vec<double> const& kernel);
double moment_product(Vec<double> const& c, Vec<double> const& mu);
cx_double moment_product(Vec<Vec<cx_double>> const& c, Vec<Vec<cx_double>> const& mu);
Vec<double> moment_transform(Vec<double> const& moments, int Mq);
Vec<Vec<cx_double>> moment_transform(Vec<Vec<cx_double>> const& moments, int Mq,
                                     Vec<double> const& kernel);
double density_product(Vec<double> const& gamma, std::function<double(double)> f, EnergyScale es);
void density_function(Vec<double> const& gamma, EnergyScale es, Vec<double>& x, Vec<double>& rho);
void density_function(Vec<Vec<cx_double>> const& gamma, EnergyScale es, Vec<double>& x,
                      Vec<double>& y, Vec<Vec<cx_double>>& rho);
void integrated_density_function(Vec<double> const& gamma, EnergyScale es, Vec<double>& x,
                                 Vec<double>& irho);
double fermi_energy(double x, double kT, double mu);
double fermi_density(double x, double kT, double mu);
// Overloads on an eigenvalue list take Vec<double> where the reference takes
// arma::vec.
double mu_to_filling(Vec<double> const& gamma, EnergyScale const& es, double kT, double mu);
double mu_to_filling(Vec<double> const& evals, double kT, double mu);
double filling_to_mu(Vec<double> const& gamma, EnergyScale const& es, double kT, double filling,
                     double delta_filling);
double filling_to_mu(Vec<double> const& evals, double kT, double filling);
double electronic_grand_energy(Vec<double> const& gamma, EnergyScale const& es, double kT,
                               double mu);
double electronic_grand_energy(Vec<double> const& evals, double kT, double mu);
double electronic_energy(Vec<double> const& gamma, EnergyScale const& es, double kT,
                         double filling, double mu);
double electronic_energy(Vec<double> const& evals, double kT, double filling);

// -- engines (src/fastkpm.h:185-250) ------------------------------------------

template <typename T>
class Engine {
 public:
  DenseMat<T> R;   // Random vectors
  DenseMat<T> R2;  // Another random vector for the 2nd version of moments

  virtual ~Engine() {}

  // Lanczos approximation to eigenvalue span
  virtual EnergyScale energy_scale(SpMatBsr<T> const& H, double extend, int iters = 128) = 0;

  // Identity matrix
  virtual void set_R_identity(int n, int j_start, int j_end);
  virtual void set_R_identity(int n);

  // Uncorrelated random elements
  virtual void set_R_uncorrelated(int n, int s, RNG& rng, int j_start, int j_end);
  virtual void set_R_uncorrelated(int n, int s, RNG& rng);

  // Correlated random elements with mostly orthogonal rows
  virtual void set_R_correlated(Vec<int> const& groups, RNG& rng, int j_start, int j_end);
  virtual void set_R_correlated(Vec<int> const& groups, RNG& rng);

  // Transfer R matrix to device
  virtual void transfer_R() {}

  // Set Hamiltonian and energy scale
  virtual void set_H(SpMatBsr<T> const& H, EnergyScale const& es) = 0;

  // Chebyshev moments: mu_m = tr T_m(Hs) ~ tr R^\dagger T_m(Hs) R
  virtual Vec<double> moments(int M) = 0;

  // Chebyshev moments: mu_{m1,m2} = tr( j1 T_{m1}(Hs) j2 T_{m2}(Hs) ).  The HIP
  // engine computes moments2_v1 on the device (R_chunk_ncols bounds the probe
  // columns per pass, a_chunk_ncols is unused: all M orders stay in HBM);
  // moments2_v2 exits with a message, as the reference GPU engine does.
  virtual Vec<Vec<cx_double>> moments2_v1(int M, SpMatBsr<T> const& j1op,
                                          SpMatBsr<T> const& j2op, int a_chunk_ncols = 0,
                                          int R_chunk_ncols = 0) = 0;
  virtual Vec<Vec<cx_double>> moments2_v2(int M, SpMatBsr<T> const& j1op,
                                          SpMatBsr<T> const& j2op, int a_chunk_ncols = 0,
                                          int R_chunk_ncols = 0) = 0;

  // Approximates D ~ (xi R^\dagger + R xi^\dagger)/2 where xi = D R
  // and D ~ (\sum_m c_m T_m(Hs))R
  virtual void stoch_matrix(Vec<double> const& c, SpMatBsr<T>& D) = 0;

  // Approximates D ~ (d/dH^T) tr g where tr g ~ tr R^\dagger g R.
  // REQUIREMENT: moments() must have been called previously (same M).
  virtual void autodiff_matrix(Vec<double> const& c, SpMatBsr<T>& D) = 0;
};

// Storage precision of the HIP engine's Chebyshev vectors.
enum class Precision { FP64 = 0, FP32 = 1 };

// HIP engine on one device (the factory this library adds).
template <typename T>
std::shared_ptr<Engine<T>> mk_engine_hip(int device = 0, Precision prec = Precision::FP64);

// Not provided: this library has no CPU engine.  Throws std::logic_error.
template <typename T>
std::shared_ptr<Engine<T>> mk_engine_cpu();

// Compatibility name of the GPU factory: the HIP engine on devices[0] (device
// 0 if empty).  More than one device throws (S-column sharding over devices
// is not built).  Returns nullptr when no HIP device exists, as the reference
// does without CUDA devices.
template <typename T>
std::shared_ptr<Engine<T>> mk_engine_cuSPARSE(Vec<int> devices = {});

// Fastest single-node engine available: the HIP engine; throws
// std::runtime_error when no HIP device is visible (no CPU fallback).
template <typename T>
std::shared_ptr<Engine<T>> mk_engine();

// Without MPI this is mk_engine<T>() (the reference's non-MPI fallback,
// src/engine_mpi.cpp:3-15).
template <typename T>
std::shared_ptr<Engine<T>> mk_engine_mpi();

// -- timer.cpp ---------------------------------------------------------------
class Timer {
 public:
  std::chrono::time_point<std::chrono::system_clock> t0;
  Timer();
  void reset();
  double measure();  // elapsed time in seconds
};
extern Timer timer[10];
}  // namespace fkpm

#endif  // FASTKPM_AMD_FASTKPM_H
