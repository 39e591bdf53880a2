/*
 * fastkpm_host.h -- C entry points into the host half of the C++ library
 * (libfastkpm.so): probing-vector generation and spectral functions, for
 * callers without C++ (the Python mirror, tests, benchmarks).  These have no
 * counterpart in the reference, whose only interface is C++ (src/fastkpm.h);
 * each one forwards to the C++ function named in its comment.
 */
#ifndef FASTKPM_HOST_H
#define FASTKPM_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed fkpm_host_* call on this thread (FKPM_E_ARG
   return): the text of the exception the C++ host library threw. */
const char* fkpm_host_last_error(void);

/* std::mt19937 (fkpm::RNG) owned by the library. */
void* fkpm_host_rng_new(uint32_t seed);
void fkpm_host_rng_free(void* rng);
uint32_t fkpm_host_rng_next(void* rng);
void fkpm_host_rng_discard(void* rng, unsigned long long k);

/* Engine<T>::set_R_identity / set_R_uncorrelated / set_R_correlated
 * (src/engine.cpp:34-116) for T given by dtype (FKPM_F32..FKPM_C64 of
 * fastkpm_hip.h).  R_out receives n x (j1-j0) values, column-major; the
 * correlated form with j0 = j1 = -1 is the two-argument overload (identity
 * when the colour count reaches n) and writes the column count to *s_out. */
int fkpm_host_set_R_identity(int dtype, int n, int j0, int j1, void* R_out);
int fkpm_host_set_R_uncorrelated(int dtype, int n, int s, void* rng, int j0, int j1, void* R_out);
int fkpm_host_set_R_correlated(int dtype, int n, const int* groups, void* rng, int j0, int j1,
                               void* R_out, int* s_out);

/* The complex-double R of set_R_correlated(groups, rng) (two-argument form)
 * in compact form, one nonzero per orbital row: rcol_out[i] = its column
 * (-1 for none), rval_out[i] (complex double) = its value; the same draws as
 * the dense form.  *s_out = column count. */
int fkpm_host_set_R_probing(int n, const int* groups, void* rng, int* rcol_out, void* rval_out,
                            int* s_out);

/* expansion_coefficients(M, Mq, f, es) with f = fermi_energy(., kT, mu)
 * (kind 0), fermi_density(., kT, mu) (kind 1) or 1 (kind 2). */
int fkpm_host_expansion_coefficients(int M, int Mq, int kind, double kT, double mu, double lo,
                                     double hi, double* c_out);
void fkpm_host_jackson_kernel(int M, double* g_out);
/* electrical_conductivity_coefficients(M, Mq, kT, mu, omega, es, kernel):
 * M x M complex written row-major as (re, im) pairs. */
int fkpm_host_conductivity_coefficients(int M, int Mq, double kT, double mu, double omega,
                                        double lo, double hi, const double* kernel, double* out);
/* moment_transform of M x M complex moments (row-major (re, im)) with a
 * kernel: Mq x Mq complex out. */
int fkpm_host_moment_transform2(int M, const double* mu, int Mq, const double* kernel,
                                double* gamma_out);
double fkpm_host_moment_product(int M, const double* c, const double* mu);
void fkpm_host_moment_transform(int M, const double* mu, int Mq, double* gamma_out);
double fkpm_host_electronic_grand_energy(int Mq, const double* gamma, double lo, double hi,
                                         double kT, double mu);
double fkpm_host_filling_to_mu(int Mq, const double* gamma, double lo, double hi, double kT,
                               double filling, double delta_filling);

#ifdef __cplusplus
}
#endif

#endif /* FASTKPM_HOST_H */
