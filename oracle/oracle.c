/*
 * oracle.c -- CPU restatement (FP64) of the FastKPM reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it (as liboracle_fkpm.so through
 * ctypes) to compare against the HIP engine.  The product library never links,
 * loads or calls it, and nothing here is a fallback for the GPU path.
 *
 * Every function restates one function of the reference (file:line cited below
 * relative to the reference root).  All arithmetic is double / double complex;
 * real engine types are handled by feeding complex values with zero imaginary
 * part, which gives identical values (every product with an exact zero is an
 * exact zero).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   mt19937            C++ standard known answer (10000th output of the default
 *                      seed is 4123659995) and a libstdc++ fixture.
 *   uniform_int        libstdc++ fixture (tests/golden/make_rng_fixture.cpp).
 *   R generation       structural invariants (unit modulus, colour groups,
 *                      draws consumed) and the libstdc++ fixture.
 *   bsr_build          hand-built unsorted/duplicate inputs vs numpy dense sums.
 *   symmetrize/scale   Hermiticity and dense-matrix identities.
 *   jackson/cheb/coef  closed forms (g_0 = 1, T_m(cos t) = cos(m t), f = 1,
 *                      polynomial f of degree < M).
 *   fermi_*            closed forms / limits.
 *   moments            R = identity gives tr T_m(Hs) = sum_k T_m(lambda_k)
 *                      (numpy eigvalsh, periodic-lattice closed form).
 *   stoch_matrix       R = identity gives p(Hs) on the pattern (numpy eigh).
 *   autodiff_matrix    central finite differences of E(H) = sum c_m mu_m(H) at
 *                      fixed R and energy scale, and R = identity vs p'(Hs).
 *   lanczos/tri-eig    numpy eigvalsh on the tridiagonal and on dense H.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;
static const double ORC_PI = 3.141592653589793238463;

/* ------------------------------------------------------------------------- */
/* std::mt19937 (fastkpm.h:17 `typedef std::mt19937 RNG`).                    */
/* Restated from the published MT19937 definition (Matsumoto & Nishimura 1998) */
/* as fixed by the C++ standard [rand.eng.mers].                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t x[624];
    int idx;
} orc_mt;

void orc_mt_seed(orc_mt* g, uint32_t seed) {
    g->x[0] = seed;
    for (int i = 1; i < 624; i++) {
        uint32_t p = g->x[i - 1];
        g->x[i] = 1812433253u * (p ^ (p >> 30)) + (uint32_t)i;
    }
    g->idx = 624;
}

static void orc_mt_twist(orc_mt* g) {
    for (int i = 0; i < 624; i++) {
        uint32_t y = (g->x[i] & 0x80000000u) | (g->x[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = g->x[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        g->x[i] = v;
    }
    g->idx = 0;
}

uint32_t orc_mt_next(orc_mt* g) {
    if (g->idx >= 624) orc_mt_twist(g);
    uint32_t y = g->x[g->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

int orc_mt_size(void) { return (int)sizeof(orc_mt); }

/* std::uniform_int_distribution<int>(a, b) applied to a 32-bit engine, as
 * libstdc++ (GCC >= 11, bits/uniform_int_dist.h) implements it: Lemire's
 * nearly-divisionless downscaling ("Fast Random Integer Generation in an
 * Interval", ACM TOMACS 29(1), 2019) with a 64-bit product.  Used by
 * random_phase (engine.cpp:12-30). */
int orc_uniform_int(orc_mt* g, int a, int b) {
    uint32_t range = (uint32_t)b - (uint32_t)a + 1u; /* b > a assumed, range < 2^32 */
    uint64_t prod = (uint64_t)orc_mt_next(g) * (uint64_t)range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t threshold = (uint32_t)(0u - range) % range;
        while (low < threshold) {
            prod = (uint64_t)orc_mt_next(g) * (uint64_t)range;
            low = (uint32_t)prod;
        }
    }
    return a + (int)(prod >> 32);
}

/* random_phase<T> (engine.cpp:7-32).  is_complex selects {+1,+i,-1,-i}
 * (uniform_int(0,3)); otherwise {+1,-1} (2*uniform_int(0,1)-1). */
static cplx orc_random_phase(orc_mt* g, int is_complex) {
    if (is_complex) {
        switch (orc_uniform_int(g, 0, 3)) {
            case 0: return 1.0;
            case 1: return I;
            case 2: return -1.0;
            default: return -I;
        }
    }
    return (cplx)(2 * orc_uniform_int(g, 0, 1) - 1);
}

/* Engine::set_R_identity (engine.cpp:34-44).  R is n x (j1-j0), column-major. */
void orc_set_R_identity(int n, int j0, int j1, cplx* R) {
    int s = j1 - j0;
    memset(R, 0, sizeof(cplx) * (size_t)n * (size_t)s);
    for (int j = j0; j < j1; j++) R[(size_t)(j - j0) * n + j] = 1.0;
}

/* Engine::set_R_uncorrelated (engine.cpp:50-70).  single != 0 rounds the scale
 * 1/sqrt(s) to float as `T x = 1.0/sqrt(s)` does for float engines. */
void orc_set_R_uncorrelated(int n, int s, orc_mt* rng, int j0, int j1, int is_complex,
                            int single, cplx* R) {
    double x = 1.0 / sqrt((double)s);
    if (single) x = (double)(float)x;
    for (int j = j0; j < j1; j++) {
        orc_mt gj;
        orc_mt_seed(&gj, orc_mt_next(rng)); /* RNG rng_j(rng()) */
        for (int i = 0; i < n; i++) R[(size_t)(j - j0) * n + i] = orc_random_phase(&gj, is_complex) * x;
    }
}

/* Engine::set_R_correlated(groups, rng, j0, j1) (engine.cpp:76-102). */
void orc_set_R_correlated(int n, const int* groups, orc_mt* rng, int j0, int j1, int is_complex,
                          cplx* R) {
    int s = j1 - j0;
    memset(R, 0, sizeof(cplx) * (size_t)n * (size_t)s);
    for (int j = j0; j < j1; j++) {
        orc_mt gj;
        orc_mt_seed(&gj, orc_mt_next(rng));
        for (int i = 0; i < n; i++)
            if (groups[i] == j) R[(size_t)(j - j0) * n + i] = orc_random_phase(&gj, is_complex);
    }
}

/* ------------------------------------------------------------------------- */
/* SpMatBsr (spmat.cpp).                                                      */
/* ------------------------------------------------------------------------- */

/* SpMatBsr::build (spmat.cpp:46-129): COO blocks -> BSR sorted by (row, col)
 * with duplicate blocks summed.  Outputs must hold nblk entries (blocks) and
 * n_rows+1 row pointers; returns the number of unique blocks. */
int orc_bsr_build(int n_rows, int b, int nblk, const int* ri, const int* ci, const cplx* v,
                  int* row_ptr, int* row_idx, int* col_idx, cplx* val) {
    int bb = b * b;
    int* order = (int*)malloc(sizeof(int) * (size_t)(nblk > 0 ? nblk : 1));
    int sorted = 1;
    for (int k = 0; k + 1 < nblk; k++) {
        if (ri[k] > ri[k + 1] || (ri[k] == ri[k + 1] && ci[k] > ci[k + 1])) { sorted = 0; break; }
    }
    if (sorted) {
        for (int p = 0; p < nblk; p++) order[p] = p;
    } else {
        /* bin by row in input order, then a stable insertion sort by column
         * within each row (the reference uses std::sort per row bin; for equal
         * (i,j) keys the summation order of duplicates may differ, which only
         * changes rounding of duplicate sums). */
        int* cnt = (int*)calloc((size_t)n_rows + 1, sizeof(int));
        for (int p = 0; p < nblk; p++) cnt[ri[p] + 1]++;
        for (int i = 0; i < n_rows; i++) cnt[i + 1] += cnt[i];
        int* fill = (int*)malloc(sizeof(int) * (size_t)(n_rows + 1));
        memcpy(fill, cnt, sizeof(int) * (size_t)(n_rows + 1));
        for (int p = 0; p < nblk; p++) order[fill[ri[p]]++] = p;
        for (int i = 0; i < n_rows; i++) {
            for (int a = cnt[i] + 1; a < cnt[i + 1]; a++) {
                int key = order[a];
                int q = a - 1;
                while (q >= cnt[i] && ci[order[q]] > ci[key]) { order[q + 1] = order[q]; q--; }
                order[q + 1] = key;
            }
        }
        free(fill);
        free(cnt);
    }
    int max_row = -1, cntu = 0;
    for (int t = 0; t < nblk; t++) {
        int p = order[t];
        int i = ri[p], j = ci[p];
        if (cntu > 0 && row_idx[cntu - 1] == i && col_idx[cntu - 1] == j) {
            for (int e = 0; e < bb; e++) val[(size_t)(cntu - 1) * bb + e] += v[(size_t)p * bb + e];
        } else {
            row_idx[cntu] = i;
            col_idx[cntu] = j;
            memcpy(&val[(size_t)cntu * bb], &v[(size_t)p * bb], sizeof(cplx) * (size_t)bb);
            while (max_row < i) row_ptr[++max_row] = cntu;
            cntu++;
        }
    }
    while (max_row < n_rows) row_ptr[++max_row] = cntu;
    free(order);
    return cntu;
}

/* SpMatBsr::find_index (spmat.cpp:145-153); returns -1 where the reference aborts. */
static int orc_find_index(const int* row_ptr, const int* col_idx, int i, int j) {
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++)
        if (col_idx[p] == j) return p;
    return -1;
}

/* SpMatBsr::symmetrize (spmat.cpp:171-189).  Returns 0, or -1 if the pattern
 * is not structurally symmetric (the reference aborts there). */
int orc_bsr_symmetrize(int b, int nnzb, const int* row_ptr, const int* row_idx,
                       const int* col_idx, cplx* val) {
    int bb = b * b;
    for (int k = 0; k < nnzb; k++) {
        int i = row_idx[k], j = col_idx[k];
        if (i >= j) {
            int k2 = orc_find_index(row_ptr, col_idx, j, i);
            if (k2 < 0) return -1;
            cplx* v1b = &val[(size_t)k * bb];
            cplx* v2b = &val[(size_t)k2 * bb];
            for (int bj = 0; bj < b; bj++)
                for (int bi = 0; bi < b; bi++) {
                    cplx* v1 = &v1b[bj * b + bi];
                    cplx* v2 = &v2b[bi * b + bj];
                    cplx sym = (*v1 + conj(*v2)) / 2.0;
                    *v1 = sym;
                    *v2 = conj(sym);
                }
        }
    }
    return 0;
}

/* SpMatBsr::scale (spmat.cpp:192-197). */
void orc_bsr_scale(int count, cplx* val, double alpha) {
    for (int k = 0; k < count; k++) val[k] *= alpha;
}

/* ------------------------------------------------------------------------- */
/* Spectral math (fastkpm.cpp).                                               */
/* ------------------------------------------------------------------------- */

/* jackson_kernel (fastkpm.cpp:13-20). */
void orc_jackson_kernel(int M, double* g) {
    double Mp = M + 1.0;
    for (int m = 0; m < M; m++)
        g[m] = (1.0 / Mp) * ((Mp - m) * cos(ORC_PI * m / Mp) + sin(ORC_PI * m / Mp) / tan(ORC_PI / Mp));
}

/* chebyshev_fill_array (fastkpm.cpp:30-39). */
void orc_chebyshev_fill_array(double x, int n, double* ret, int kind) {
    if (n > 0) ret[0] = 1.0;
    if (n > 1) ret[1] = kind * x;
    for (int m = 2; m < n; m++) ret[m] = 2 * x * ret[m - 1] - ret[m - 2];
}

/* fermi_energy / fermi_density (fastkpm.cpp:375-393). */
double orc_fermi_energy(double x, double kT, double mu) {
    double alpha = (x - mu) / fabs(kT);
    if (kT < 1e-15 || fabs(alpha) > 20) return (x < mu) ? (x - mu) : 0.0;
    return -kT * log(1 + exp(-alpha));
}

double orc_fermi_density(double x, double kT, double mu) {
    double alpha = (x - mu) / fabs(kT);
    if (kT < 1e-15 || fabs(alpha) > 20) return (x < mu) ? 1.0 : 0.0;
    return 1.0 / (exp(alpha) + 1.0);
}

/* Chebyshev-Gauss nodes x_i = cos(pi (i+1/2)/Mq) mapped by EnergyScale::unscale
 * (fastkpm.h:118): the points at which expansion_coefficients samples f. */
void orc_chebyshev_points(int Mq, double lo, double hi, double* x_unscaled) {
    double avg = (hi + lo) / 2.0, mag = (hi - lo) / 2.0;
    for (int i = 0; i < Mq; i++) x_unscaled[i] = cos(ORC_PI * (i + 0.5) / Mq) * mag + avg;
}

/* expansion_coefficients, direct-sum branch (fastkpm.cpp:41-44, 62-78), with
 * f given by its samples f_i = f(es.unscale(x_i)). */
void orc_expansion_coefficients_samples(int M, int Mq, const double* f_samples, double* c) {
    double* kern = (double*)malloc(sizeof(double) * (size_t)M);
    double* fp = (double*)calloc((size_t)M, sizeof(double));
    double* T = (double*)malloc(sizeof(double) * (size_t)M);
    orc_jackson_kernel(M, kern);
    for (int i = 0; i < Mq; i++) {
        double x_i = cos(ORC_PI * (i + 0.5) / Mq);
        orc_chebyshev_fill_array(x_i, M, T, 1);
        for (int m = 0; m < M; m++) fp[m] += f_samples[i] * T[m];
    }
    for (int m = 0; m < M; m++) c[m] = (m == 0 ? 1.0 : 2.0) * kern[m] * fp[m] / Mq;
    free(kern);
    free(fp);
    free(T);
}

/* expansion_coefficients(M, Mq, f, es) with f = fermi_energy(., kT, mu)
 * (kind 0) or fermi_density(., kT, mu) (kind 1) or the constant 1 (kind 2). */
void orc_expansion_coefficients_fermi(int M, int Mq, int kind, double kT, double mu, double lo,
                                      double hi, double* c) {
    double* xs = (double*)malloc(sizeof(double) * (size_t)Mq);
    double* fs = (double*)malloc(sizeof(double) * (size_t)Mq);
    orc_chebyshev_points(Mq, lo, hi, xs);
    for (int i = 0; i < Mq; i++) {
        fs[i] = kind == 0 ? orc_fermi_energy(xs[i], kT, mu)
              : kind == 1 ? orc_fermi_density(xs[i], kT, mu) : 1.0;
    }
    orc_expansion_coefficients_samples(M, Mq, fs, c);
    free(xs);
    free(fs);
}

/* moment_product (fastkpm.cpp:291-298). */
double orc_moment_product(int M, const double* c, const double* mu) {
    double r = 0;
    for (int i = 0; i < M; i++) r += c[i] * mu[i];
    return r;
}

/* moment_transform (fastkpm.cpp:236-256). */
void orc_moment_transform(int M, const double* mu, int Mq, double* gamma) {
    double* kern = (double*)malloc(sizeof(double) * (size_t)M);
    double* T = (double*)malloc(sizeof(double) * (size_t)M);
    orc_jackson_kernel(M, kern);
    for (int i = 0; i < Mq; i++) {
        double x_i = cos(ORC_PI * (i + 0.5) / Mq);
        orc_chebyshev_fill_array(x_i, M, T, 1);
        double g = 0.0;
        for (int m = 0; m < M; m++) g += (m == 0 ? 1 : 2) * (mu[m] * kern[m]) * T[m];
        gamma[i] = g;
    }
    free(kern);
    free(T);
}

/* density_product with f = fermi_energy (electronic_grand_energy,
 * fastkpm.cpp:314-322 and 440-443). */
double orc_electronic_grand_energy(int Mq, const double* gamma, double lo, double hi, double kT,
                                   double mu) {
    double avg = (hi + lo) / 2.0, mag = (hi - lo) / 2.0, r = 0.0;
    for (int i = 0; i < Mq; i++) {
        double x_i = cos(ORC_PI * (i + 0.5) / Mq);
        r += gamma[i] * orc_fermi_energy(x_i * mag + avg, kT, mu);
    }
    return r / Mq;
}

/* ------------------------------------------------------------------------- */
/* Symmetric tridiagonal extreme eigenvalues (replaces arma::eig_sym in        */
/* energy_scale, engine_cpu.cpp:30-46).  Sturm-sequence bisection.             */
/* ------------------------------------------------------------------------- */
static int orc_sturm_count(int n, const double* a, const double* b2, double x) {
    /* number of eigenvalues < x; b2[j] = beta_j^2 (j >= 1) */
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) cnt++;
    for (int j = 1; j < n; j++) {
        if (q == 0) q = 1e-300;
        q = a[j] - x - b2[j] / q;
        if (q < 0) cnt++;
    }
    return cnt;
}

static double orc_bisect_kth(int n, const double* a, const double* b2, double lo, double hi, int k) {
    /* k-th smallest eigenvalue (0-based) within [lo, hi] */
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (orc_sturm_count(n, a, b2, mid) > k) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

/* alpha[0..n-1] diagonal, beta[1..n-1] off-diagonal (beta[0] unused). */
void orc_tridiag_extremes(int n, const double* alpha, const double* beta, double* emin, double* emax) {
    double* b2 = (double*)malloc(sizeof(double) * (size_t)n);
    double glo = alpha[0], ghi = alpha[0];
    b2[0] = 0;
    for (int j = 1; j < n; j++) b2[j] = beta[j] * beta[j];
    for (int j = 0; j < n; j++) {
        double r = (j > 0 ? fabs(beta[j]) : 0) + (j + 1 < n ? fabs(beta[j + 1]) : 0);
        if (alpha[j] - r < glo) glo = alpha[j] - r;
        if (alpha[j] + r > ghi) ghi = alpha[j] + r;
    }
    double pad = 1e-12 * (fabs(glo) + fabs(ghi) + 1.0);
    *emin = orc_bisect_kth(n, alpha, b2, glo - pad, ghi + pad, 0);
    *emax = orc_bisect_kth(n, alpha, b2, glo - pad, ghi + pad, n - 1);
    free(b2);
}

/* ------------------------------------------------------------------------- */
/* The CPU engine (engine_cpu.cpp).  Vectors are N x S, row-major internally.  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nb, b, n, nnzb;
    int *row_ptr, *col_idx;
    cplx* hs; /* scaled BSR values, column-major blocks */
    double avg, mag;
    int s;
    cplx *R, *a0, *a1, *a2;
    int have_alpha; /* moments() state valid for autodiff */
} orc_engine;

orc_engine* orc_engine_new(void) { return (orc_engine*)calloc(1, sizeof(orc_engine)); }

static void orc_engine_free_vecs(orc_engine* e) {
    free(e->R); free(e->a0); free(e->a1); free(e->a2);
    e->R = e->a0 = e->a1 = e->a2 = NULL;
}

void orc_engine_free(orc_engine* e) {
    if (!e) return;
    free(e->row_ptr); free(e->col_idx); free(e->hs);
    orc_engine_free_vecs(e);
    free(e);
}

/* Engine_CPU::set_H (engine_cpu.cpp:53-61): Hs = (H - avg I) / mag.  All
 * diagonal blocks must be present (as the GPU engine asserts,
 * engine_cusparse.cpp:716-727); returns -1 otherwise. */
int orc_engine_set_H(orc_engine* e, int nb, int b, int nnzb, const int* row_ptr, const int* col_idx,
                     const cplx* val, double lo, double hi) {
    int bb = b * b;
    free(e->row_ptr); free(e->col_idx); free(e->hs);
    e->nb = nb; e->b = b; e->n = nb * b; e->nnzb = nnzb;
    e->avg = (hi + lo) / 2.0; e->mag = (hi - lo) / 2.0;
    e->row_ptr = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
    e->col_idx = (int*)malloc(sizeof(int) * (size_t)(nnzb > 0 ? nnzb : 1));
    e->hs = (cplx*)malloc(sizeof(cplx) * (size_t)(nnzb > 0 ? nnzb : 1) * bb);
    memcpy(e->row_ptr, row_ptr, sizeof(int) * (size_t)(nb + 1));
    memcpy(e->col_idx, col_idx, sizeof(int) * (size_t)nnzb);
    int diag = 0;
    for (int i = 0; i < nb; i++)
        for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) {
            for (int q = 0; q < bb; q++) e->hs[(size_t)p * bb + q] = val[(size_t)p * bb + q];
            if (col_idx[p] == i) {
                diag++;
                for (int d = 0; d < b; d++) e->hs[(size_t)p * bb + d * b + d] -= e->avg;
            }
            for (int q = 0; q < bb; q++) e->hs[(size_t)p * bb + q] /= e->mag;
        }
    e->have_alpha = 0;
    return diag == nb ? 0 : -1;
}

/* R given column-major n x s (the layout of Engine::R, fastkpm.h:188). */
int orc_engine_set_R(orc_engine* e, int n, int s, const cplx* R_colmajor) {
    orc_engine_free_vecs(e);
    e->s = s;
    size_t ns = (size_t)n * (size_t)s;
    e->R = (cplx*)malloc(sizeof(cplx) * ns);
    e->a0 = (cplx*)calloc(ns, sizeof(cplx));
    e->a1 = (cplx*)calloc(ns, sizeof(cplx));
    e->a2 = (cplx*)calloc(ns, sizeof(cplx));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < s; j++) e->R[(size_t)i * s + j] = R_colmajor[(size_t)j * n + i];
    e->have_alpha = 0;
    return 0;
}

/* y = alpha * Hs x + beta * z  (rows r0..r1 of the N x S block), the product
 * Hs * a of engine_cpu.cpp:72-76 restricted to a row range. */
static void orc_spmm_rows(const orc_engine* e, double alpha, const cplx* x, double beta, const cplx* z,
                          cplx* y, int ib0, int ib1, int c0, int c1) {
    int b = e->b, bb = b * b, s = e->s;
    cplx acc[64];
    for (int i = ib0; i < ib1; i++)
        for (int bi = 0; bi < b; bi++) {
            size_t r = (size_t)(b * i + bi);
            for (int c = c0; c < c1; c += 64) {
                int w = (c1 - c) < 64 ? (c1 - c) : 64;
                for (int l = 0; l < w; l++) acc[l] = 0;
                for (int p = e->row_ptr[i]; p < e->row_ptr[i + 1]; p++) {
                    int j = e->col_idx[p];
                    for (int bj = 0; bj < b; bj++) {
                        cplx h = e->hs[(size_t)p * bb + bj * b + bi];
                        const cplx* xr = &x[(size_t)(b * j + bj) * s + c];
                        for (int l = 0; l < w; l++) acc[l] += h * xr[l];
                    }
                }
                cplx* yr = &y[r * s + c];
                if (z) {
                    const cplx* zr = &z[r * s + c];
                    for (int l = 0; l < w; l++) yr[l] = alpha * acc[l] + beta * zr[l];
                } else {
                    for (int l = 0; l < w; l++) yr[l] = alpha * acc[l];
                }
            }
        }
}

static double orc_re_cdot(size_t len, const cplx* x, const cplx* y) {
    /* Re sum conj(x) y, i.e. std::real(arma::cdot(x, y)) */
    double r = 0;
    for (size_t k = 0; k < len; k++) r += creal(x[k]) * creal(y[k]) + cimag(x[k]) * cimag(y[k]);
    return r;
}

/* Engine_CPU::moments (engine_cpu.cpp:63-84), columns c0..c1 of R only (the
 * S-column sharding of Engine_Threaded, engine_cusparse.cpp:1067-1117, when
 * c0..c1 is a slice; the full call uses 0..s).  M must be even. */
static void orc_moments_cols(orc_engine* e, int M, double* mu, int c0, int c1) {
    int n = e->n, s = e->s;
    /* work on the column slice in place (rows are length s, slice c0..c1) */
    for (int i = 0; i < n; i++)
        for (int c = c0; c < c1; c++) e->a0[(size_t)i * s + c] = e->R[(size_t)i * s + c];
    orc_spmm_rows(e, 1.0, e->R, 0.0, NULL, e->a1, 0, e->nb, c0, c1);
    double m0 = 0, m1 = 0;
    for (int i = 0; i < n; i++) {
        m0 += orc_re_cdot((size_t)(c1 - c0), &e->a0[(size_t)i * s + c0], &e->a0[(size_t)i * s + c0]);
        m1 += orc_re_cdot((size_t)(c1 - c0), &e->a1[(size_t)i * s + c0], &e->a0[(size_t)i * s + c0]);
    }
    mu[0] = m0;
    mu[1] = m1;
    for (int m = 1; m < M / 2; m++) {
        orc_spmm_rows(e, 2.0, e->a1, -1.0, e->a0, e->a2, 0, e->nb, c0, c1); /* a2 = 2 Hs a1 - a0 */
        cplx* t = e->a0; e->a0 = e->a1; e->a1 = e->a2; e->a2 = t;
        double d0 = 0, d1 = 0;
        for (int i = 0; i < n; i++) {
            d0 += orc_re_cdot((size_t)(c1 - c0), &e->a0[(size_t)i * s + c0], &e->a0[(size_t)i * s + c0]);
            d1 += orc_re_cdot((size_t)(c1 - c0), &e->a1[(size_t)i * s + c0], &e->a0[(size_t)i * s + c0]);
        }
        mu[2 * m] = 2 * d0 - mu[0];
        mu[2 * m + 1] = 2 * d1 - mu[1];
    }
}

int orc_engine_moments(orc_engine* e, int M, double* mu) {
    if (M % 2 != 0 || M < 2 || !e->R) return -1;
    orc_moments_cols(e, M, mu, 0, e->s);
    e->have_alpha = M;
    return 0;
}

/* Moments over the column slice c0..c1 only (a bounded CPU-baseline sample
 * and the per-device share of Engine_Threaded). Leaves no autodiff state. */
int orc_engine_moments_cols(orc_engine* e, int M, double* mu, int c0, int c1) {
    if (M % 2 != 0 || M < 2 || !e->R || c0 < 0 || c1 > e->s || c0 >= c1) return -1;
    orc_moments_cols(e, M, mu, c0, c1);
    e->have_alpha = 0;
    return 0;
}

/* Engine_CPU::outer_product (engine_cpu.cpp:213-226):
 * D_k(bi,bj) += alpha * sum_l A[b i + bi, l] conj(B[b j + bj, l]). */
/* y = A x for one column (length n = nb * b) of a BSR matrix A with column-
 * major blocks (SpMatBsr layout, fastkpm.h:64-106). */
static void orc_bsr_apply_col(int nb, int b, const int* rp, const int* ci, const cplx* val,
                              const cplx* x, cplx* y) {
    int bb = b * b;
    for (int i = 0; i < nb; i++)
        for (int bi = 0; bi < b; bi++) {
            cplx acc = 0;
            for (int p = rp[i]; p < rp[i + 1]; p++) {
                int j = ci[p];
                for (int bj = 0; bj < b; bj++) acc += val[(size_t)p * bb + bj * b + bi] * x[(size_t)b * j + bj];
            }
            y[(size_t)b * i + bi] = acc;
        }
}

/* Engine_CPU::moments2_v1 (engine_cpu.cpp:86-160): for every probe column r_k,
 * alpha_m = T_m(Hs) r_k and atild_m = T_m(Hs) j1 r_k (m < M), and
 * mu[m1][m2] += cdot(alpha_m1, j2 atild_m2) = sum_i conj(alpha_m1[i]) (j2 atild_m2)[i].
 * j1, j2 are BSR on H's block rows, unscaled.  mu is M x M row-major.  The
 * reference's column chunking (a_chunk_ncols) only bounds its memory; the sum
 * is the same. */
int orc_engine_moments2_v1(orc_engine* e, int M, int nnzb1, const int* rp1, const int* ci1,
                           const cplx* v1, int nnzb2, const int* rp2, const int* ci2, const cplx* v2,
                           cplx* mu) {
    (void)nnzb1; (void)nnzb2;
    if (M < 2 || M % 2 != 0 || !e->R) return -1;
    int n = e->n, b = e->b, nb = e->nb, s = e->s;
    cplx* al = (cplx*)malloc(sizeof(cplx) * (size_t)M * n);
    cplx* at = (cplx*)malloc(sizeof(cplx) * (size_t)M * n);
    cplx* bt = (cplx*)malloc(sizeof(cplx) * (size_t)M * n);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (size_t q = 0; q < (size_t)M * M; q++) mu[q] = 0;
    for (int k = 0; k < s; k++) {
        for (int i = 0; i < n; i++) al[i] = e->R[(size_t)i * s + k];
        orc_bsr_apply_col(nb, b, rp1, ci1, v1, al, at); /* atild_0 = j1 r */
        for (int m = 1; m < M; m++) {
            cplx* am = al + (size_t)m * n;
            cplx* tm = at + (size_t)m * n;
            orc_bsr_apply_col(nb, b, e->row_ptr, e->col_idx, e->hs, al + (size_t)(m - 1) * n, am);
            orc_bsr_apply_col(nb, b, e->row_ptr, e->col_idx, e->hs, at + (size_t)(m - 1) * n, tm);
            if (m >= 2)
                for (int i = 0; i < n; i++) {
                    am[i] = 2.0 * am[i] - al[(size_t)(m - 2) * n + i];
                    tm[i] = 2.0 * tm[i] - at[(size_t)(m - 2) * n + i];
                }
        }
        for (int m = 0; m < M; m++) orc_bsr_apply_col(nb, b, rp2, ci2, v2, at + (size_t)m * n, bt + (size_t)m * n);
        for (int m1 = 0; m1 < M; m1++)
            for (int m2 = 0; m2 < M; m2++) {
                cplx acc = 0;
                const cplx* x = al + (size_t)m1 * n;
                const cplx* y = bt + (size_t)m2 * n;
                for (int i = 0; i < n; i++) acc += conj(x[i]) * y[i];
                mu[(size_t)m1 * M + m2] += acc;
            }
    }
    free(al); free(at); free(bt); free(tmp);
    return 0;
}

static void orc_outer_product(const orc_engine* e, double alpha, const cplx* A, const cplx* B, int nnzbD,
                              const int* D_row_idx, const int* D_col_idx, cplx* D_val) {
    int b = e->b, bb = b * b, s = e->s;
    for (int k = 0; k < nnzbD; k++) {
        int i = D_row_idx[k], j = D_col_idx[k];
        cplx* v = &D_val[(size_t)k * bb];
        for (int bj = 0; bj < b; bj++)
            for (int bi = 0; bi < b; bi++) {
                const cplx* Br = &B[(size_t)(b * j + bj) * s];
                const cplx* Ar = &A[(size_t)(b * i + bi) * s];
                cplx acc = 0;
                for (int l = 0; l < s; l++) acc += conj(Br[l]) * Ar[l];
                v[b * bj + bi] += alpha * acc;
            }
    }
}

/* Engine_CPU::stoch_matrix (engine_cpu.cpp:229-253). c has M entries. */
int orc_engine_stoch_matrix(orc_engine* e, int M, const double* c, int nnzbD, const int* D_row_idx,
                            const int* D_col_idx, cplx* D_val) {
    if (!e->R || M < 2) return -1;
    size_t ns = (size_t)e->n * e->s;
    cplx* xi = (cplx*)malloc(sizeof(cplx) * ns);
    memcpy(e->a0, e->R, sizeof(cplx) * ns);
    orc_spmm_rows(e, 1.0, e->R, 0.0, NULL, e->a1, 0, e->nb, 0, e->s);
    for (size_t k = 0; k < ns; k++) xi[k] = c[0] * e->a0[k] + c[1] * e->a1[k];
    for (int m = 2; m < M; m++) {
        orc_spmm_rows(e, 2.0, e->a1, -1.0, e->a0, e->a2, 0, e->nb, 0, e->s);
        for (size_t k = 0; k < ns; k++) xi[k] += c[m] * e->a2[k];
        cplx* t = e->a0; e->a0 = e->a1; e->a1 = e->a2; e->a2 = t;
    }
    memset(D_val, 0, sizeof(cplx) * (size_t)nnzbD * e->b * e->b);
    orc_outer_product(e, 0.5, e->R, xi, nnzbD, D_row_idx, D_col_idx, D_val);
    orc_outer_product(e, 0.5, xi, e->R, nnzbD, D_row_idx, D_col_idx, D_val);
    free(xi);
    e->have_alpha = 0; /* engine_cpu.cpp:250-252 */
    return 0;
}

/* Engine_CPU::autodiff_matrix (engine_cpu.cpp:255-304).  Requires moments(M)
 * with M == c size just before (fastkpm.h:232).  D pattern needs row_ptr for
 * symmetrize.  Returns -1 on a missing precondition, -2 on an asymmetric
 * pattern. */
int orc_engine_autodiff_matrix(orc_engine* e, int M, const double* c, int nbD, int nnzbD,
                               const int* D_row_ptr, const int* D_row_idx, const int* D_col_idx,
                               cplx* D_val) {
    (void)nbD;
    if (!e->R || e->have_alpha != M) return -1;
    int n = e->n, s = e->s;
    size_t ns = (size_t)n * s;
    double diag = c[1];
    for (int m = 1; m < M / 2; m++) diag -= c[2 * m + 1];
    memset(D_val, 0, sizeof(cplx) * (size_t)nnzbD * e->b * e->b);
    orc_outer_product(e, diag, e->R, e->R, nnzbD, D_row_idx, D_col_idx, D_val);
    cplx* b0 = (cplx*)calloc(ns, sizeof(cplx));
    cplx* b1 = (cplx*)calloc(ns, sizeof(cplx));
    cplx* b2 = (cplx*)calloc(ns, sizeof(cplx));
    if (M > 2)
        for (size_t k = 0; k < ns; k++) b0[k] = 2 * c[M - 1] * e->a0[k];
    for (int m = M / 2 - 1; m >= 1; m--) {
        orc_outer_product(e, 2.0, e->a0, b0, nnzbD, D_row_idx, D_col_idx, D_val);
        /* a2 = a1; a1 = a0; a0 = 2 Hs a1 - a2 */
        cplx* t = e->a2; e->a2 = e->a1; e->a1 = e->a0; e->a0 = t;
        orc_spmm_rows(e, 2.0, e->a1, -1.0, e->a2, e->a0, 0, e->nb, 0, s);
        /* b2 = b1; b1 = b0; b0 = 4 c[2m] a1 + 2 c[2m+1] a2 + 2 Hs b1 - b2 (+ 2 c[2m-1] a0) */
        t = b2; b2 = b1; b1 = b0; b0 = t;
        orc_spmm_rows(e, 2.0, b1, -1.0, b2, b0, 0, e->nb, 0, s);
        for (size_t k = 0; k < ns; k++) {
            cplx v = 4 * c[2 * m] * e->a1[k] + 2 * c[2 * m + 1] * e->a2[k];
            b0[k] = v + b0[k];
            if (m > 1) b0[k] += 2 * c[2 * m - 1] * e->a0[k];
        }
    }
    orc_outer_product(e, 1.0, e->a0, b0, nnzbD, D_row_idx, D_col_idx, D_val);
    free(b0); free(b1); free(b2);
    int rc = orc_bsr_symmetrize(e->b, nnzbD, D_row_ptr, D_row_idx, D_col_idx, D_val);
    orc_bsr_scale(nnzbD * e->b * e->b, D_val, 1.0 / e->mag);
    e->have_alpha = 0; /* engine_cpu.cpp:301-303 */
    return rc == 0 ? 0 : -2;
}

/* Engine_CPU::energy_scale (engine_cpu.cpp:14-51) with the start vector given
 * (the reference draws it with arma::randn, which cannot be reproduced here).
 * H is the unscaled BSR matrix; v_start has n entries. */
int orc_energy_scale(int nb, int b, int nnzb, const int* row_ptr, const int* col_idx, const cplx* val,
                     double extend, int iters, const cplx* v_start, double* alpha_out, double* beta_out,
                     double* lo, double* hi) {
    orc_engine h;
    memset(&h, 0, sizeof(h));
    h.nb = nb; h.b = b; h.n = nb * b; h.nnzb = nnzb; h.s = 1;
    h.row_ptr = (int*)row_ptr; h.col_idx = (int*)col_idx; h.hs = (cplx*)val;
    int n = h.n;
    cplx* v0 = (cplx*)calloc((size_t)n, sizeof(cplx));
    cplx* v1 = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    double* al = alpha_out;
    double* be = beta_out;
    memcpy(v1, v_start, sizeof(cplx) * (size_t)n);
    double nrm = sqrt(orc_re_cdot((size_t)n, v1, v1));
    for (int k = 0; k < n; k++) v1[k] /= nrm;
    be[0] = 0;
    for (int j = 1; j < iters; j++) {
        orc_spmm_rows(&h, 1.0, v1, 0.0, NULL, w, 0, nb, 0, 1);
        al[j - 1] = orc_re_cdot((size_t)n, w, v1);
        for (int k = 0; k < n; k++) w[k] = w[k] - al[j - 1] * v1[k] - be[j - 1] * v0[k];
        be[j] = sqrt(orc_re_cdot((size_t)n, w, w));
        memcpy(v0, v1, sizeof(cplx) * (size_t)n);
        for (int k = 0; k < n; k++) v1[k] = w[k] / be[j];
    }
    orc_spmm_rows(&h, 1.0, v1, 0.0, NULL, w, 0, nb, 0, 1);
    al[iters - 1] = orc_re_cdot((size_t)n, w, v1);
    double emin, emax;
    orc_tridiag_extremes(iters, al, be, &emin, &emax);
    double slack = extend * (emax - emin);
    *lo = emin - slack;
    *hi = emax + slack;
    free(v0); free(v1); free(w);
    return 0;
}
