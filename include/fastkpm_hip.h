/*
 * fastkpm_hip.h -- C-ABI of the MI355X (gfx950) stochastic Chebyshev engine.
 *
 * This is the drop-in boundary below the C++ fkpm::Engine<T> interface
 * (include/fastkpm.h).  It replaces the device half of the reference GPU
 * engine, Engine_cuSPARSE (src/engine_cusparse.cpp:486-1040), and its one CUDA
 * kernel (src/outer_product.cu:38-68).  Plain C types only: no C++ or torch
 * types cross it.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative FKPM_E* code; the
 *    message of the last failure on the calling thread is fkpm_last_error().
 *    (The reference aborts / exits instead, engine_cusparse.cpp:25,60-66.)
 *  - A handle owns one HIP device, one HIP stream and all device buffers.  A
 *    handle is not thread-safe (neither is a reference Engine).
 *  - Host arrays are caller-owned.  Value arrays hold the engine element type
 *    chosen at fkpm_dev_open (FKPM_F32, FKPM_F64, FKPM_C32, FKPM_C64 = float,
 *    double, std::complex<float>, std::complex<double>), complex values
 *    interleaved (re, im).
 *  - BSR blocks are b x b, column-major inside a block, block rows sorted by
 *    column, as SpMatBsr (src/fastkpm.h:64-106).
 *  - Calls are synchronous with respect to the host arrays they take: results
 *    are in host memory on return, unless a NULL result pointer asks for the
 *    result to stay device-resident (fkpm_get_D copies it later).
 */
#ifndef FASTKPM_HIP_H
#define FASTKPM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of host arrays (the T of Engine<T>, src/engine.cpp:118-121) */
#define FKPM_F32 0
#define FKPM_F64 1
#define FKPM_C32 2
#define FKPM_C64 3

/* device storage precision of the Chebyshev vectors.  Accumulation of moment
 * dots, xi and D is FP64 in both. */
#define FKPM_PREC_FP64 0 /* vectors and H in FP64 */
#define FKPM_PREC_FP32 1 /* vectors and H in FP32, row sums and dots in FP64 */

/* error codes */
#define FKPM_OK 0
#define FKPM_E_ARG (-1)     /* invalid argument / shape mismatch */
#define FKPM_E_STATE (-2)   /* call order violated (e.g. autodiff without moments) */
#define FKPM_E_HIP (-3)     /* HIP runtime error */
#define FKPM_E_NODEV (-4)   /* no usable HIP device */
#define FKPM_E_UNSUP (-5)   /* unsupported configuration */

typedef struct fkpm_handle fkpm_handle;

/* Message of the last failed call on this thread ("" if none). */
const char* fkpm_last_error(void);

/* Number of visible HIP devices (replaces the cudaGetDeviceCount probe of
 * mk_engine_cuSPARSE, engine_cusparse.cpp:1171-1215).  Returns 0 when none. */
int fkpm_device_count(void);

/* Open an engine on `device` for element type `dtype` and storage precision
 * `prec`.  Replaces the Engine_cuSPARSE constructor (engine_cusparse.cpp:514-523).
 * FKPM_PREC_FP32 with an FP64 dtype runs the FP32-mixed recursion on FP64
 * host data. */
int fkpm_dev_open(int device, int dtype, int prec, fkpm_handle** out);

/* Release all device memory and the stream (engine_cusparse.cpp:525-531). */
int fkpm_dev_close(fkpm_handle* h);

/* The HIP stream (hipStream_t) all work of this handle is queued on. */
void* fkpm_stream(fkpm_handle* h);

/* Wait until all work queued on the handle's stream has finished. */
int fkpm_synchronize(fkpm_handle* h);

/* Lanczos estimate of the spectral bounds of the unscaled H, FP64 on the
 * device; the iters x iters tridiagonal eigenproblem is solved on the host.
 * Replaces Engine_cuSPARSE::energy_scale (engine_cusparse.cpp:598-673).
 * v_start: n = nb*b start-vector entries of the engine type (the reference
 * draws it with arma::randn).  Writes lo/hi = [lmin - slack, lmax + slack],
 * slack = extend*(lmax - lmin). */
int fkpm_energy_scale(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr,
                      const int* col_idx, const void* val, double extend, int iters,
                      const void* v_start, double* lo, double* hi);

/* Upload the scaled Hamiltonian Hs = (H - avg I)/mag.  Every diagonal block
 * must be present.  Replaces Engine_cuSPARSE::set_H (engine_cusparse.cpp:701-730). */
int fkpm_set_H(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr,
               const int* col_idx, const void* val, double avg, double mag);

/* Upload R (n x s, column-major, the layout of Engine::R) and allocate the
 * recursion state.  Replaces Engine_cuSPARSE::transfer_R (engine_cusparse.cpp:675-699). */
int fkpm_set_R(fkpm_handle* h, int n, int s, const void* R_colmajor);

/* Chebyshev moments mu[0..M-1], M even, with the two-moments-per-product
 * identity.  Leaves alpha_{M/2-1}, alpha_{M/2} on the device for
 * fkpm_autodiff_matrix.  Replaces Engine_cuSPARSE::moments (engine_cusparse.cpp:755-791). */
int fkpm_moments(fkpm_handle* h, int M, double* mu);

/* Block pattern of D (nbD block rows; row_ptr[nbD+1], row_idx/col_idx[nnzbD]),
 * the sparsity structure of SpMatBsr D in stoch_matrix/autodiff_matrix.  Kept
 * until replaced; a pattern equal to H's enables the fused autodiff kernel. */
int fkpm_set_D_pattern(fkpm_handle* h, int nbD, int nnzbD, const int* row_ptr,
                       const int* row_idx, const int* col_idx);

/* D = (R xi^+ + xi R^+)/2 on D's pattern with xi = sum_{m<M} c_m T_m(Hs) R.
 * D_val: nnzbD*b*b values of the engine type, or NULL to leave D on the
 * device.  Invalidates the moments state.  Replaces
 * Engine_cuSPARSE::stoch_matrix (engine_cusparse.cpp:910-953). */
int fkpm_stoch_matrix(fkpm_handle* h, int M, const double* c, void* D_val);

/* D = d tr g / dH^T on D's pattern by reverse-mode differentiation of the
 * moment recursion, symmetrized and scaled by 1/mag.  Requires fkpm_moments
 * with the same M just before.  D_val as in fkpm_stoch_matrix.  Replaces
 * Engine_cuSPARSE::autodiff_matrix (engine_cusparse.cpp:955-1039). */
int fkpm_autodiff_matrix(fkpm_handle* h, int M, const double* c, void* D_val);

/* Two-current moments mu[m1][m2] = sum_k <T_m1(Hs) r_k, j2 T_m2(Hs) j1 r_k>
 * (M even), written row-major as M*M complex doubles (re, im), for the
 * current operators j1, j2 given as BSR matrices on H's block rows (nb block
 * rows of b x b blocks of the engine type, any pattern).  The probe columns
 * are processed in chunks of at most col_chunk (0: all), halved until the
 * 2M + 3 vectors of a chunk fit the device-memory budget (half the free
 * memory, at most 64 GiB, or the byte count in the environment variable
 * FKPM_MOMENTS2_BYTES); FKPM_E_ARG if one column does not fit.  Replaces
 * Engine_cuSPARSE::moments2_v1 (engine_cusparse.cpp:793-902).  The state left
 * by fkpm_moments (for fkpm_autodiff_matrix) is kept. */
int fkpm_moments2(fkpm_handle* h, int M, int nnzb1, const int* rp1, const int* ci1,
                  const void* v1, int nnzb2, const int* rp2, const int* ci2, const void* v2,
                  int col_chunk, double* mu);

/* Copy the last device-resident D to the host (engine type). */
int fkpm_get_D(fkpm_handle* h, void* D_val);

/* Per-call timing of the dominant recursion kernels, measured with HIP events
 * on the handle's stream when enabled: total milliseconds and launch counts of
 * the Chebyshev step kernels of the last fkpm_moments / fkpm_stoch_matrix /
 * fkpm_autodiff_matrix call. */
int fkpm_set_timing(fkpm_handle* h, int on);
int fkpm_get_timing(fkpm_handle* h, double* ms_moments, int* n_moments, double* ms_stoch,
                    int* n_stoch, double* ms_autodiff, int* n_autodiff);

/* Kernel geometry in use: lanes per block row, packets per lane, padded S,
 * workgroups per step launch, and whether the fused autodiff path is active. */
int fkpm_get_config(fkpm_handle* h, int* lanes, int* kq, int* s_padded, int* n_wg,
                    int* fused_autodiff);

/* Force one kernel geometry (lanes per block row, packets per lane, workgroup
 * cap; lanes = 0 returns to the default / tuned choice) for every kernel kind.
 * Takes effect at the next call; discards the moments state. */
int fkpm_set_geometry(fkpm_handle* h, int lanes, int kq, int max_wg);

/* Chebyshev moments with no device allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * work of fkpm_set_H + fkpm_set_R + fkpm_moments in one call, on inputs that
 * are already in device memory: H as BSR (nb block rows of b x b
 * complex-double blocks, column-major blocks, rows sorted, every diagonal
 * block present; row_ptr, col_idx, val all device pointers; the pattern is not
 * checked), scaled as (H - avg) / mag; R (n x s complex double, column-major,
 * n = nb * b, device); mu: M doubles in device memory.  The work is queued on
 * `stream` (a hipStream_t; NULL: the null stream) and the call returns without
 * waiting; prec = FKPM_PREC_FP64 or FKPM_PREC_FP32.  Working memory is a
 * static device buffer of FKPM_NOALLOC_SCRATCH_MB MiB per device (the scaled
 * H, the packed R, three recursion vectors and the moment partials); a
 * problem that needs more returns FKPM_E_ARG.  Every call on one device shares
 * that buffer: calls that overlap on one device (two streams or two threads)
 * race.  The untuned vector-unit kernels run (no matrix cores, no tuner). */
#define FKPM_NOALLOC_SCRATCH_MB 64
int fkpm_moments_noalloc(int device, int prec, int nb, int b, int nnzb, const int* row_ptr,
                         const int* col_idx, const void* val, double avg, double mag, int n,
                         int s, const void* R_colmajor, int M, double* mu, void* stream);

/* Options (fkpm_set_option). */
#define FKPM_OPT_STOCH_METHOD 1 /* value: FKPM_STOCH_AUTO (default), FKPM_STOCH_CLENSHAW or
                                   FKPM_STOCH_FORWARD */
#define FKPM_OPT_AUTOTUNE 2     /* value 1: tune the kernel geometry at first use of a shape
                                   (default 0; the environment variable FKPM_AUTOTUNE=1 at
                                   fkpm_dev_open sets 1) */
#define FKPM_OPT_MFMA 3         /* value 1 (default): matrix-core (MFMA) kernels where the block
                                   shape has them (complex 8x8 blocks); 2: the same, with the
                                   row-resident variants (rows of at most 8 blocks) as the
                                   untuned choice of every step that has one; 0: vector-unit
                                   kernels */
#define FKPM_OPT_STORE_ALPHA 5  /* value 2: fkpm_moments keeps every alpha_k when the table
                                   fits half the free device memory (or the byte count in the
                                   environment variable FKPM_ALPHA_TABLE_BYTES), and
                                   fkpm_autodiff_matrix then reads alpha_{m-1} instead of
                                   recomputing it (one gathered vector per step instead of two);
                                   1 (default): the same when three vectors exceed the
                                   32 MB of L2; 0: the three-vector ring and the
                                   recomputation */
#define FKPM_OPT_SWEEP 4        /* value 0: every kernel kind sweeps the rows per XCD; s in
                                   1..64: as one front in segments of s passes per XCD
                                   (fkpm_get_kernel_sweep); -1 (default): the configuration's
                                   own (the tuner's choice, else one front with s = 2) */
#define FKPM_OPT_SLICE 6        /* value 1: the vector-unit step kernels sweep all their block
                                   rows once per column pass (the probe columns of a row in
                                   slices of one pass width), so the rows an XCD gathers
                                   between a row and its far neighbours take fewer bytes of
                                   its L2; 0: a row's column passes back to back; -1
                                   (default): the configuration's own (the tuner's choice,
                                   else 0) */
#define FKPM_OPT_GRAPH 8        /* value 1: the launches of every step loop (moments, the
                                   stoch_matrix recurrences, the autodiff backward steps)
                                   replay as one hipGraph, captured the first time the same
                                   launches (kernels, grids, argument values) occur and kept
                                   for the handle (the 16 most recent); 0: direct launches;
                                   -1 (default): graphs when the recursion vectors hold at most
                                   2^22 scalars (steps short enough for launches to set the
                                   time); the environment variable FKPM_GRAPH (-1, 0 or 1) at
                                   fkpm_dev_open sets the handle's default */
#define FKPM_OPT_ROW_ORDER 9    /* value 1: when H's pattern is a periodic 3-D lattice in
                                   natural order (block row x + n1 (y + n2 z), every side at
                                   least 8; recognised at fkpm_set_H from the block
                                   offsets), the vector-unit step kernels visit the rows
                                   slab by slab (one x-slab of about n1/8 sites per XCD), in
                                   y-tiles along z, so the next plane's rows are still in the
                                   XCD's L2; results per row are the same as in row order, the
                                   moment sums add in another order; 0: row order; -1
                                   (default): the configuration's own (the tuner's choice,
                                   else 0).  Patterns of another shape keep row order. */
#define FKPM_OPT_ALPHA_TABLE_MB 7 /* value >= 0: the most MiB the table of forward vectors
                                   (FKPM_OPT_STORE_ALPHA) may hold on this handle; -1
                                   (default): a quarter of the device's memory and at most
                                   half of what is free when fkpm_moments runs (the
                                   environment variable FKPM_ALPHA_TABLE_BYTES, read at each
                                   call, takes precedence over both) */
#define FKPM_STOCH_CLENSHAW 0   /* xi by the Clenshaw recurrence: one 2-vector pass per order */
#define FKPM_STOCH_FORWARD 1    /* xi accumulated in FP64 along the forward recursion, the
                                   reference's summation order (engine_cpu.cpp:236-244) */
#define FKPM_STOCH_TABLE 2      /* (reported only) xi from the forward vectors a preceding
                                   fkpm_moments kept (FKPM_OPT_STORE_ALPHA), by
                                   T_{K+j} = 2 T_K T_j - T_{K-j}: one pass over the table and
                                   K = ceil(M/2) Chebyshev steps instead of M - 1 */
#define FKPM_STOCH_AUTO 3       /* FKPM_STOCH_TABLE when the held forward vectors (same H and
                                   R) reach alpha_K, else FKPM_STOCH_CLENSHAW */
int fkpm_set_option(fkpm_handle* h, int option, int value);

/* Kernel kinds of fkpm_get_kernel_config. */
#define FKPM_KERNEL_MOMENTS 0  /* forward step (also the Clenshaw steps of stoch_matrix) */
#define FKPM_KERNEL_XI 1       /* forward step with FP64 xi accumulation */
#define FKPM_KERNEL_AUTODIFF 2 /* backward step with the fused outer product */
#define FKPM_KERNEL_OUTER 3    /* sparse outer product on D's pattern */
#define FKPM_KERNEL_MOMENTS_TABLE 4  /* forward step writing the kept forward vectors
                                        (FKPM_OPT_STORE_ALPHA) */
#define FKPM_KERNEL_AUTODIFF_TABLE 5 /* backward step reading them */

/* Autotune the geometry (lanes per block row, packets per lane, workgroup
 * count, nontemporal hints) of every kernel kind on the current H and R by
 * timing the candidates on the device; the result is kept for this handle and
 * for every later handle of the same problem shape in the process.  Discards
 * the moments state.  No counterpart in the reference. */
int fkpm_autotune(fkpm_handle* h);

/* Geometry in use for one kernel kind. */
int fkpm_get_kernel_config(fkpm_handle* h, int kind, int* lanes, int* kq, int* n_wg, int* nt);

/* *mfma = 0 if that kernel kind runs on the vector units, 1 on the matrix
 * cores, 2 on the matrix cores in the row-resident form (rows of at most 8
 * blocks).  No counterpart in the reference. */
int fkpm_get_kernel_mfma(fkpm_handle* h, int kind, int* mfma);

/* How the workgroups of a kernel kind cover the block rows: 0 = each of the 8
 * XCDs sweeps its own contiguous region (neighbour rows stay in that XCD's
 * L2); s >= 1 = all workgroups sweep one front over all rows (far neighbours
 * stay in the shared last-level cache), the rows dealt to the XCDs in
 * segments of s workgroup passes.  The autotuner times 0, 1, 2 and 8. */
int fkpm_get_kernel_sweep(fkpm_handle* h, int kind, int* sweep);

/* *stored = 1 when the last fkpm_moments kept its forward vectors for the
 * next fkpm_autodiff_matrix (FKPM_OPT_STORE_ALPHA). */
int fkpm_get_alpha_table(fkpm_handle* h, int* stored);

/* Step loops replayed as a hipGraph and graphs captured so far (FKPM_OPT_GRAPH). */
int fkpm_get_graph_stats(fkpm_handle* h, long long* launches, long long* captures);

/* The lattice row order fkpm_set_H would build for a block pattern
 * (FKPM_OPT_ROW_ORDER), a host computation that needs no device: returns 1
 * and fills dims[3] = (n1, n2, n3) and order[nb] (the block row at each sweep
 * position: XCD k's slab, x in [floor(k n1/8), floor((k+1) n1/8)), after
 * those of XCDs 0 .. k-1) when the pattern
 * is a periodic 3-D lattice in natural order; 0 (order and dims untouched)
 * otherwise; FKPM_E_ARG for bad arguments.  No counterpart in the reference:
 * a diagnostic of this library's sweep. */
int fkpm_lattice_order(int nb, int nnzb, const int* row_ptr, const int* col_idx, int* order,
                       int* dims);

/* *method = the FKPM_STOCH_* method the last fkpm_stoch_matrix ran (-1 if
 * none yet). */
int fkpm_get_stoch_method(fkpm_handle* h, int* method);

/* r_compact: R has at most one nonzero per row (probing colours) and the
 * Clenshaw source term reads the compact form; tuned: the geometry came from
 * the autotuner; moments_M: M of the held moments state (0 if none). */
int fkpm_get_state(fkpm_handle* h, int* r_compact, int* tuned, int* moments_M);

/* Bandwidth probe: copies `bytes` (a multiple of 16) between two device
 * buffers `reps` times with a plain copy kernel (k_copy16) and with a
 * streaming one (k_copy16_nt: four packets per thread, nontemporal hints),
 * and reports the better read+write GB/s.  The measured ceiling the step
 * kernels are set against; k_copy16 is also the known-size kernel that
 * calibrates HBM byte counters.  No counterpart in the reference. */
int fkpm_bandwidth_probe(fkpm_handle* h, size_t bytes, int reps, double* gbs);

/* ------------------------------------------------------------------------
 * Batched engine (serving runtime).  No counterpart in the reference: it runs
 * many independent Engine solves -- moments(M) then autodiff_matrix(c, D) on
 * H's own pattern, the energy-and-force request of a simulation loop -- on one
 * GPU at once.  A scheduler thread admits queued requests while a slot is free
 * and the request's HBM footprint fits the batch's arena, advances every
 * admitted request by one Chebyshev step per tick in batched launches
 * (captured in hipGraphs), and retires finished ones.  Results do not depend on
 * what else is in flight.  Complex-double inputs (the cx_double Engine types),
 * block sizes 2 and 8, every diagonal block present, pattern structurally
 * symmetric.
 * ------------------------------------------------------------------------ */
typedef struct fkpm_batch fkpm_batch;

#define FKPM_BATCH_OPT_GRAPH 1 /* value 1 (default): capture runs of ticks in hipGraphs */
#define FKPM_BATCH_OPT_TICKS 2 /* value: most ticks per replay between admissions (16) */

/* HBM bytes one request occupies while admitted (its arena footprint; for
 * 2x2 blocks and nb >= 4096 with room for the lattice row order a 3-D
 * lattice pattern takes, so an upper bound for other patterns). */
int fkpm_batch_footprint(int prec, int nb, int b, int nnzb, int s, int M, size_t* bytes);

/* Open a batched engine on `device` with an arena of `hbm_bytes` and at most
 * `max_slots` requests in flight; prec = FKPM_PREC_FP64 or FKPM_PREC_FP32. */
int fkpm_batch_open(int device, int prec, size_t hbm_bytes, int max_slots, fkpm_batch** out);
/* Stop the scheduler, fail every request still queued or running with
 * FKPM_E_STATE (threads blocked in fkpm_batch_wait on them return that code),
 * wait until those threads have left fkpm_batch_wait, then release the
 * device memory.  No other call may start on the handle once close has
 * begun. */
int fkpm_batch_close(fkpm_batch* b);
int fkpm_batch_set_option(fkpm_batch* b, int option, int value);

/* Queue one request: H (BSR, nb block rows of b x b complex-double blocks,
 * column-major), its energy scale [lo, hi], R (n x s complex double,
 * column-major, n = nb * b), M (even) and the coefficients c[0..M-1] of the
 * energy (expansion_coefficients of fermi_energy).  The inputs are copied.
 * *id identifies the request. */
int fkpm_batch_submit(fkpm_batch* b, int nb, int blk, int nnzb, const int* row_ptr,
                      const int* col_idx, const void* val, double lo, double hi, int n, int s,
                      const void* R_colmajor, int M, const double* c, long long* id);

/* The same with probing vectors in compact form, one nonzero per orbital row
 * (the R of Engine::set_R_correlated): rcol[i] in [-1, s) (-1: zero row) and
 * rval[i] (complex double) for i < n. */
int fkpm_batch_submit_probing(fkpm_batch* b, int nb, int blk, int nnzb, const int* row_ptr,
                              const int* col_idx, const void* val, double lo, double hi, int n,
                              int s, const int* rcol, const void* rval, int M, const double* c,
                              long long* id);

/* The same with the probing vectors drawn on the device: groups[i] in
 * [0, s) is the colour of orbital row i (s < n) and col_seeds[j] the seed of
 * column j's std::mt19937, i.e. the j-th draw of the caller's master engine;
 * R is then bit-identical to Engine::set_R_correlated(groups, rng)
 * (src/engine.cpp:76-89, complex phases).  Only n colours and s seeds cross
 * PCIe. */
int fkpm_batch_submit_seeded(fkpm_batch* b, int nb, int blk, int nnzb, const int* row_ptr,
                             const int* col_idx, const void* val, double lo, double hi, int n,
                             int s, const int* groups, const unsigned* col_seeds, int M,
                             const double* c, long long* id);

/* *done = 1 once the request has finished (or failed). */
int fkpm_batch_poll(fkpm_batch* b, long long id, int* done);

/* Wait for a request and take its results: mu[0..M-1] (moments), *energy =
 * sum c_m mu_m, D_val (nnzb * b * b complex doubles, dE/dH on H's pattern as
 * autodiff_matrix returns it; may be NULL) and times_ms[3] = queued, running,
 * total milliseconds (may be NULL).  The request is forgotten afterwards. */
int fkpm_batch_wait(fkpm_batch* b, long long id, double* mu, double* energy, double* D_val,
                    double* times_ms);

/* Ticks run, graph replays, graph captures, most requests in flight at once
 * and the arena's high-water mark. */
int fkpm_batch_stats(fkpm_batch* b, long long* ticks, long long* replays, long long* captures,
                     int* peak_slots, size_t* peak_bytes);

#ifdef __cplusplus
}
#endif

#endif /* FASTKPM_HIP_H */
