// kpm_dispatch.h -- host-side table of kernel instantiations.
//
// One KSet holds the host stubs of every kernel for one (storage type S,
// real/complex, block size B, lanes per group L, packets per lane KQ).  The
// instantiations live in four translation units (kpm_inst_{d,f}{c,r}.hip) that
// compile in parallel; kpm_capi.hip looks sets up by runtime parameters.
#pragma once
#include <hip/hip_runtime.h>

namespace fkd {

template <typename S>
struct RSource;

template <class S>
struct KSet {
  void (*mom)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int, double,
              double, double*, int, RSource<S>) = nullptr;
  void (*xi)(const int*, const int*, const S*, const S*, const S*, S*, double*, int, int, int,
             double, double, double, double, int, int) = nullptr;
  void (*ad_fused)(const int*, const int*, const S*, const S*, const S*, const S*, const S*, S*,
                   S*, double*, int, int, int, double, double, double, double, double,
                   double, int) = nullptr;
  void (*ad_plain)(const int*, const int*, const S*, const S*, const S*, const S*, const S*, S*,
                   S*, double*, int, int, int, double, double, double, double, double,
                   double, int) = nullptr;
  // backward steps reading alpha_{m-1} from the forward table (D fused on H's
  // pattern, or not)
  void (*ad_stored)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                    const S*, S*, double*, int, int, int, double, double, double, double, double,
                    double, int) = nullptr;
  void (*ad_stored_plain)(const int*, const int*, const S*, const S*, const S*, const S*,
                          const S*, const S*, S*, double*, int, int, int, double, double, double,
                          double, double, double, int) = nullptr;
  void (*outer_ss)(const int*, const int*, const S*, const S*, double*, int, int, int,
                   double) = nullptr;
  void (*outer_sd)(const int*, const int*, const S*, const double*, double*, int, int, int,
                   double) = nullptr;
  void (*outer_ds)(const int*, const int*, const double*, const S*, double*, int, int, int,
                   double) = nullptr;
  // matrix-core kernels (complex 8x8 blocks only; null otherwise), same
  // signatures as mom / ad_fused / ad_plain
  void (*mfma_mom)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int, double,
                   double, double*, int, RSource<S>) = nullptr;
  void (*mfma_ad_fused)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                        S*, S*, double*, int, int, int, double, double, double, double, double,
                        double, int) = nullptr;
  void (*mfma_ad_plain)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                        S*, S*, double*, int, int, int, double, double, double, double, double,
                        double, int) = nullptr;
  // row-resident matrix-core variants (rows of at most 8 blocks, vectors
  // under 4 GiB): forward step, fused autodiff step
  void (*mfma_mom_rr)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int,
                      double, double, double*, int, RSource<S>) = nullptr;
  void (*mfma_ad_rr)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                     S*, S*, double*, int, int, int, double, double, double, double, double,
                     double, int) = nullptr;
  // the same for rows of at most 7 blocks (no padding slot)
  void (*mfma_mom_rr7)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int,
                       double, double, double*, int, RSource<S>) = nullptr;
  void (*mfma_ad_rr7)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                      S*, S*, double*, int, int, int, double, double, double, double, double,
                      double, int) = nullptr;
  // FP32 only: the row-resident forward and autodiff steps on pairs of column
  // tiles (the padded width a multiple of 16 packets), 8 and 7 block slots
  void (*mfma_mom_rr2)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int,
                       double, double, double*, int, RSource<S>) = nullptr;
  void (*mfma_mom_rr2_7)(const int*, const int*, const S*, const S*, const S*, S*, int, int, int,
                         double, double, double*, int, RSource<S>) = nullptr;
  void (*mfma_ad_rr2)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                      S*, S*, double*, int, int, int, double, double, double, double, double,
                      double, int) = nullptr;
  void (*mfma_ad_rr2_7)(const int*, const int*, const S*, const S*, const S*, const S*, const S*,
                        S*, S*, double*, int, int, int, double, double, double, double, double,
                        double, int) = nullptr;
  int B = 0, L = 0, KQ = 0;
};

// Geometries compiled for each block size (lanes per group, packets per lane).
struct GeomChoice {
  int L, KQ;
};
// Block sizes with compiled kernels.
inline bool block_supported(int b) { return b == 1 || b == 2 || b == 4 || b == 8; }
// List of (L, KQ) compiled for block size b; returns the count.
int geometries_for_block(int b, GeomChoice* out, int cap);

bool find_kset_dc(int B, int L, int KQ, KSet<double>* out);  // double, complex
bool find_kset_dr(int B, int L, int KQ, KSet<double>* out);  // double, real
bool find_kset_fc(int B, int L, int KQ, KSet<float>* out);   // float, complex
bool find_kset_fr(int B, int L, int KQ, KSet<float>* out);   // float, real

}  // namespace fkd

// The list of compiled (B, L, KQ) geometries.  B = 1, 2 take up to 4 packets
// per lane; B = 4 up to 2; B = 8 one (its accumulators are 8x larger).  One
// packet per lane keeps the register count (and so the occupancy) of the
// fused autodiff step low at a given pass width; four packets over eight
// lanes gives it longer loads per lane (the FP32 autodiff step at S = 64).
// For 2x2 blocks, two packets over a whole wave carry 128 FP64 columns in one
// pass with the row wave-uniform (scalar row offsets, indices and H) at 3
// waves per SIMD in the autodiff step, where 32 lanes x 4 packets hold 2 (the
// tuner takes it for C3's FP64 forward step, with the lattice row order).
#define FKD_FOR_EACH_GEOMETRY(X)                                                       \
  X(1, 4, 1) X(1, 4, 2) X(1, 8, 2) X(1, 8, 4) X(1, 16, 1) X(1, 16, 2) X(1, 16, 4)      \
  X(1, 32, 1) X(1, 32, 4) X(1, 64, 1) X(1, 64, 4)                                      \
  X(2, 4, 1) X(2, 4, 2) X(2, 8, 2) X(2, 8, 4) X(2, 16, 1) X(2, 16, 2) X(2, 16, 4)      \
  X(2, 32, 1) X(2, 32, 4) X(2, 64, 1) X(2, 64, 2) X(2, 64, 4)                          \
  X(4, 4, 1) X(4, 4, 2) X(4, 8, 2) X(4, 16, 1) X(4, 16, 2) X(4, 32, 1) X(4, 32, 2)     \
  X(4, 64, 1) X(4, 64, 2)                                                              \
  X(8, 4, 1) X(8, 8, 1) X(8, 16, 1) X(8, 32, 1) X(8, 64, 1)
