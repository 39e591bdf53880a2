// kpm_internal.h -- host functions shared between the translation units of
// libfastkpm_hip.so (not exported).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace fkpm_internal {

// Hs = (H - shift I) / mag from complex-double host-layout values `raw` (device)
// into `dst` (FP32 or FP64 interleaved complex); row_of_block[p] = block row of p.
int scale_H_cx(hipStream_t st, const double* raw, void* dst, bool single, const int* row_of_block,
               const int* ci, int nnzb, int b, double mag, double shift);

// R (n x s column-major complex double, device) -> n x sp row-major packed
// (FP32 or FP64), padded columns zero.
int pack_R_cx(hipStream_t st, const double* raw, void* dst, bool single, int n, int s, int sp);

// Record an error message for fkpm_last_error() and return `code`.
int set_error(int code, const char* msg);

// The lattice row order the single-request engine builds for a periodic 3-D
// lattice pattern in natural order (FKPM_OPT_ROW_ORDER): the block row at each
// sweep position, XCD k's slab in the k-th eighth; false for other patterns.
bool lattice_row_order(int nb, const int* rp, const int* ci, std::vector<int>& order);

}  // namespace fkpm_internal
