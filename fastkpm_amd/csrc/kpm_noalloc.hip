// kpm_noalloc.hip -- fkpm_moments_noalloc: the Chebyshev moments of a
// Hamiltonian and a probing block that are already in device memory, with no
// device allocation at all.
//
// The work is that of fkpm_set_H + fkpm_set_R + fkpm_moments (the forward
// recursion alpha_{m+1} = 2 Hs alpha_m - alpha_{m-1} with both moment dots
// fused, engine_cpu.cpp:63-84), on the same step kernels.  Everything beyond
// the caller's inputs and outputs -- the scaled H values, the block-row index
// of each block, the packed probe block, the three recursion vectors, the
// per-workgroup moment partials and their per-step sums -- lives in one
// file-scope device array (`scratch`, kNoallocScratchBytes per device, held
// by every device the library runs on until the process ends).  Its device
// address is looked up once per device with hipGetSymbolAddress at the first
// call on that device; kernels receive that address, never the array's host
// name.  Nothing in the scratch needs to start at zero: every byte a call
// reads was written earlier in the same call.  Two calls that overlap on one
// device (two streams or two threads) share the scratch and race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "fastkpm_hip.h"
#include "kpm_dispatch.h"
#include "kpm_internal.h"
#include "kpm_kernels.h"

static constexpr size_t kNoallocScratchBytes = (size_t)FKPM_NOALLOC_SCRATCH_MB << 20;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device

int nerr(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return fkpm_internal::set_error(code, buf);
}

// The scratch's device address on `device` (current), looked up once.
int scratch_on(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return nerr(FKPM_E_HIP, "moments_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                  device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return 0;
}

// row_of_block[p] = block row of block p (one thread per block row).
__global__ void k_row_of_block(const int* __restrict__ rp, int nb, int* __restrict__ rob) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x)
    for (int p = rp[i]; p < rp[i + 1]; p++) rob[p] = i;
}

// sums[2 m + c] = sum_w partial[m][w][c] in a fixed order, one workgroup per
// step m.
__global__ __launch_bounds__(fkd::WG) void k_sum_partials(const double* __restrict__ partial,
                                                          int nwg, double* __restrict__ sums) {
  const double* p = partial + (size_t)blockIdx.x * nwg * 2;
  double a = 0.0, b = 0.0;
  for (int w = threadIdx.x; w < nwg; w += fkd::WG) {
    a += p[2 * w];
    b += p[2 * w + 1];
  }
  fkd::wg_sum2(a, b, sums + 2 * (size_t)blockIdx.x);
}

// mu from the per-step dots (the two-moments-per-product identity,
// engine_cpu.cpp:75-81): mu_0, mu_1 as summed; mu_2m = 2 s_2m - mu_0,
// mu_2m+1 = 2 s_2m+1 - mu_1.
__global__ void k_mu_from_sums(const double* __restrict__ sums, int nsteps,
                               double* __restrict__ mu) {
  const double m0 = sums[0], m1 = sums[1];
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nsteps; m += gridDim.x * blockDim.x) {
    mu[2 * m] = m == 0 ? m0 : 2.0 * sums[2 * m] - m0;
    mu[2 * m + 1] = m == 0 ? m1 : 2.0 * sums[2 * m + 1] - m1;
  }
}

template <class T>
struct TypeId {
  typedef T type;
};

// Launch a kernel through its host stub with exactly its parameter types.
template <typename... P>
int launch(void (*fp)(P...), dim3 grid, dim3 block, hipStream_t st,
           typename TypeId<P>::type... args) {
  void* argv[] = {(void*)&args...};
  const hipError_t e = hipLaunchKernel((const void*)fp, grid, block, argv, 0, st);
  if (e != hipSuccess) return nerr(FKPM_E_HIP, "moments_noalloc: launch: %s", hipGetErrorString(e));
  return 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The step kernels' geometry for a row of qp0 packets: the narrowest compiled
// pass width that covers it (else the widest), widest lane groups first.
bool pick_geometry(int b, int qp0, int* L, int* KQ, int* W) {
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(b, gs, 64);
  int w = 0, widest = 0;
  for (int k = 0; k < ng; k++) {
    const int x = gs[k].L * gs[k].KQ;
    widest = x > widest ? x : widest;
    if (x >= qp0 && (w == 0 || x < w)) w = x;
  }
  if (w == 0) w = widest;
  static const int pref[] = {64, 32, 16, 8, 4};
  for (int L0 : pref)
    for (int k = 0; k < ng; k++)
      if (gs[k].L == L0 && gs[k].L * gs[k].KQ == w) {
        *L = gs[k].L;
        *KQ = gs[k].KQ;
        *W = w;
        return true;
      }
  return false;
}

template <typename S>
int run(int device, unsigned char* base, int nb, int b, int nnzb, const int* rp, const int* ci,
        const double* val, double avg, double mag, int n, int s, const double* R, int M,
        double* mu, hipStream_t st) {
  const int ne = (int)(16 / sizeof(S)) / 2;  // complex elements per 16-byte packet
  const int qp0 = (s + ne - 1) / ne;
  int L = 0, KQ = 0, W = 0;
  if (!pick_geometry(b, qp0, &L, &KQ, &W))
    return nerr(FKPM_E_UNSUP, "moments_noalloc: no kernel geometry for block %d", b);
  fkd::KSet<S> ks;
  bool found = false;
  if constexpr (sizeof(S) == 8)
    found = fkd::find_kset_dc(b, L, KQ, &ks);
  else
    found = fkd::find_kset_fc(b, L, KQ, &ks);
  if (!found) return nerr(FKPM_E_UNSUP, "moments_noalloc: no kernels for block %d", b);
  const int qp = (qp0 + W - 1) / W * W;
  const int sp = qp * ne;
  // one front in segments of two passes, about one workgroup per 4 block
  // rows, at most 4096 (the engine's untuned default)
  const int gwg = (fkd::WG / fkd::WAVE) * (fkd::WAVE / L);
  int nwg = (nb + gwg - 1) / gwg;
  nwg = nwg > 4096 ? 4096 : (nwg < 1 ? 1 : nwg);
  nwg = (nwg + 15) / 16 * 16;
  const int chunk = -2;
  const int nsteps = M / 2;
  // scratch layout
  const size_t hv_b = (size_t)nnzb * b * b * 2 * sizeof(S);
  const size_t rob_b = (size_t)nnzb * sizeof(int);
  const size_t vec_b = (size_t)n * sp * 2 * sizeof(S);
  const size_t part_b = (size_t)nsteps * nwg * 2 * sizeof(double);
  const size_t sums_b = (size_t)nsteps * 2 * sizeof(double);
  size_t off = 0;
  const size_t o_hv = off;
  off += align256(hv_b);
  const size_t o_rob = off;
  off += align256(rob_b);
  size_t o_vec[4];
  for (int k = 0; k < 4; k++) {
    o_vec[k] = off;
    off += align256(vec_b);
  }
  const size_t o_part = off;
  off += align256(part_b);
  const size_t o_sums = off;
  off += align256(sums_b);
  if (off > kNoallocScratchBytes)
    return nerr(FKPM_E_ARG,
                "moments_noalloc: the problem needs %zu bytes of working memory, more than the "
                "%zu-byte static scratch (nb %d, b %d, nnzb %d, s %d, M %d)",
                off, kNoallocScratchBytes, nb, b, nnzb, s, M);
  S* hv = reinterpret_cast<S*>(base + o_hv);
  int* rob = reinterpret_cast<int*>(base + o_rob);
  S* vec[4];
  for (int k = 0; k < 4; k++) vec[k] = reinterpret_cast<S*>(base + o_vec[k]);
  double* part = reinterpret_cast<double*>(base + o_part);
  double* sums = reinterpret_cast<double*>(base + o_sums);
  const dim3 blk(fkd::WG);
  // Hs = (H - avg I) / mag in the storage type; R packed n x sp, zero padded
  if (launch(&k_row_of_block, dim3((nb + 255) / 256 < 4096 ? (nb + 255) / 256 : 4096), dim3(256),
             st, rp, nb, rob))
    return FKPM_E_HIP;
  const size_t cnt = (size_t)nnzb * b * b;
  const int sgrid = (int)((cnt + 255) / 256 < 4096 ? (cnt + 255) / 256 : 4096);
  if (launch(&fkd::k_scale_H<double, S, true>, dim3(sgrid > 0 ? sgrid : 1), dim3(256), st, val, hv,
             (const int*)rob, ci, nnzb, b, mag, avg / mag))
    return FKPM_E_HIP;
  if (launch(&fkd::k_pack_R<double, S, true>, dim3((n + 63) / 64, (sp + 15) / 16), blk, st, R,
             vec[0], n, s, sp))
    return FKPM_E_HIP;
  // alpha_0 = R in vec[0]; alpha_k (k >= 1) in vec[1 + (k - 1) % 3]
  auto alpha = [&](int k) { return k == 0 ? vec[0] : vec[1 + (k - 1) % 3]; };
  fkd::RSource<S> nors{nullptr, nullptr, nullptr, 0.0};
  for (int m = 0; m < nsteps; m++) {
    const S* x = alpha(m);
    const S* z = m == 0 ? nullptr : alpha(m - 1);
    S* y = alpha(m + 1);
    if (launch(ks.mom, dim3(nwg), blk, st, (const int*)rp, (const int*)ci, (const S*)hv, x, z, y,
               nb, qp, chunk, m == 0 ? 1.0 : 2.0, m == 0 ? 0.0 : -1.0,
               part + (size_t)m * nwg * 2, 0, nors))
      return FKPM_E_HIP;
  }
  if (launch(&k_sum_partials, dim3(nsteps), blk, st, (const double*)part, nwg, sums))
    return FKPM_E_HIP;
  if (launch(&k_mu_from_sums, dim3((nsteps + 255) / 256), dim3(256), st, (const double*)sums,
             nsteps, mu))
    return FKPM_E_HIP;
  (void)device;
  return 0;
}

}  // namespace

extern "C" {

int fkpm_moments_noalloc(int device, int prec, int nb, int b, int nnzb, const int* row_ptr,
                         const int* col_idx, const void* val, double avg, double mag, int n,
                         int s, const void* R_colmajor, int M, double* mu, void* stream) {
  if (!row_ptr || !col_idx || !val || !R_colmajor || !mu)
    return nerr(FKPM_E_ARG, "moments_noalloc: null argument");
  if (prec != FKPM_PREC_FP64 && prec != FKPM_PREC_FP32)
    return nerr(FKPM_E_ARG, "moments_noalloc: bad precision %d", prec);
  if (!fkd::block_supported(b)) return nerr(FKPM_E_UNSUP, "moments_noalloc: block size %d", b);
  if (nb <= 0 || nnzb < nb || n != nb * b || s <= 0)
    return nerr(FKPM_E_ARG, "moments_noalloc: bad sizes (nb %d, b %d, nnzb %d, n %d, s %d)", nb,
                b, nnzb, n, s);
  if (M < 2 || M % 2 != 0) return nerr(FKPM_E_ARG, "moments_noalloc: M = %d must be even", M);
  if (!(mag > 0) || !std::isfinite(avg))
    return nerr(FKPM_E_ARG, "moments_noalloc: bad energy scale avg %g mag %g", avg, mag);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev ||
      device >= kMaxDevices)
    return nerr(FKPM_E_NODEV, "moments_noalloc: invalid device #%d", device);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return nerr(FKPM_E_HIP, "moments_noalloc: hipSetDevice: %s", hipGetErrorString(e));
  unsigned char* base = nullptr;
  if (int rc = scratch_on(device, &base)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int* rp = row_ptr;
  const int* ci = col_idx;
  const double* hv = static_cast<const double*>(val);
  const double* R = static_cast<const double*>(R_colmajor);
  return prec == FKPM_PREC_FP32
             ? run<float>(device, base, nb, b, nnzb, rp, ci, hv, avg, mag, n, s, R, M, mu, st)
             : run<double>(device, base, nb, b, nnzb, rp, ci, hv, avg, mag, n, s, R, M, mu, st);
}

}  // extern "C"
