// kpm_batch.hip -- batched multi-Hamiltonian engine: the serving runtime.
//
// Many KPM requests (each: a Hamiltonian, probing vectors R, M Chebyshev
// moments and coefficients c) share the GPU.  A request admitted to a slot
// runs, entirely on the device, the pipeline of one Engine solve:
//   moments(M)                     ticks 0 .. M/2-1    (forward steps)
//   D = diag R R^+                 tick  M/2
//   autodiff backward steps        ticks M/2+1 .. M-1
//   D += alpha_0 beta_1^+          tick  M            (M > 2)
//   symmetrize + scale(1/mag),
//   reduction of the moment dots   tick  M+1
// (Engine_CPU::moments / autodiff_matrix, engine_cpu.cpp:63-84, 255-304).
// One tick advances every active slot by one step: per class of slots
// (precision, block size, column pass width) one forward, one backward and
// one auxiliary launch cover all slots of the class, each workgroup decoding
// its slot from a prefix table; slots in another phase exit at once.  A run
// of ticks is captured once in a hipGraph and replayed while the slot set is
// unchanged.  Between replays a scheduler thread retires finished requests
// (moments and D to the host) and admits queued ones while their HBM
// footprint fits the arena and a slot is free (continuous batching).
//
// Each slot's kernels see only that slot's buffers with a workgroup count
// fixed by the request's shape, so a request's results are bitwise the same
// alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fastkpm_hip.h"
#include "kpm_internal.h"
#include "kpm_kernels.h"
#include "kpm_mfma.h"

namespace fkd {

// Device descriptor of one admitted request.
struct SlotDev {
  const int* rp;
  const int* ci;
  const void* hv;
  const int* tpos;
  const void* R;
  void* A[3];
  void* Bv[3];
  double* dacc;
  double* dout;
  double* partial;  // [M/2][nwg][2] per-workgroup moment dots
  double* psum;     // [M/2][2]
  const double* c;  // M coefficients
  const int* rord;  // lattice row order: {row, first block, end block} per sweep position
  int nb, qp, chunk, nwg, M, nt;  // nt: NT_ORDER when the step kernels sweep rord
  double inv_mag, diag, sb;
};

// Slot of workgroup `bid`: the last s with wg_off[s] <= bid.
__device__ __forceinline__ int batch_find(const int* __restrict__ wg_off, int ns, int bid) {
  int lo = 0, hi = ns;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (wg_off[mid] <= bid) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename S>
__device__ __forceinline__ S* vec_a(const SlotDev& d, int k) {
  return k == 0 ? (S*)d.R : (S*)d.A[k % 3];
}

// Forward Chebyshev step of every slot of the class in its forward phase.
// MF: 0 vector units, 1 matrix cores, 2 matrix cores row-resident (rows of
// at most MF_DM blocks, vectors below 4 GiB), 3 the same for rows of at most
// 7 blocks (no padding slot).
template <typename S, int B, int L, int KQ, int MF>
__global__ __launch_bounds__(WG) void k_batch_fwd(const SlotDev* __restrict__ sd,
                                                  const int* __restrict__ st,
                                                  const int* __restrict__ slots,
                                                  const int* __restrict__ wg_off, int ns) {
  const int s = batch_find(wg_off, ns, (int)blockIdx.x);
  const int slot = slots[s];
  const SlotDev& d = sd[slot];
  const int step = st[slot];
  const int h = d.M / 2;
  if (step >= h) return;
  const WgId wid{(int)blockIdx.x - wg_off[s], d.nwg};
  const S* x = vec_a<S>(d, step);
  const S* z = step == 0 ? nullptr : vec_a<S>(d, step - 1);
  S* y = vec_a<S>(d, step + 1);
  const double cx = step == 0 ? 1.0 : 2.0, cz = step == 0 ? 0.0 : -1.0;
  double* part = d.partial + (size_t)step * d.nwg * 2;
  RSource<S> nors{nullptr, nullptr, nullptr, 0.0};
  if constexpr (MF >= 2)
    mfma_step_rr_dev<S, MF == 3 ? 7 : MF_DM, 4>(wid, d.rp, d.ci, (const S*)d.hv, x, z, y, d.nb,
                                                d.qp, d.chunk, cx, cz, part, 0, nors);
  else if constexpr (MF == 1)
    mfma_step_dev<S>(wid, d.rp, d.ci, (const S*)d.hv, x, z, y, d.nb, d.qp, d.chunk, cx, cz, part,
                     0, nors);
  else
    step_moments_dev<S, true, B, L, KQ>(wid, (d.nt & NT_ORDER) ? d.rord : d.rp, d.ci,
                                        (const S*)d.hv, x, z, y, d.nb, d.qp, d.chunk, cx, cz, part,
                                        d.nt, nors);
}

// Backward (autodiff) step of every slot of the class in its backward phase;
// the sequence of autodiff_t in the single-request engine.  The row-resident
// autodiff step is used in FP64 only (in FP32 the general one is faster).
template <typename S, int B, int L, int KQ, int MF>
__global__ __launch_bounds__(WG) void k_batch_bwd(const SlotDev* __restrict__ sd,
                                                  const int* __restrict__ st,
                                                  const int* __restrict__ slots,
                                                  const int* __restrict__ wg_off, int ns) {
  const int s = batch_find(wg_off, ns, (int)blockIdx.x);
  const int slot = slots[s];
  const SlotDev& d = sd[slot];
  const int step = st[slot];
  const int h = d.M / 2;
  if (step <= h || step >= d.M) return;
  const int m = h - 1 - (step - h - 1);  // m = M/2-1 .. 1
  const WgId wid{(int)blockIdx.x - wg_off[s], d.nwg};
  const S* am = (const S*)d.A[m % 3];
  const S* am1 = (const S*)d.A[(m + 1) % 3];
  S* aout = (S*)d.A[(m - 1) % 3];
  const bool first = (m == h - 1), second = (m == h - 2);
  const S* bm1 = first ? nullptr : (const S*)d.Bv[(m + 1) % 3];
  const S* bm2 = (first || second) ? nullptr : (const S*)d.Bv[(m + 2) % 3];
  const double sb2 = second ? d.sb : 0.0;
  S* bout = (S*)d.Bv[m % 3];
  const double c4 = 4 * d.c[2 * m], c2 = 2 * d.c[2 * m + 1];
  const double w = m > 1 ? 2 * d.c[2 * m - 1] : 0.0;
  if constexpr (MF >= 2 && sizeof(S) == 8)
    mfma_autodiff_rr_dev<S, true, MF == 3 ? 7 : MF_DM, mf_rr_group<S, MF == 3 ? 7 : MF_DM>()>(
        wid, d.rp, d.ci, (const S*)d.hv, am, am1, bm1, bm2, aout, bout, d.dacc, d.nb, d.qp,
        d.chunk, c4, c2, w, d.sb, sb2, 2.0, 0);
  else if constexpr (MF >= 1)
    mfma_autodiff_dev<S, true>(wid, d.rp, d.ci, (const S*)d.hv, am, am1, bm1, bm2, aout, bout,
                               d.dacc, d.nb, d.qp, d.chunk, c4, c2, w, d.sb, sb2, 2.0, 0);
  else
    step_autodiff_dev<S, true, B, L, KQ, true>(wid, (d.nt & NT_ORDER) ? d.rord : d.rp, d.ci,
                                               (const S*)d.hv, am, am1, bm1, bm2, aout, bout,
                                               d.dacc, d.nb, d.qp, d.chunk, c4, c2, w, d.sb, sb2,
                                               2.0, d.nt);
}

// Auxiliary work of a tick: zero D (tick 0), D = diag R R^+ (tick M/2),
// D += alpha_0 beta_1^+ (tick M), symmetrize/scale D and reduce the moment
// dots (tick M+1).
template <typename S, int B, int L, int KQ, int MF>
__global__ __launch_bounds__(WG) void k_batch_aux(const SlotDev* __restrict__ sd,
                                                  const int* __restrict__ st,
                                                  const int* __restrict__ slots,
                                                  const int* __restrict__ wg_off, int ns) {
  const int s = batch_find(wg_off, ns, (int)blockIdx.x);
  const int slot = slots[s];
  const SlotDev& d = sd[slot];
  const int step = st[slot];
  const int h = d.M / 2;
  const WgId wid{(int)blockIdx.x - wg_off[s], d.nwg};
  const int* rp = d.rp;
  const size_t nnzb = (size_t)rp[d.nb];
  const size_t delem = nnzb * B * B;  // complex elements of D
  const size_t t0 = (size_t)wid.k * WG + threadIdx.x, tstride = (size_t)wid.n * WG;
  if (step == 0) {
    for (size_t e = t0; e < 2 * delem; e += tstride) d.dacc[e] = 0.0;
  } else if (step == h) {
    outer_dev<S, S, S, true, B, L, KQ>(wid, rp, d.ci, (const S*)d.R, (const S*)d.R, d.dacc, d.nb,
                                       d.qp, d.chunk, d.diag);
  } else if (step == d.M && d.M > 2) {
    outer_dev<S, S, S, true, B, L, KQ>(wid, rp, d.ci, (const S*)d.A[0], (const S*)d.Bv[1], d.dacc,
                                       d.nb, d.qp, d.chunk, 1.0);
  } else if (step == d.M + 1) {
    // D_out = ((D + D^+) / 2) / mag on the block pattern (symmetrize + scale)
    for (size_t e = t0; e < delem; e += tstride) {
      const size_t k = e / (B * B);
      const int r = (int)(e % (B * B));
      const int bj = r / B, bi = r % B;
      const int tp = d.tpos[k];
      if (tp < 0) {
        d.dout[2 * e] = d.dacc[2 * e] * d.inv_mag;
        d.dout[2 * e + 1] = d.dacc[2 * e + 1] * d.inv_mag;
      } else {
        const size_t e2 = (size_t)tp * B * B + (size_t)bi * B + bj;
        d.dout[2 * e] = (d.dacc[2 * e] + d.dacc[2 * e2]) / 2.0 * d.inv_mag;
        d.dout[2 * e + 1] = (d.dacc[2 * e + 1] - d.dacc[2 * e2 + 1]) / 2.0 * d.inv_mag;
      }
    }
    // moment dots of forward step m: fixed-order sum over the nwg partials
    for (int m = wid.k; m < h; m += wid.n) {
      const double* p = d.partial + (size_t)m * d.nwg * 2;
      double a = 0.0, b = 0.0;
      for (int w = threadIdx.x; w < d.nwg; w += WG) {
        a += p[2 * w];
        b += p[2 * w + 1];
      }
      wg_sum2(a, b, d.psum + 2 * (size_t)m);
    }
  }
}

// Probing vectors from their compact form (one nonzero per orbital row:
// column rcol[i] (-1 for none), value rval[i]) into the packed n x sp block.
template <typename S>
__global__ __launch_bounds__(WG) void k_batch_scatter_R(const int* __restrict__ rcol,
                                                        const double* __restrict__ rval, int n,
                                                        int sp, S* __restrict__ R) {
  const size_t total = (size_t)n * sp;
  for (size_t e = (size_t)blockIdx.x * WG + threadIdx.x; e < total; e += (size_t)gridDim.x * WG) {
    const int i = (int)(e / sp), j = (int)(e % sp);
    const bool hit = rcol[i] == j;
    R[2 * e] = hit ? (S)rval[2 * (size_t)i] : S(0);
    R[2 * e + 1] = hit ? (S)rval[2 * (size_t)i + 1] : S(0);
  }
}

// Probing vectors generated on the device (Engine::set_R_correlated,
// src/engine.cpp:76-89): column j's generator is std::mt19937(seed_j), seed_j
// the j-th draw of the caller's master engine, and the rows of colour j draw
// their phases from it in row order through
// std::uniform_int_distribution<int>(0, 3), which libstdc++ (Lemire's
// downscaling, no rejection for a range of 4) reduces to the top two bits of
// the 32-bit draw: phase = (+1, +i, -1, -i)[y >> 30].
//
// One workgroup of one wave per column.  The wave walks the rows in chunks of
// 64; the rows of colour j in a chunk take consecutive outputs of the
// generator (their rank among the chunk's hits).  The generator state (624
// words) sits in LDS and is regenerated ("twisted") by the whole wave in four
// phases whose reads never see a word of the wrong generation:
//   k in [0, 227):   uses old mt[k], mt[k+1], mt[k+397]
//   k in [227, 454): uses new mt[k-227], old mt[k], mt[k+1]
//   k in [454, 623): same, mt[k-227] from the previous phase
//   k = 623:         uses new mt[396], new mt[0], old mt[623]
// Every row has exactly one colour in [0, s), so every rval is written once.
__global__ __launch_bounds__(64) void k_probe_phases(const int* __restrict__ groups, int n,
                                                     const unsigned* __restrict__ seeds,
                                                     double* __restrict__ rval) {
  __shared__ unsigned mt[624];
  const int j = blockIdx.x, lane = threadIdx.x;
  if (lane == 0) {
    unsigned x = seeds[j];
    mt[0] = x;
    for (int i = 1; i < 624; i++) {
      x = 1812433253u * (x ^ (x >> 30)) + (unsigned)i;
      mt[i] = x;
    }
  }
  __syncthreads();
  auto twist_range = [&](int k0, int k1) {
    for (int kb = k0; kb < k1; kb += 64) {
      const int k = kb + lane;
      unsigned nv = 0;
      if (k < k1) {
        const unsigned y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
        nv = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      __syncthreads();
      if (k < k1) mt[k] = nv;
      __syncthreads();
    }
  };
  auto twist = [&]() {
    twist_range(0, 227);
    twist_range(227, 454);
    twist_range(454, 623);
    twist_range(623, 624);
  };
  int pos = 624;  // next unused output of the current block (624: none left)
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    const bool hit = i < n && groups[i] == j;
    const unsigned long long mask = __ballot(hit);
    const int cnt = __popcll(mask);
    if (cnt == 0) continue;  // the same for the whole wave
    const int idx = pos + __popcll(mask & ((1ull << lane) - 1ull));
    unsigned y = 0;
    if (hit && idx < 624) y = mt[idx];
    if (pos + cnt > 624) {
      __syncthreads();
      twist();
      if (hit && idx >= 624) y = mt[idx - 624];
      pos += cnt - 624;
    } else {
      pos += cnt;
    }
    if (hit) {
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      const unsigned ph = y >> 30;
      rval[2 * (size_t)i] = ph == 0 ? 1.0 : (ph == 2 ? -1.0 : 0.0);
      rval[2 * (size_t)i + 1] = ph == 1 ? 1.0 : (ph == 3 ? -1.0 : 0.0);
    }
  }
}

// End of a tick: every active slot moves to its next step.
__global__ void k_batch_advance(int* __restrict__ st, const int* __restrict__ act, int na) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < na) st[act[t]] += 1;
}

}  // namespace fkd

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;
using fkd::SlotDev;
typedef void (*BatchKernel)(const SlotDev*, const int*, const int*, const int*, int);

// Kernel classes: (storage type, block size, lanes, packets per lane, MFMA).
// W = pass width in 16-byte packets (the padding granularity of S).
struct ClassDef {
  int B, L, KQ, mf, W;
  BatchKernel fwd, bwd, aux;
};

template <typename S, int B, int L, int KQ, int MF>
ClassDef mk_class(int W) {
  return ClassDef{B, L, KQ, MF, W, &fkd::k_batch_fwd<S, B, L, KQ, MF>,
                  &fkd::k_batch_bwd<S, B, L, KQ, MF>, &fkd::k_batch_aux<S, B, L, KQ, MF>};
}

// Complex 2x2 blocks: pass widths 8, 16, 32, 64 packets (a packet holds one
// FP64 or two FP32 complex columns; the narrowest class keeps a 16-column
// FP32 request from being padded to 32); complex 8x8 blocks: matrix cores
// (tiles of 16 columns) with the outer products at width 16 (FP64) or 8
// (FP32) packets: general, row-resident on 8 block slots (requests whose rows
// all have at most MF_DM blocks) and on 7 (at most 7 blocks: 3-D lattices).
const std::vector<ClassDef>& classes(bool single) {
  static const std::vector<ClassDef> d = {
      mk_class<double, 2, 4, 2, 0>(8),    mk_class<double, 2, 8, 2, 0>(16),
      mk_class<double, 2, 16, 2, 0>(32),  mk_class<double, 2, 64, 1, 0>(64),
      mk_class<double, 8, 16, 1, 1>(16),  mk_class<double, 8, 16, 1, 2>(16),
      mk_class<double, 8, 16, 1, 3>(16)};
  static const std::vector<ClassDef> f = {
      mk_class<float, 2, 4, 2, 0>(8),    mk_class<float, 2, 8, 2, 0>(16),
      mk_class<float, 2, 16, 2, 0>(32),  mk_class<float, 2, 64, 1, 0>(64),
      mk_class<float, 8, 8, 1, 1>(8),    mk_class<float, 8, 8, 1, 2>(8),
      mk_class<float, 8, 8, 1, 3>(8)};
  return single ? f : d;
}

// First-fit arena over one device allocation; offsets aligned to 256 bytes.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0, peak = 0;
  std::map<size_t, size_t> holes;  // offset -> length
  void init(char* b, size_t n) {
    base = b;
    size = n;
    holes.clear();
    holes[0] = n;
  }
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t)255;
    if (n == 0) n = 256;
    for (auto it = holes.begin(); it != holes.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, o = it->first + n;
        holes.erase(it);
        if (rest) holes[o] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    if (n == 0) n = 256;
    used -= n;
    auto it = holes.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != holes.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      holes.erase(nx);
    }
    if (it != holes.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        holes.erase(it);
      }
    }
  }
};

struct Request {
  long long id = 0;
  int nb = 0, b = 0, nnzb = 0, n = 0, s = 0, M = 0;
  double lo = 0, hi = 0;
  std::vector<int> rp, ci, rowblk, tpos;
  std::vector<int> rdesc;  // lattice row order descriptors (large 3-D lattices), else empty
  int max_deg = 0;  // most blocks in a row
  std::vector<double> val, R, c;
  bool compact = false;  // R given as (rcol, rval) per orbital row (R holds rval)
  bool seeded = false;   // compact R generated on the device: rcol = colours
  std::vector<int> rcol;
  std::vector<unsigned> seeds;  // seeded: one generator seed per column
  // results
  std::vector<double> mu, D;
  double energy = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, < 0 error code
  std::string err;
  Clock::time_point t_submit, t_admit, t_done;
};

// Per-request layout of the slot's buffers in the arena.
struct Layout {
  int cls = 0, W = 0, qp = 0, sp = 0, nwg = 0, chunk = 0;
  size_t off[24];
  size_t len[24];
  int nbuf = 0;
  size_t total = 0;
};

enum Buf { B_RP, B_CI, B_RB, B_TP, B_HV, B_HRAW, B_R, B_A0, B_A1, B_A2, B_B0, B_B1, B_B2, B_RRAW,
           B_DACC, B_DOUT, B_PART, B_PSUM, B_C, B_RD, B_COUNT };

// mf: the matrix-core form a b = 8 request takes (1 general, 2 row-resident on
// 8 slots, 3 on 7); vector units (0) for b = 2.
int pick_class(bool single, int b, int s, int mf, int* W) {
  const auto& cl = classes(single);
  const int ne = single ? 2 : 1;  // complex elements per 16-byte packet
  const int need = (s + ne - 1) / ne;
  const int want = b == 8 ? mf : 0;
  int best = -1;
  for (int k = 0; k < (int)cl.size(); k++) {
    if (cl[k].B != b || cl[k].mf != want) continue;
    if (best < 0) {
      best = k;
      continue;
    }
    // smallest width covering the row, else the widest
    const bool cov_k = cl[k].W >= need, cov_b = cl[best].W >= need;
    if ((cov_k && (!cov_b || cl[k].W < cl[best].W)) || (!cov_k && !cov_b && cl[k].W > cl[best].W))
      best = k;
  }
  if (best >= 0) *W = cl[best].W;
  return best;
}

void plan(bool single, const Request& r, Layout* L) {
  const size_t es = single ? 4 : 8;
  const int ne = single ? 2 : 1;
  const int need = (r.s + ne - 1) / ne;
  // row-resident matrix-core kernels: rows of at most MF_DM blocks and
  // vector byte offsets within 32 bits (at the class's pass width of 16 / 8
  // packets)
  const int wb8 = single ? 8 : 16;
  const size_t vb8 = (size_t)r.n * ((need + wb8 - 1) / wb8) * wb8 * ne * 2 * es;
  const bool rr = r.b == 8 && r.max_deg <= fkd::MF_DM && vb8 < ((size_t)1 << 32);
  L->cls = pick_class(single, r.b, r.s, !rr ? 1 : r.max_deg <= 7 ? 3 : 2, &L->W);
  if (L->cls < 0) {
    L->total = 0;
    return;
  }
  L->qp = ((need + L->W - 1) / L->W) * L->W;
  L->sp = L->qp * ne;
  // workgroups per slot: one per 4 block rows up to 2048 (fixed by the
  // shape); they sweep the slot's rows as one front (chunk -1), or, with the
  // lattice row order, the sweep positions one range per XCD (chunk > 0).  A
  // slot's first workgroup sits at any offset of the launch, so workgroup k
  // of the slot runs on XCD (offset + k) mod 8: the ranges keep to one XCD
  // each, only which XCD takes which range shifts.
  int nwg = std::min(std::max((r.nb + 3) / 4, 1), 2048);
  int chunk = (r.nb + nwg - 1) / nwg;
  if (!r.rdesc.empty()) {
    const int gwg = (fkd::WG / fkd::WAVE) * (fkd::WAVE / classes(single)[L->cls].L);
    chunk = (chunk + gwg - 1) / gwg * gwg;
  }
  nwg = (r.nb + chunk - 1) / chunk;
  L->nwg = nwg;
  L->chunk = r.rdesc.empty() ? -1 : chunk;
  const size_t vec = (size_t)r.n * L->sp * 2 * es;
  const size_t dbl = (size_t)r.nnzb * r.b * r.b * 2 * sizeof(double);
  const size_t h = r.M / 2;
  size_t len[B_COUNT];
  len[B_RP] = sizeof(int) * (r.nb + 1);
  len[B_CI] = sizeof(int) * r.nnzb;
  len[B_RB] = sizeof(int) * r.nnzb;
  len[B_TP] = sizeof(int) * r.nnzb;
  len[B_HV] = (size_t)r.nnzb * r.b * r.b * 2 * es;
  len[B_HRAW] = dbl;
  len[B_R] = len[B_A0] = len[B_A1] = len[B_A2] = len[B_B0] = len[B_B1] = len[B_B2] = vec;
  len[B_RRAW] = r.compact ? (size_t)r.n * (sizeof(int) + 2 * sizeof(double)) + 512 +
                                (r.seeded ? sizeof(unsigned) * r.seeds.size() : 0)
                         : (size_t)r.n * r.s * 2 * sizeof(double);
  len[B_DACC] = len[B_DOUT] = dbl;
  len[B_PART] = h * nwg * 2 * sizeof(double);
  len[B_PSUM] = h * 2 * sizeof(double);
  len[B_C] = sizeof(double) * r.M;
  len[B_RD] = sizeof(int) * r.rdesc.size();
  L->nbuf = B_COUNT;
  L->total = 0;
  for (int k = 0; k < B_COUNT; k++) {
    L->len[k] = len[k];
    L->total += (len[k] + 255) & ~(size_t)255;
  }
}

struct Slot {
  std::shared_ptr<Request> rq;
  Layout lay;
  int idx = -1;
  bool staging_live = true;
  SlotDev dev{};
};

}  // namespace

struct fkpm_batch {
  int device = 0;
  bool single = false;
  int max_slots = 0;
  hipStream_t stream = nullptr;
  char* arena_mem = nullptr;
  Arena arena;
  // device tables
  SlotDev* d_slots = nullptr;
  int* d_st = nullptr;
  int* d_act = nullptr;
  std::vector<int*> d_cls_slots, d_cls_off;
  // host mirrors
  std::vector<SlotDev> h_slots;
  std::vector<int> h_act, h_st;
  std::vector<std::vector<int>> h_cls_slots, h_cls_off;
  std::vector<int> free_idx;
  // scheduling
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<long long, std::shared_ptr<Request>> reqs;
  std::vector<Slot> active;
  long long next_id = 1;
  bool stop = false;
  int waiters = 0;  // threads inside fkpm_batch_wait (close waits until none is left)
  std::condition_variable cv_idle;
  std::string broken;  // sticky failure of the scheduler
  std::thread worker;
  // options and statistics
  bool use_graph = true;
  int max_ticks = 16;
  long long n_ticks = 0, n_replays = 0, n_captures = 0;
  int peak_slots = 0;
  // graph cache for the current slot set
  long long set_version = 0, graph_version = -1;
  std::map<int, hipGraphExec_t> graphs;  // by tick count

  void clear_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
};

namespace {


#define BCHK(x)                                                                       \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      char m_[256];                                                                   \
      snprintf(m_, sizeof(m_), "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),    \
               __FILE__, __LINE__);                                                   \
      throw std::runtime_error(m_);                                                   \
    }                                                                                 \
  } while (0)

// Host-to-device copy of the first `bytes` of a request's host array; a byte
// count beyond the array is a planning error and is refused before the copy.
template <class T>
void h2d(void* dst, const std::vector<T>& src, size_t bytes, hipStream_t st) {
  if (bytes > src.size() * sizeof(T)) {
    char m[160];
    snprintf(m, sizeof(m), "admission copy of %zu bytes from a host array of %zu bytes", bytes,
             src.size() * sizeof(T));
    throw std::runtime_error(m);
  }
  if (bytes > 0) BCHK(hipMemcpyAsync(dst, src.data(), bytes, hipMemcpyHostToDevice, st));
}

void launch_k(BatchKernel k, int grid, hipStream_t st, const SlotDev* sd, const int* stp,
              const int* slots, const int* off, int ns) {
  void* argv[] = {(void*)&sd, (void*)&stp, (void*)&slots, (void*)&off, (void*)&ns};
  BCHK(hipLaunchKernel((const void*)k, dim3(grid), dim3(fkd::WG), argv, 0, st));
}

// Upload a request into its slot's buffers and prepare H, R on the device.
void admit_one(fkpm_batch* B, Slot& sl) {
  Request& r = *sl.rq;
  const Layout& L = sl.lay;
  char* base = B->arena.base;
  auto P = [&](int k) { return (void*)(base + L.off[k]); };
  hipStream_t st = B->stream;
  h2d(P(B_RP), r.rp, L.len[B_RP], st);
  h2d(P(B_CI), r.ci, L.len[B_CI], st);
  h2d(P(B_RB), r.rowblk, L.len[B_RB], st);
  h2d(P(B_TP), r.tpos, L.len[B_TP], st);
  h2d(P(B_HRAW), r.val, L.len[B_HRAW], st);
  // compact R: rcol then rval (16-byte aligned) in the staging buffer
  const size_t rv_off = ((size_t)r.n * sizeof(int) + 255) & ~(size_t)255;
  if (r.seeded) {
    // colours and column seeds up; the phases are drawn on the device
    const size_t sd_off = (rv_off + sizeof(double) * 2 * r.n + 255) & ~(size_t)255;
    h2d(P(B_RRAW), r.rcol, sizeof(int) * r.n, st);
    h2d((char*)P(B_RRAW) + sd_off, r.seeds, sizeof(unsigned) * r.s, st);
    const int* g = (const int*)P(B_RRAW);
    const unsigned* sd = (const unsigned*)((char*)P(B_RRAW) + sd_off);
    double* rv = (double*)((char*)P(B_RRAW) + rv_off);
    const int n = r.n;
    void* argv[] = {(void*)&g, (void*)&n, (void*)&sd, (void*)&rv};
    BCHK(hipLaunchKernel((const void*)&fkd::k_probe_phases, dim3(r.s), dim3(64), argv, 0, st));
  } else if (r.compact) {
    h2d(P(B_RRAW), r.rcol, sizeof(int) * r.n, st);
    h2d((char*)P(B_RRAW) + rv_off, r.R, sizeof(double) * 2 * r.n, st);
  } else {
    h2d(P(B_RRAW), r.R, L.len[B_RRAW], st);
  }
  h2d(P(B_C), r.c, L.len[B_C], st);
  h2d(P(B_RD), r.rdesc, L.len[B_RD], st);
  const double avg = (r.hi + r.lo) / 2, mag = (r.hi - r.lo) / 2;
  if (fkpm_internal::scale_H_cx(st, (const double*)P(B_HRAW), P(B_HV), B->single,
                                (const int*)P(B_RB), (const int*)P(B_CI), r.nnzb, r.b, mag,
                                avg / mag) != 0)
    throw std::runtime_error(fkpm_last_error());
  if (r.compact) {
    const int* rc = (const int*)P(B_RRAW);
    const double* rv = (const double*)((char*)P(B_RRAW) + rv_off);
    const size_t tot = (size_t)r.n * L.sp;
    const int grid = (int)std::min<size_t>((tot + fkd::WG - 1) / fkd::WG, 65536);
    const int n = r.n, sp = L.sp;
    if (B->single) {
      float* Rd = (float*)P(B_R);
      void* argv[] = {(void*)&rc, (void*)&rv, (void*)&n, (void*)&sp, (void*)&Rd};
      BCHK(hipLaunchKernel((const void*)&fkd::k_batch_scatter_R<float>, dim3(grid), dim3(fkd::WG),
                           argv, 0, st));
    } else {
      double* Rd = (double*)P(B_R);
      void* argv[] = {(void*)&rc, (void*)&rv, (void*)&n, (void*)&sp, (void*)&Rd};
      BCHK(hipLaunchKernel((const void*)&fkd::k_batch_scatter_R<double>, dim3(grid),
                           dim3(fkd::WG), argv, 0, st));
    }
  } else if (fkpm_internal::pack_R_cx(st, (const double*)P(B_RRAW), P(B_R), B->single, r.n, r.s,
                                      L.sp) != 0) {
    throw std::runtime_error(fkpm_last_error());
  }
  SlotDev& d = sl.dev;
  std::memset(&d, 0, sizeof(d));
  d.rp = (const int*)P(B_RP);
  d.ci = (const int*)P(B_CI);
  d.hv = P(B_HV);
  d.tpos = (const int*)P(B_TP);
  d.R = P(B_R);
  d.A[0] = P(B_A0);
  d.A[1] = P(B_A1);
  d.A[2] = P(B_A2);
  d.Bv[0] = P(B_B0);
  d.Bv[1] = P(B_B1);
  d.Bv[2] = P(B_B2);
  d.dacc = (double*)P(B_DACC);
  d.dout = (double*)P(B_DOUT);
  d.partial = (double*)P(B_PART);
  d.psum = (double*)P(B_PSUM);
  d.c = (const double*)P(B_C);
  d.rord = r.rdesc.empty() ? nullptr : (const int*)P(B_RD);
  d.nt = r.rdesc.empty() ? 0 : fkd::NT_ORDER;
  d.nb = r.nb;
  d.qp = L.qp;
  d.chunk = L.chunk;
  d.nwg = L.nwg;
  d.M = r.M;
  d.inv_mag = 1.0 / mag;
  double diag = r.c[1];
  for (int m = 1; m < r.M / 2; m++) diag -= r.c[2 * m + 1];
  d.diag = diag;
  d.sb = 2 * r.c[r.M - 1];
  B->h_slots[sl.idx] = d;
  BCHK(hipMemcpyAsync(B->d_slots + sl.idx, &B->h_slots[sl.idx], sizeof(SlotDev),
                      hipMemcpyHostToDevice, st));
  // the slot's step counter starts at 0 (the host mirror holds it; the copy
  // completes in upload_tables' synchronisation before the mirror is reread)
  B->h_st[sl.idx] = 0;
  BCHK(hipMemcpyAsync(B->d_st + sl.idx, &B->h_st[sl.idx], sizeof(int), hipMemcpyHostToDevice, st));
}

void release_staging(fkpm_batch* B, Slot& sl) {
  if (!sl.staging_live) return;
  B->arena.release(sl.lay.off[B_HRAW], sl.lay.len[B_HRAW]);
  B->arena.release(sl.lay.off[B_RRAW], sl.lay.len[B_RRAW]);
  B->arena.release(sl.lay.off[B_RB], sl.lay.len[B_RB]);
  sl.lay.len[B_HRAW] = sl.lay.len[B_RRAW] = sl.lay.len[B_RB] = 0;
  sl.staging_live = false;
}

void release_slot(fkpm_batch* B, Slot& sl) {
  release_staging(B, sl);
  for (int k = 0; k < B_COUNT; k++)
    if (k != B_HRAW && k != B_RRAW && k != B_RB) B->arena.release(sl.lay.off[k], sl.lay.len[k]);
  B->free_idx.push_back(sl.idx);
}

// Try to allocate every buffer of a request; all or nothing.
bool try_alloc(fkpm_batch* B, Layout& L) {
  int k = 0;
  for (; k < B_COUNT; k++)
    if (!B->arena.alloc(L.len[k], &L.off[k])) break;
  if (k == B_COUNT) return true;
  for (int j = 0; j < k; j++) B->arena.release(L.off[j], L.len[j]);
  return false;
}

// Per-class slot lists, workgroup prefix tables and the active list.
void upload_tables(fkpm_batch* B) {
  const auto& cl = classes(B->single);
  const size_t nc = cl.size();
  for (size_t c = 0; c < nc; c++) {
    B->h_cls_slots[c].clear();
    B->h_cls_off[c].assign(1, 0);
  }
  B->h_act.clear();
  for (const Slot& sl : B->active) {
    const int c = sl.lay.cls;
    B->h_cls_slots[c].push_back(sl.idx);
    B->h_cls_off[c].push_back(B->h_cls_off[c].back() + sl.lay.nwg);
    B->h_act.push_back(sl.idx);
  }
  for (size_t c = 0; c < nc; c++) {
    if (B->h_cls_slots[c].empty()) continue;
    BCHK(hipMemcpyAsync(B->d_cls_slots[c], B->h_cls_slots[c].data(),
                        sizeof(int) * B->h_cls_slots[c].size(), hipMemcpyHostToDevice, B->stream));
    BCHK(hipMemcpyAsync(B->d_cls_off[c], B->h_cls_off[c].data(),
                        sizeof(int) * B->h_cls_off[c].size(), hipMemcpyHostToDevice, B->stream));
  }
  if (!B->h_act.empty())
    BCHK(hipMemcpyAsync(B->d_act, B->h_act.data(), sizeof(int) * B->h_act.size(),
                        hipMemcpyHostToDevice, B->stream));
  BCHK(hipStreamSynchronize(B->stream));  // host tables are reused next time
}

void enqueue_ticks(fkpm_batch* B, int T) {
  const auto& cl = classes(B->single);
  const int na = (int)B->h_act.size();
  for (int t = 0; t < T; t++) {
    for (size_t c = 0; c < cl.size(); c++) {
      const int ns = (int)B->h_cls_slots[c].size();
      if (ns == 0) continue;
      const int grid = B->h_cls_off[c].back();
      launch_k(cl[c].fwd, grid, B->stream, B->d_slots, B->d_st, B->d_cls_slots[c],
               B->d_cls_off[c], ns);
      launch_k(cl[c].bwd, grid, B->stream, B->d_slots, B->d_st, B->d_cls_slots[c],
               B->d_cls_off[c], ns);
      launch_k(cl[c].aux, grid, B->stream, B->d_slots, B->d_st, B->d_cls_slots[c],
               B->d_cls_off[c], ns);
    }
    int* st = B->d_st;
    const int* act = B->d_act;
    void* argv[] = {(void*)&st, (void*)&act, (void*)&na};
    BCHK(hipLaunchKernel((const void*)&fkd::k_batch_advance, dim3((na + 255) / 256), dim3(256),
                         argv, 0, B->stream));
  }
}

void run_ticks(fkpm_batch* B, int T, bool use_graph) {
  if (!use_graph) {
    enqueue_ticks(B, T);
  } else {
    if (B->graph_version != B->set_version) {
      B->clear_graphs();
      B->graph_version = B->set_version;
    }
    auto it = B->graphs.find(T);
    if (it == B->graphs.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t ex = nullptr;
      BCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_ticks(B, T);
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(B->stream, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
      }
      BCHK(hipStreamEndCapture(B->stream, &g));
      BCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      it = B->graphs.emplace(T, ex).first;
      B->n_captures++;
    }
    BCHK(hipGraphLaunch(it->second, B->stream));
  }
  B->n_ticks += T;
  B->n_replays++;
}

// Moments and D of a finished slot to its request.
void retire(fkpm_batch* B, Slot& sl) {
  Request& r = *sl.rq;
  const int h = r.M / 2;
  std::vector<double> ps((size_t)h * 2);
  r.D.resize((size_t)r.nnzb * r.b * r.b * 2);
  BCHK(hipMemcpyAsync(ps.data(), sl.dev.psum, ps.size() * sizeof(double), hipMemcpyDeviceToHost,
                      B->stream));
  BCHK(hipMemcpyAsync(r.D.data(), sl.dev.dout, r.D.size() * sizeof(double), hipMemcpyDeviceToHost,
                      B->stream));
  BCHK(hipStreamSynchronize(B->stream));
  r.mu.assign(r.M, 0.0);
  r.mu[0] = ps[0];
  r.mu[1] = ps[1];
  for (int m = 1; m < h; m++) {
    r.mu[2 * m] = 2 * ps[2 * m] - r.mu[0];
    r.mu[2 * m + 1] = 2 * ps[2 * m + 1] - r.mu[1];
  }
  double e = 0.0;
  for (int m = 0; m < r.M; m++) e += r.c[m] * r.mu[m];
  r.energy = e;
  // inputs are no longer needed
  std::vector<double>().swap(r.val);
  std::vector<double>().swap(r.R);
}

void fail_all(fkpm_batch* B, const std::string& msg) {
  std::lock_guard<std::mutex> lk(B->mu);
  B->broken = msg;
  for (Slot& sl : B->active) {
    sl.rq->status = FKPM_E_HIP;
    sl.rq->err = msg;
  }
  B->active.clear();
  for (auto& r : B->pending) {
    r->status = FKPM_E_HIP;
    r->err = msg;
  }
  B->pending.clear();
  B->cv_done.notify_all();
}

// Closing: every request still queued or running fails with FKPM_E_STATE and
// its waiters are woken (called with B->mu held).
void fail_unfinished_locked(fkpm_batch* B) {
  for (Slot& sl : B->active) {
    sl.rq->status = FKPM_E_STATE;
    sl.rq->err = "the batch engine was closed before the request finished";
  }
  for (auto& r : B->pending) {
    r->status = FKPM_E_STATE;
    r->err = "the batch engine was closed before the request was admitted";
  }
  B->pending.clear();
  B->cv_done.notify_all();
}

void scheduler(fkpm_batch* B) {
  try {
    BCHK(hipSetDevice(B->device));
    for (;;) {
      std::vector<size_t> fresh;  // positions in `active` of slots admitted now
      bool use_graph = true;
      int max_ticks = 1;
      {
        std::unique_lock<std::mutex> lk(B->mu);
        B->cv_work.wait(lk, [&] { return B->stop || !B->pending.empty() || !B->active.empty(); });
        if (B->stop) {
          // the previous iteration synchronised the stream: nothing is in flight
          fail_unfinished_locked(B);
          return;
        }
        use_graph = B->use_graph;  // options are written under the lock
        max_ticks = B->max_ticks;
        // admission: FIFO while a slot is free and the footprint fits
        while (!B->pending.empty() && (int)B->active.size() < B->max_slots) {
          std::shared_ptr<Request> r = B->pending.front();
          Slot sl;
          plan(B->single, *r, &sl.lay);
          if (!try_alloc(B, sl.lay)) {
            if (B->active.empty()) {  // cannot fit even alone: refuse it
              B->pending.pop_front();
              r->status = FKPM_E_ARG;
              r->err = "request does not fit the batch's HBM budget";
              B->cv_done.notify_all();
              continue;
            }
            break;
          }
          B->pending.pop_front();
          sl.rq = r;
          sl.idx = B->free_idx.back();
          B->free_idx.pop_back();
          B->h_st[sl.idx] = 0;
          r->status = 1;
          r->t_admit = Clock::now();
          fresh.push_back(B->active.size());
          B->active.push_back(sl);
        }
        B->peak_slots = std::max(B->peak_slots, (int)B->active.size());
      }
      if (!fresh.empty()) {
        for (size_t k : fresh) admit_one(B, B->active[k]);
        upload_tables(B);
        B->set_version++;
      }
      if (B->active.empty()) continue;
      // ticks until the first slot finishes (at most max_ticks)
      int T = max_ticks;
      for (const Slot& sl : B->active) {
        const int left = sl.rq->M + 2 - B->h_st[sl.idx];
        T = std::min(T, std::max(left, 1));
      }
      run_ticks(B, T, use_graph);
      BCHK(hipStreamSynchronize(B->stream));
      // new slots' staging buffers were consumed by their first ticks
      for (Slot& sl : B->active) release_staging(B, sl);
      BCHK(hipMemcpy(B->h_st.data(), B->d_st, sizeof(int) * B->max_slots, hipMemcpyDeviceToHost));
      std::vector<Slot> keep;
      std::vector<Slot> done;
      for (Slot& sl : B->active) {
        if (B->h_st[sl.idx] >= sl.rq->M + 2) done.push_back(sl); else keep.push_back(sl);
      }
      for (Slot& sl : done) retire(B, sl);
      if (!done.empty()) {
        B->active.swap(keep);
        upload_tables(B);
        B->set_version++;
        std::lock_guard<std::mutex> lk(B->mu);
        const auto now = Clock::now();
        for (Slot& sl : done) {
          sl.rq->t_done = now;
          sl.rq->status = 2;
          release_slot(B, sl);
        }
        B->cv_done.notify_all();
      }
    }
  } catch (const std::exception& e) {
    fail_all(B, e.what());
  }
}

int berr(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return fkpm_internal::set_error(code, buf);
}

}  // namespace

extern "C" {

int fkpm_batch_footprint(int prec, int nb, int b, int nnzb, int s, int M, size_t* bytes) {
  if (!bytes || nb <= 0 || nnzb < nb || s <= 0 || M < 2 || (b != 2 && b != 8))
    return berr(FKPM_E_ARG, "batch_footprint: bad arguments");
  Request r;
  r.nb = nb;
  r.b = b;
  r.nnzb = nnzb;
  r.n = nb * b;
  r.s = s;
  r.M = M;
  Layout L;
  plan(prec == FKPM_PREC_FP32, r, &L);
  *bytes = L.total;
  if (b == 2 && nb >= 4096) {  // a 3-D lattice pattern also holds its row order
    r.rdesc.assign(3 * (size_t)nb, 0);
    plan(prec == FKPM_PREC_FP32, r, &L);
    *bytes = std::max(*bytes, L.total);
  }
  return 0;
}

int fkpm_batch_open(int device, int prec, size_t hbm_bytes, int max_slots, fkpm_batch** out) {
  if (!out || max_slots < 1 || max_slots > 4096 || hbm_bytes < (1u << 20))
    return berr(FKPM_E_ARG, "batch_open: bad arguments");
  if (prec != FKPM_PREC_FP64 && prec != FKPM_PREC_FP32)
    return berr(FKPM_E_ARG, "batch_open: bad precision %d", prec);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return berr(FKPM_E_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return berr(FKPM_E_ARG, "bad device %d", device);
  std::unique_ptr<fkpm_batch> B(new fkpm_batch());
  B->device = device;
  B->single = prec == FKPM_PREC_FP32;
  B->max_slots = max_slots;
  try {
    BCHK(hipSetDevice(device));
    BCHK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    BCHK(hipMalloc((void**)&B->arena_mem, hbm_bytes));
    B->arena.init(B->arena_mem, hbm_bytes);
    BCHK(hipMalloc((void**)&B->d_slots, sizeof(SlotDev) * max_slots));
    BCHK(hipMalloc((void**)&B->d_st, sizeof(int) * max_slots));
    BCHK(hipMalloc((void**)&B->d_act, sizeof(int) * max_slots));
    const size_t nc = classes(B->single).size();
    B->d_cls_slots.assign(nc, nullptr);
    B->d_cls_off.assign(nc, nullptr);
    for (size_t c = 0; c < nc; c++) {
      BCHK(hipMalloc((void**)&B->d_cls_slots[c], sizeof(int) * max_slots));
      BCHK(hipMalloc((void**)&B->d_cls_off[c], sizeof(int) * (max_slots + 1)));
    }
    B->h_cls_slots.assign(nc, {});
    B->h_cls_off.assign(nc, {});
    B->h_slots.assign(max_slots, SlotDev());
    B->h_st.assign(max_slots, 0);
    std::vector<int> zeros(max_slots, 0);
    BCHK(hipMemcpy(B->d_st, zeros.data(), sizeof(int) * max_slots, hipMemcpyHostToDevice));
    for (int k = max_slots - 1; k >= 0; k--) B->free_idx.push_back(k);
  } catch (const std::exception& e) {
    return berr(FKPM_E_HIP, "%s", e.what());
  }
  B->worker = std::thread(scheduler, B.get());
  *out = B.release();
  return 0;
}

int fkpm_batch_close(fkpm_batch* B) {
  if (!B) return berr(FKPM_E_ARG, "null batch");
  {
    std::lock_guard<std::mutex> lk(B->mu);
    B->stop = true;
  }
  B->cv_work.notify_all();
  if (B->worker.joinable()) B->worker.join();
  {
    // the scheduler failed every unfinished request on its way out; wait
    // until every thread woken in fkpm_batch_wait has left it before the
    // mutex and condition variables go away
    std::unique_lock<std::mutex> lk(B->mu);
    if (B->broken.empty()) fail_unfinished_locked(B);  // scheduler never started its loop
    B->cv_idle.wait(lk, [&] { return B->waiters == 0; });
  }
  (void)hipSetDevice(B->device);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  B->clear_graphs();
  for (size_t c = 0; c < B->d_cls_slots.size(); c++) {
    (void)hipFree(B->d_cls_slots[c]);
    (void)hipFree(B->d_cls_off[c]);
  }
  (void)hipFree(B->d_slots);
  (void)hipFree(B->d_st);
  (void)hipFree(B->d_act);
  (void)hipFree(B->arena_mem);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return 0;
}

int fkpm_batch_set_option(fkpm_batch* B, int option, int value) {
  if (!B) return berr(FKPM_E_ARG, "null batch");
  std::lock_guard<std::mutex> lk(B->mu);
  switch (option) {
    case FKPM_BATCH_OPT_GRAPH:
      B->use_graph = value != 0;
      return 0;
    case FKPM_BATCH_OPT_TICKS:
      if (value < 1 || value > 4096) return berr(FKPM_E_ARG, "bad tick count %d", value);
      B->max_ticks = value;
      return 0;
  }
  return berr(FKPM_E_ARG, "unknown batch option %d", option);
}

static int submit_common(fkpm_batch* B, int nb, int b, int nnzb, const int* row_ptr,
                         const int* col_idx, const void* val, double lo, double hi, int n, int s,
                         const void* R_colmajor, const int* rcol, const void* rval, int M,
                         const double* c, long long* id, const unsigned* seeds = nullptr) {
  if (!B || !row_ptr || !col_idx || !val || !(R_colmajor || (rcol && (rval || seeds))) || !c ||
      !id)
    return berr(FKPM_E_ARG, "batch_submit: null argument");
  if (b != 2 && b != 8) return berr(FKPM_E_UNSUP, "batch: block size %d (2 or 8 supported)", b);
  if (nb <= 0 || nnzb < nb || n != nb * b || s <= 0 || M < 2 || M % 2 != 0 || !(hi > lo))
    return berr(FKPM_E_ARG, "batch_submit: bad sizes (nb %d, b %d, n %d, s %d, M %d)", nb, b, n,
                s, M);
  auto r = std::make_shared<Request>();
  r->nb = nb;
  r->b = b;
  r->nnzb = nnzb;
  r->n = n;
  r->s = s;
  r->M = M;
  r->lo = lo;
  r->hi = hi;
  r->rp.assign(row_ptr, row_ptr + nb + 1);
  r->ci.assign(col_idx, col_idx + nnzb);
  if (r->rp[0] != 0 || r->rp[nb] != nnzb) return berr(FKPM_E_ARG, "batch_submit: bad row_ptr");
  r->rowblk.resize(nnzb);
  for (int i = 0; i < nb; i++) {
    if (r->rp[i + 1] < r->rp[i]) return berr(FKPM_E_ARG, "batch_submit: row_ptr not monotone");
    bool diag = false;
    for (int p = r->rp[i]; p < r->rp[i + 1]; p++) {
      const int j = r->ci[p];
      if (j < 0 || j >= nb || (p > r->rp[i] && j <= r->ci[p - 1]))
        return berr(FKPM_E_ARG, "batch_submit: column indices of row %d unsorted/out of range", i);
      diag |= (j == i);
      r->rowblk[p] = i;
    }
    if (!diag) return berr(FKPM_E_ARG, "batch_submit: diagonal block %d missing", i);
    r->max_deg = std::max(r->max_deg, r->rp[i + 1] - r->rp[i]);
  }
  // transpose positions (structurally symmetric pattern, as symmetrize needs)
  r->tpos.assign(nnzb, -1);
  for (int i = 0; i < nb; i++)
    for (int p = r->rp[i]; p < r->rp[i + 1]; p++) {
      const int j = r->ci[p];
      const int* b0 = r->ci.data() + r->rp[j];
      const int* b1 = r->ci.data() + r->rp[j + 1];
      const int* f = std::lower_bound(b0, b1, i);
      if (f == b1 || *f != i) return berr(FKPM_E_ARG, "batch_submit: pattern not symmetric");
      r->tpos[p] = (int)(f - r->ci.data());
    }
  const size_t hvn = (size_t)nnzb * b * b * 2;
  r->val.assign((const double*)val, (const double*)val + hvn);
  if (R_colmajor) {
    r->R.assign((const double*)R_colmajor, (const double*)R_colmajor + (size_t)n * s * 2);
  } else if (seeds) {
    r->compact = r->seeded = true;
    r->rcol.assign(rcol, rcol + n);
    for (int i = 0; i < n; i++)
      if (r->rcol[i] < 0 || r->rcol[i] >= s)
        return berr(FKPM_E_ARG, "batch_submit_seeded: colour %d of row %d not in [0, %d)",
                    r->rcol[i], i, s);
    r->seeds.assign(seeds, seeds + s);
  } else {
    r->compact = true;
    r->rcol.assign(rcol, rcol + n);
    for (int i = 0; i < n; i++)
      if (r->rcol[i] < -1 || r->rcol[i] >= s)
        return berr(FKPM_E_ARG, "batch_submit: probing column %d of row %d out of range",
                    r->rcol[i], i);
    r->R.assign((const double*)rval, (const double*)rval + (size_t)n * 2);
  }
  r->c.assign(c, c + M);
  // large 3-D lattices with 2x2 blocks (the vector-unit classes) sweep their
  // rows in the lattice row order, as the single-request engine's tuner
  // chooses for them
  if (b == 2 && nb >= 4096) {
    std::vector<int> ord;
    if (fkpm_internal::lattice_row_order(nb, r->rp.data(), r->ci.data(), ord)) {
      r->rdesc.resize(3 * (size_t)nb);
      for (int k = 0; k < nb; k++) {
        r->rdesc[3 * (size_t)k] = ord[k];
        r->rdesc[3 * (size_t)k + 1] = r->rp[ord[k]];
        r->rdesc[3 * (size_t)k + 2] = r->rp[ord[k] + 1];
      }
    }
  }
  Layout L;
  plan(B->single, *r, &L);
  if (L.cls < 0) return berr(FKPM_E_UNSUP, "batch: no kernel class for block %d", b);
  std::lock_guard<std::mutex> lk(B->mu);
  if (!B->broken.empty()) return berr(FKPM_E_HIP, "batch engine failed: %s", B->broken.c_str());
  if (L.total > B->arena.size)
    return berr(FKPM_E_ARG, "request needs %zu bytes, more than the batch's HBM budget %zu",
                L.total, B->arena.size);
  r->id = B->next_id++;
  r->t_submit = Clock::now();
  B->pending.push_back(r);
  B->reqs[r->id] = r;
  *id = r->id;
  B->cv_work.notify_one();
  return 0;
}

int fkpm_batch_submit(fkpm_batch* B, int nb, int b, int nnzb, const int* row_ptr,
                      const int* col_idx, const void* val, double lo, double hi, int n, int s,
                      const void* R_colmajor, int M, const double* c, long long* id) {
  return submit_common(B, nb, b, nnzb, row_ptr, col_idx, val, lo, hi, n, s, R_colmajor, nullptr,
                       nullptr, M, c, id);
}

int fkpm_batch_submit_probing(fkpm_batch* B, int nb, int b, int nnzb, const int* row_ptr,
                              const int* col_idx, const void* val, double lo, double hi, int n,
                              int s, const int* rcol, const void* rval, int M, const double* c,
                              long long* id) {
  if (!rcol || !rval) return berr(FKPM_E_ARG, "batch_submit_probing: null probing vectors");
  return submit_common(B, nb, b, nnzb, row_ptr, col_idx, val, lo, hi, n, s, nullptr, rcol, rval,
                       M, c, id);
}

int fkpm_batch_submit_seeded(fkpm_batch* B, int nb, int b, int nnzb, const int* row_ptr,
                             const int* col_idx, const void* val, double lo, double hi, int n,
                             int s, const int* groups, const unsigned* col_seeds, int M,
                             const double* c, long long* id) {
  if (!groups || !col_seeds) return berr(FKPM_E_ARG, "batch_submit_seeded: null colours/seeds");
  if (s >= n)
    return berr(FKPM_E_ARG, "batch_submit_seeded: %d colours for %d rows (R = I: submit it)", s,
                n);
  return submit_common(B, nb, b, nnzb, row_ptr, col_idx, val, lo, hi, n, s, nullptr, groups,
                       nullptr, M, c, id, col_seeds);
}

int fkpm_batch_poll(fkpm_batch* B, long long id, int* done) {
  if (!B || !done) return berr(FKPM_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(B->mu);
  auto it = B->reqs.find(id);
  if (it == B->reqs.end()) return berr(FKPM_E_ARG, "unknown request %lld", id);
  *done = it->second->status == 2 || it->second->status < 0 ? 1 : 0;
  return 0;
}

int fkpm_batch_wait(fkpm_batch* B, long long id, double* mu, double* energy, double* D_val,
                    double* times_ms) {
  if (!B) return berr(FKPM_E_ARG, "null batch");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(B->mu);
    auto it = B->reqs.find(id);
    if (it == B->reqs.end()) return berr(FKPM_E_ARG, "unknown request %lld", id);
    r = it->second;
    B->waiters++;
    B->cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0; });
    B->reqs.erase(r->id);
    if (--B->waiters == 0) B->cv_idle.notify_all();
  }
  if (r->status < 0) return berr(r->status, "request %lld failed: %s", id, r->err.c_str());
  if (mu) std::copy(r->mu.begin(), r->mu.end(), mu);
  if (energy) *energy = r->energy;
  if (D_val) std::copy(r->D.begin(), r->D.end(), D_val);
  if (times_ms) {
    auto ms = [](Clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
    times_ms[0] = ms(r->t_admit - r->t_submit);
    times_ms[1] = ms(r->t_done - r->t_admit);
    times_ms[2] = ms(r->t_done - r->t_submit);
  }
  return 0;
}

int fkpm_batch_stats(fkpm_batch* B, long long* ticks, long long* replays, long long* captures,
                     int* peak_slots, size_t* peak_bytes) {
  if (!B) return berr(FKPM_E_ARG, "null batch");
  std::lock_guard<std::mutex> lk(B->mu);
  if (ticks) *ticks = B->n_ticks;
  if (replays) *replays = B->n_replays;
  if (captures) *captures = B->n_captures;
  if (peak_slots) *peak_slots = B->peak_slots;
  if (peak_bytes) *peak_bytes = B->arena.peak;
  return 0;
}

}  // extern "C"
