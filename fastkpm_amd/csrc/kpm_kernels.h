// kpm_kernels.h -- gfx950 kernels of the stochastic Chebyshev engine.
//
// Layout in HBM (all kernels):
//   Hs        BSR: row_ptr[nb+1], col_idx[nnzb] (int32), val[nnzb*B*B] in the
//             storage type S (complex interleaved), blocks column-major.
//   vectors   alpha_m, beta_m, R: N x Sp row-major (the Sp probe columns of one
//             orbital contiguous), Sp = S padded so that a row is a whole
//             number of 16-byte packets per lane group; padded columns are 0
//             and stay 0 under the recursion.
//   xi, D     FP64 accumulators (complex interleaved when CX).
//
// Work decomposition: a wavefront is split into 64/L groups of L lanes; a
// group owns one block row i (B orbital rows) and sweeps its Sp columns in
// passes of L*KQ packets, each lane holding KQ 16-byte packets per orbital row.
// The neighbour blocks of a row are gathered U at a time: the column indices,
// the U*B*KQ vector packets and the U H blocks are all requested before any
// arithmetic, so that a wave keeps many loads in flight.  Workgroups take
// contiguous ranges of block rows, and the workgroup index is remapped so that
// each of the 8 XCDs (which receive workgroups round-robin) owns one
// contiguous eighth of the rows: the gathers of a neighbour row then hit that
// XCD's L2.
//
// Precision: the row sums of the SpMM run in the storage precision S (FP64 or
// FP32); moment dots, xi and D are accumulated in FP64.
//
// Reductions (moment dots, D blocks) are deterministic: fixed DPP lane
// patterns, fixed LDS order, per-workgroup partials reduced in index order by a
// second kernel.  No floating-point atomics anywhere.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fkd {

constexpr int WAVE = 64;
constexpr int WG = 256;

template <typename T> struct Vec16;
template <> struct Vec16<double> {
  typedef double V __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};
template <> struct Vec16<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};

// Load NSC scalars of type T (a whole number of 16-byte vectors) as type A.
template <typename T, int NSC, typename A>
__device__ __forceinline__ void ldv(const T* __restrict__ p, A* v) {
  typedef typename Vec16<T>::V VT;
  constexpr int PER = Vec16<T>::N;
  static_assert(NSC % PER == 0, "load width");
#pragma unroll
  for (int c = 0; c < NSC / PER; c++) {
    VT t = *reinterpret_cast<const VT*>(p + c * PER);
#pragma unroll
    for (int e = 0; e < PER; e++) v[c * PER + e] = (A)t[e];
  }
}

// Store NSC values of type A as scalars of type T.
template <typename T, int NSC, typename A>
__device__ __forceinline__ void stv(T* __restrict__ p, const A* v) {
  typedef typename Vec16<T>::V VT;
  constexpr int PER = Vec16<T>::N;
  static_assert(NSC % PER == 0, "store width");
#pragma unroll
  for (int c = 0; c < NSC / PER; c++) {
    VT t;
#pragma unroll
    for (int e = 0; e < PER; e++) t[e] = (T)v[c * PER + e];
    *reinterpret_cast<VT*>(p + c * PER) = t;
  }
}

// Streaming variants: the nontemporal hint keeps data that is read or
// written once per step from displacing the gathered vector in the caches.
template <typename T, int NSC, typename A>
__device__ __forceinline__ void ldv_nt(const T* __restrict__ p, A* v, bool nt) {
  if (!nt) {
    ldv<T, NSC>(p, v);
    return;
  }
  typedef typename Vec16<T>::V VT;
  constexpr int PER = Vec16<T>::N;
#pragma unroll
  for (int c = 0; c < NSC / PER; c++) {
    VT t = __builtin_nontemporal_load(reinterpret_cast<const VT*>(p + c * PER));
#pragma unroll
    for (int e = 0; e < PER; e++) v[c * PER + e] = (A)t[e];
  }
}

template <typename T, int NSC, typename A>
__device__ __forceinline__ void stv_nt(T* __restrict__ p, const A* v, bool nt) {
  if (!nt) {
    stv<T, NSC>(p, v);
    return;
  }
  typedef typename Vec16<T>::V VT;
  constexpr int PER = Vec16<T>::N;
#pragma unroll
  for (int c = 0; c < NSC / PER; c++) {
    VT t;
#pragma unroll
    for (int e = 0; e < PER; e++) t[e] = (T)v[c * PER + e];
    __builtin_nontemporal_store(t, reinterpret_cast<VT*>(p + c * PER));
  }
}

// Bits of the `nt` kernel argument.
constexpr int NT_LOAD = 1;   // loads of the operands read once per step
constexpr int NT_STORE = 2;  // stores of the step's outputs
// Loop order of the step kernels: without NT_SLICE a lane group finishes
// every column pass of a block row before the next row; with it the
// workgroups sweep all their rows once per column pass (L*KQ packets), so the
// rows one XCD touches between a row and its far neighbours (a 3-D lattice's
// next plane) take 1/passes of the bytes and stay in its L2, at the cost of
// reading H's indices and values once per pass.
constexpr int NT_SLICE = 4;
// Row order: `rp` holds three ints per position k of the sweep, {row, first
// block, end block}, instead of the row offsets, and the workgroups sweep the
// positions (one contiguous range per XCD) rather than the rows.  The host
// builds it for a periodic 3-D lattice so that each XCD walks its slab tile by
// tile and a row's far neighbours (the next plane) are still in its L2.  A
// row's arithmetic is the same in either order.
constexpr int NT_ORDER = 8;

// DPP lane move inside 16-lane rows (gfx9 encodings: quad_perm 0x00-0xFF,
// row_mirror 0x140, row_half_mirror 0x141).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over aligned groups of L lanes, result in every lane of the group.
// Levels 1, 2: quad_perm [1,0,3,2] and [2,3,0,1]; 4: row_half_mirror (lane
// l <- 7-l, the other quad of the half row); 8: row_mirror (l <- 15-l, the
// other half of the row); 16, 32: cross-row swizzles.  Every lane of a group
// adds the same two partial sums (in either order), so all lanes agree.
template <int L>
__device__ __forceinline__ double group_sum(double v) {
  if constexpr (L >= 2) v += dpp_mov<0xB1>(v);
  if constexpr (L >= 4) v += dpp_mov<0x4E>(v);
  if constexpr (L >= 8) v += dpp_mov<0x141>(v);
  if constexpr (L >= 16) v += dpp_mov<0x140>(v);
  if constexpr (L >= 32) v += __shfl_xor(v, 16, WAVE);
  if constexpr (L >= 64) v += __shfl_xor(v, 32, WAVE);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum<WAVE>(v); }

// The value held by lane l ^ M (M a power of two < 64): quad_perm DPP for
// M = 1, 2; ds_swizzle in bitmask mode (lane ((l & 0x1F) ^ M) of the 32-lane
// half) for M = 4, 8, 16; a shuffle for M = 32.  Doubles move as two dwords.
template <int M>
__device__ __forceinline__ double xor_move(double v) {
  if constexpr (M == 1) {
    return dpp_mov<0xB1>(v);
  } else if constexpr (M == 2) {
    return dpp_mov<0x4E>(v);
  } else if constexpr (M < 32) {
    constexpr int PAT = (M << 10) | 0x1F;
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)b, PAT);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), PAT);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
  } else {
    return __shfl_xor(v, M, WAVE);
  }
}

template <int M>
__device__ __forceinline__ float xor_move(float v) {
  const int b = __builtin_bit_cast(int, v);
  if constexpr (M == 1) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, b, 0xB1, 0xF, 0xF, false));
  } else if constexpr (M == 2) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, b, 0x4E, 0xF, 0xF, false));
  } else if constexpr (M < 32) {
    constexpr int PAT = (M << 10) | 0x1F;
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(b, PAT));
  } else {
    return __shfl_xor(v, M, WAVE);
  }
}

// Reduce-scatter of CNT values over the lanes M, 2M, ... < L of a group: at
// each level a lane keeps the half of its values selected by lane bit M, adds
// the partner's copy of that half and sends the other half; once one value
// is left the remaining levels are plain butterflies.  Afterwards lane l holds
// in v[0 .. max(NV/L, 1)) the group sums of indices rs_base<L, NV>(l) + t.
template <int M, int L, int CNT, typename T>
__device__ __forceinline__ void rs_level(T* v, int l) {
  if constexpr (M < L) {
    if constexpr (CNT > 1) {
      constexpr int H = CNT / 2;
      const bool up = (l & M) != 0;
#pragma unroll
      for (int t = 0; t < H; t++) {
        const T send = up ? v[t] : v[t + H];
        const T keep = up ? v[t + H] : v[t];
        v[t] = keep + xor_move<M>(send);
      }
      rs_level<2 * M, L, H>(v, l);
    } else {
      v[0] += xor_move<M>(v[0]);
      rs_level<2 * M, L, 1>(v, l);
    }
  }
}

template <int L, int NV, typename T>
__device__ __forceinline__ void reduce_scatter(T* v, int l) {
  static_assert((NV & (NV - 1)) == 0, "power of two");
  rs_level<1, L, NV>(v, l);
}

// First index owned by lane l after reduce_scatter<L, NV>.
template <int L, int NV>
__device__ __forceinline__ int rs_base(int l) {
  int base = 0, cnt = NV;
#pragma unroll
  for (int m = 1; m < L && cnt > 1; m <<= 1) {
    cnt >>= 1;
    if (l & m) base += cnt;
  }
  return base;
}

// Sum two per-thread doubles over the workgroup; thread 0 writes them.
__device__ __forceinline__ void wg_sum2(double a, double b, double* out) {
  __shared__ double sm[2][WG / WAVE];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    sm[0][w] = a;
    sm[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int k = 0; k < WG / WAVE; k++) {
      sa += sm[0][k];
      sb += sm[1][k];
    }
    out[0] = sa;
    out[1] = sb;
  }
}

// Complex (or real) multiply-accumulate on interleaved scalars:
// acc[e] += h * x[e] for the elements of a packet.
template <bool CX, int NSC, typename A>
__device__ __forceinline__ void cmac(A* acc, A hr, A hi, const A* x) {
  if constexpr (CX) {
#pragma unroll
    for (int e = 0; e < NSC; e += 2) {
      acc[e] = fma(hr, x[e], acc[e]);
      acc[e] = fma(-hi, x[e + 1], acc[e]);
      acc[e + 1] = fma(hr, x[e + 1], acc[e + 1]);
      acc[e + 1] = fma(hi, x[e], acc[e + 1]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < NSC; e++) acc[e] = fma(hr, x[e], acc[e]);
  }
}

// Static geometry of one kernel instantiation.
template <typename S, bool CX, int B, int L, int KQ>
struct Geo {
  static constexpr int NS = Vec16<S>::N;        // scalars per packet
  static constexpr int HS = CX ? 2 : 1;          // scalars per element
  static constexpr int NE = NS / HS;             // elements (columns) per packet
  static constexpr int G = WAVE / L;             // groups per wave
  static constexpr int GWG = (WG / WAVE) * G;    // groups (block rows) per workgroup pass
  static constexpr int BBS = B * B * HS;         // scalars per H block
  static constexpr int DE = B * B;               // elements per D block
  static constexpr int DOWN = (DE + L - 1) / L;  // D elements owned per lane
  // Neighbour blocks gathered per batch for NX gathered vectors: a batch holds
  // U * (NX*B*KQ*NS + B*B*HS) values; a budget of about 48 (up to 4 blocks)
  // keeps the registers of the autodiff steps in check while giving several
  // blocks of loads in flight; the forward steps, with fewer live values,
  // take 64 (up to 8 blocks: a 2-D lattice row in one batch on 64 lanes,
  // FP64 forward step 80 -> 79 us at C2; the autodiff step lost a wave per
  // SIMD and 4 us with it).
  template <int NX, int BUDGET = 48>
  static constexpr int batch() {
    constexpr int per = NX * B * KQ * NS + B * B * HS;
    constexpr int u = BUDGET / per;
    constexpr int cap = BUDGET > 48 ? 8 : 4;
    return u < 1 ? 1 : (u > cap ? cap : u);
  }
};

// Rows of one workgroup.  Workgroups reach the 8 XCDs round-robin
// (blockIdx % 8); XCD x owns one contiguous region of block rows, the union
// of the chunks of its workgroups, and its workgroups sweep that region
// together: workgroup r of the XCD takes passes r, r + n_x, r + 2 n_x, ... of
// `step` = (block rows per workgroup pass) rows.  The region the XCD touches at
// a time is then about n_x * step rows wide, so the neighbour rows a gather
// needs (a lattice's next rows) are still in that XCD's L2.
struct RowRange {
  int r0, r1, stride;
};

// A workgroup's index k among the n workgroups that share one sweep: the
// launch's own (blockIdx, gridDim), or a sub-range of a batched launch.
struct WgId {
  int k, n;
};

__device__ __forceinline__ WgId wg_self() { return WgId{(int)blockIdx.x, (int)gridDim.x}; }

//
// chunk = -s (s >= 1) selects the other sweep: all workgroups advance one
// front over all rows, the rows cut into segments of s passes (s * step
// rows) dealt to the XCDs in turn (segment j to XCD j % 8); the workgroups of
// an XCD take its passes round-robin.  With n a multiple of 8 s, workgroup
// (x = k % 8, r = k / 8) starts at ((r / s) * 8 + x) * s * step + (r % s) * step
// and advances by n * step.  The rows in flight across the chip span one
// narrow window, whose far neighbours (a 3-D lattice's next plane) stay in the
// shared last-level cache, a neighbour d rows away is on the same XCD when
// d / (s * step) is a multiple of 8, and only rows at segment edges have a
// near neighbour on another XCD (s = 1: workgroup k takes passes k, k + n, ...).
__device__ __forceinline__ RowRange wg_rows(int nb, int chunk, int step, WgId wid) {
  const int k = wid.k, n = wid.n;
  if (chunk < 0) {
    const int sp = -chunk, x = k & 7, r = k >> 3;
    RowRange rr;
    rr.r0 = ((r / sp) * 8 + x) * sp * step + (r % sp) * step;
    rr.r1 = nb;
    rr.stride = n * step;
    return rr;
  }
  const int x = k & 7, r = k >> 3;
  const int per = n >> 3, extra = n & 7;
  const int nx = per + (x < extra ? 1 : 0);        // workgroups of XCD x
  const int first = x * per + (x < extra ? x : extra);  // its first chunk
  RowRange rr;
  const int start = first * chunk;
  rr.r1 = min(nb, (first + nx) * chunk);
  rr.r0 = start + r * step;
  rr.stride = nx * step;
  return rr;
}

// The block row at sweep position i and its block range: row i and rp[i],
// rp[i+1], or with NT_ORDER the descriptor rp[3i .. 3i+2] (empty past r1).
// The row loops load the next position's entry before working on the current
// row.
struct RowPtr {
  int row, p0, p1;
};

__device__ __forceinline__ RowPtr row_ptr(const int* __restrict__ rp, int i, int r1, int nt) {
  if (i >= r1) return RowPtr{i, 0, 0};
  if (nt & NT_ORDER) return RowPtr{rp[3 * i], rp[3 * i + 1], rp[3 * i + 2]};
  return RowPtr{i, rp[i], rp[i + 1]};
}

// A group's first row; with one group per wave (L = 64) the index is made
// wave-uniform, so that the row's offsets, column indices and H values are
// scalar loads.
template <int L>
__device__ __forceinline__ int first_row(int i) {
  if constexpr (L == WAVE) return __builtin_amdgcn_readfirstlane(i);
  return i;
}

// Block row i's gather: acc[x][k][bi] += sum_p Hs_p(bi, bj) X_x[B j_p + bj]
// over the blocks p0..p1 of the row, for NX vectors sharing the pattern, then
// blk(p, j_p, xv, pv) with the gathered packets xv[x][bj][k][.] of block p and
// pv = pre(p), a value pre() loads for block p while block p's packets are in
// flight (the D entries a fused outer product updates), so that its latency
// overlaps the gather's.  qo is the scalar offset of this lane's first packet
// in a row.  The column indices of the next batch are loaded with the current
// batch's packets.
// With OWN, the packets of X[0] in the row's own block (column iown) are taken
// from `own` (already in registers) instead of being loaded again.
struct NoPre {};

template <typename S, typename A, bool CX, int B, int L, int KQ, int NX, bool OWN, class F,
          class P, int BUDGET = 48>
__device__ __forceinline__ void gather_row_pre(const int* __restrict__ ci,
                                               const S* __restrict__ hv, const S* const (&X)[NX],
                                               size_t rs, size_t qo, int p0, int p1,
                                               A (&acc)[NX][KQ][B][Vec16<S>::N], F&& blk,
                                               P&& pre, int iown,
                                               const A (&own)[KQ][B][Vec16<S>::N]) {
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS, HS = GE::HS, BBS = GE::BBS;
  constexpr int U = GE::template batch<NX, BUDGET>();
  using PV = decltype(pre(0));
  int jn[U];
#pragma unroll
  for (int u = 0; u < U; u++) jn[u] = (p0 + u < p1) ? ci[p0 + u] : -1;
  for (int p = p0; p < p1; p += U) {
    int j[U];
#pragma unroll
    for (int u = 0; u < U; u++) j[u] = jn[u];
    A xv[U][NX][B][KQ][NS];
    A hh[U][BBS];
    PV pv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (j[u] >= 0) {
#pragma unroll
        for (int x = 0; x < NX; x++) {
          if (OWN && x == 0 && j[u] == iown) {
#pragma unroll
            for (int bj = 0; bj < B; bj++)
#pragma unroll
              for (int k = 0; k < KQ; k++)
#pragma unroll
                for (int e = 0; e < NS; e++) xv[u][0][bj][k][e] = own[k][bj][e];
            continue;
          }
#pragma unroll
          for (int bj = 0; bj < B; bj++)
#pragma unroll
            for (int k = 0; k < KQ; k++)
              ldv<S, NS>(X[x] + (size_t)(B * j[u] + bj) * rs + qo + (size_t)k * L * NS,
                         xv[u][x][bj][k]);
        }
#pragma unroll
        for (int t = 0; t < BBS; t++) hh[u][t] = (A)hv[(size_t)(p + u) * BBS + t];
        pv[u] = pre(p + u);
      } else {
        // defined zeros past the row's end keep the batch's registers from
        // staying live across the (skipped) arithmetic
#pragma unroll
        for (int x = 0; x < NX; x++)
#pragma unroll
          for (int bj = 0; bj < B; bj++)
#pragma unroll
            for (int k = 0; k < KQ; k++)
#pragma unroll
              for (int e = 0; e < NS; e++) xv[u][x][bj][k][e] = A(0);
#pragma unroll
        for (int t = 0; t < BBS; t++) hh[u][t] = A(0);
        pv[u] = PV{};
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) jn[u] = (p + U + u < p1) ? ci[p + U + u] : -1;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (j[u] < 0) continue;
#pragma unroll
      for (int bj = 0; bj < B; bj++)
#pragma unroll
        for (int x = 0; x < NX; x++)
#pragma unroll
          for (int k = 0; k < KQ; k++)
#pragma unroll
            for (int bi = 0; bi < B; bi++)
              cmac<CX, NS>(acc[x][k][bi], hh[u][(bj * B + bi) * HS],
                           CX ? hh[u][(bj * B + bi) * HS + HS - 1] : A(0), xv[u][x][bj][k]);
      blk(p + u, j[u], xv[u], pv[u]);
    }
  }
}

// The forward steps' gather (no prefetched values, the larger batch budget).
template <typename S, typename A, bool CX, int B, int L, int KQ, int NX, class F>
__device__ __forceinline__ void gather_row(const int* __restrict__ ci, const S* __restrict__ hv,
                                           const S* const (&X)[NX], size_t rs, size_t qo, int p0,
                                           int p1, A (&acc)[NX][KQ][B][Vec16<S>::N], F&& blk) {
  const A none[KQ][B][Vec16<S>::N] = {};
  auto nopre = [](int) { return NoPre{}; };
  auto blk4 = [&](int p, int jb, auto& xv, NoPre) { blk(p, jb, xv); };
  gather_row_pre<S, A, CX, B, L, KQ, NX, false, decltype(blk4)&, decltype(nopre)&, 64>(
      ci, hv, X, rs, qo, p0, p1, acc, blk4, nopre, -1, none);
}

// Own rows of block row i: v[k][bi] = X[B i + bi] packets of this lane.
template <typename S, typename A, int B, int L, int KQ>
__device__ __forceinline__ void load_rows(const S* __restrict__ X, int i, size_t rs, size_t qo,
                                          A (&v)[KQ][B][Vec16<S>::N]) {
  constexpr int NS = Vec16<S>::N;
#pragma unroll
  for (int bi = 0; bi < B; bi++)
#pragma unroll
    for (int k = 0; k < KQ; k++)
      ldv<S, NS>(X + (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS, v[k][bi]);
}

// The same with the nontemporal hint chosen once for all packets, so that
// every load of the row is issued before the first use.
template <typename S, typename A, int B, int L, int KQ>
__device__ __forceinline__ void load_rows_nt(const S* __restrict__ X, int i, size_t rs, size_t qo,
                                             A (&v)[KQ][B][Vec16<S>::N], bool nt) {
  constexpr int NS = Vec16<S>::N;
  if (nt) {
#pragma unroll
    for (int bi = 0; bi < B; bi++)
#pragma unroll
      for (int k = 0; k < KQ; k++)
        ldv_nt<S, NS>(X + (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS, v[k][bi], true);
  } else {
    load_rows<S, A, B, L, KQ>(X, i, rs, qo, v);
  }
}

// Optional source term cr * R_i of a Chebyshev step (the Clenshaw form of
// stoch_matrix): R dense (rd, the N x Sp probe block) or compact (one nonzero
// per orbital row: column rcol[r] (-1 for none) with value rval[r]).
template <typename S>
struct RSource {
  const S* rd;
  const int* rcol;
  const S* rval;
  double cr;
};

// y[k][bi] += cr * R(B i + bi, packet ql + k L) for all packets of a lane in
// block row i (ql = q0 + l), every load issued before the first use.
template <typename S, bool CX, typename A, int B, int L, int KQ>
__device__ __forceinline__ void add_rsource_rows(const RSource<S>& rs_, int i, size_t rs,
                                                 size_t qo, int ql,
                                                 A (&y)[KQ][B][Vec16<S>::N]) {
  constexpr int NS = Vec16<S>::N, HS = CX ? 2 : 1, NE = NS / HS;
  const A cr = (A)rs_.cr;
  if (rs_.rd != nullptr) {
    A r[KQ][B][NS];
    load_rows<S, A, B, L, KQ>(rs_.rd, i, rs, qo, r);
#pragma unroll
    for (int k = 0; k < KQ; k++)
#pragma unroll
      for (int bi = 0; bi < B; bi++)
#pragma unroll
        for (int e = 0; e < NS; e++) y[k][bi][e] = fma(cr, r[k][bi][e], y[k][bi][e]);
  } else {
    int col[B];
    A val[B][HS];
#pragma unroll
    for (int bi = 0; bi < B; bi++) {
      const size_t row = (size_t)B * i + bi;
      col[bi] = rs_.rcol[row];
#pragma unroll
      for (int h = 0; h < HS; h++) val[bi][h] = (A)rs_.rval[row * HS + h];
    }
#pragma unroll
    for (int k = 0; k < KQ; k++)
#pragma unroll
      for (int bi = 0; bi < B; bi++) {
        const int c = col[bi] - (ql + k * L) * NE;  // element of this packet holding the nonzero
        // selects rather than a branch on c, which the compiler would turn
        // into a dynamic index into y (and y into scratch memory)
#pragma unroll
        for (int e = 0; e < NE; e++)
#pragma unroll
          for (int h = 0; h < HS; h++)
            y[k][bi][e * HS + h] =
                fma(cr, e == c ? val[bi][h] : A(0), y[k][bi][e * HS + h]);
      }
  }
}

// ---------------------------------------------------------------------------
// Moments step: y = cx * (Hs x) + cz * z  (z may be null) [+ cr * R], and the
// two moment dots sum |x_i|^2 and Re<y_i, x_i> over this workgroup's rows
// (i = own rows; skipped when partial is null).  This is the forward
// recursion alpha_{m+1} = 2 Hs alpha_m - alpha_{m-1} of Engine_CPU::moments
// (engine_cpu.cpp:63-84) fused with both cdot's; with the R term it is one
// step b_k = c_k R + 2 Hs b_{k+1} - b_{k+2} of the Clenshaw recurrence.
// ---------------------------------------------------------------------------
template <typename S, bool CX, int B, int L, int KQ>
__device__ __forceinline__ void step_moments_dev(
    WgId wid,
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  using A = S;
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane % L;
  const int g = (threadIdx.x / WAVE) * GE::G + lane / L;
  const RowRange rr = wg_rows(nb, chunk, GE::GWG, wid);
  const size_t rs = (size_t)qp * NS;  // scalars per orbital row
  const A acx = (A)cx, acz = (A)cz;
  double dmm = 0.0, dm1 = 0.0;
  const S* const X[1] = {x};

  const int qw = (nt & NT_SLICE) ? L * KQ : qp;  // columns per sweep over the rows
  for (int qs = 0; qs < qp; qs += qw) {
  RowPtr nxt = row_ptr(rp, first_row<L>(rr.r0 + g), rr.r1, nt);
  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    int i = base + g;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const RowPtr cur = nxt;
    nxt = row_ptr(rp, i + rr.stride, rr.r1, nt);
    if (i >= rr.r1) continue;
    i = cur.row;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const int p0 = cur.p0, p1 = cur.p1;
    for (int q0 = qs; q0 < qs + qw; q0 += L * KQ) {
      const size_t qo = (size_t)(q0 + l) * NS;
      A xd[KQ][B][NS];
      A acc[1][KQ][B][NS];
      // the own rows of x (for the dots) arrive with the diagonal block of
      // the gather; every row has one (set_H requires it)
#pragma unroll
      for (int k = 0; k < KQ; k++)
#pragma unroll
        for (int bi = 0; bi < B; bi++)
#pragma unroll
          for (int e = 0; e < NS; e++) {
            acc[0][k][bi][e] = A(0);
            xd[k][bi][e] = A(0);
          }
      gather_row<S, A, CX, B, L, KQ, 1>(ci, hv, X, rs, qo, p0, p1, acc,
                                         [&](int, int jb, auto& xv) {
                                           if (jb != i) return;
#pragma unroll
                                           for (int k = 0; k < KQ; k++)
#pragma unroll
                                             for (int bi = 0; bi < B; bi++)
#pragma unroll
                                               for (int e = 0; e < NS; e++)
                                                 xd[k][bi][e] = xv[0][bi][k][e];
                                         });
      A out[KQ][B][NS];
      if (z != nullptr) {
        load_rows_nt<S, A, B, L, KQ>(z, i, rs, qo, out, nt & NT_LOAD);
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++)
              out[k][bi][e] = fma(acx, acc[0][k][bi][e], acz * out[k][bi][e]);
      } else {
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++) out[k][bi][e] = acx * acc[0][k][bi][e];
      }
      if (rsrc.cr != 0.0) add_rsource_rows<S, CX, A, B, L, KQ>(rsrc, i, rs, qo, q0 + l, out);
#pragma unroll
      for (int bi = 0; bi < B; bi++) {
#pragma unroll
        for (int k = 0; k < KQ; k++) {
          const size_t off = (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS;
          stv_nt<S, NS>(y + off, out[k][bi], nt & NT_STORE);
#pragma unroll
          for (int e = 0; e < NS; e++) {
            const double xe = (double)xd[k][bi][e];
            dmm = fma(xe, xe, dmm);
            dm1 = fma((double)out[k][bi][e], xe, dm1);
          }
        }
      }
    }
  }
  }
  if (partial != nullptr) wg_sum2(dmm, dm1, partial + 2 * (size_t)wid.k);
}

template <typename S, bool CX, int B, int L, int KQ>
__global__ __launch_bounds__(WG) void k_step_moments(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  step_moments_dev<S, CX, B, L, KQ>(wg_self(), rp, ci, hv, x, z, y, nb, qp, chunk, cx, cz, partial, nt, rsrc);
}

// ---------------------------------------------------------------------------
// Stochastic-matrix step: y = cx * (Hs x) + cz * z, and xi (FP64) updated as
// xi = c0 * x_i + c1 * y  (first step, x = R) or xi += c1 * y.  This is the
// loop body of Engine_CPU::stoch_matrix (engine_cpu.cpp:229-244).
// ---------------------------------------------------------------------------
template <typename S, bool CX, int B, int L, int KQ>
__global__ __launch_bounds__(WG) void k_step_xi(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y,
    double* __restrict__ xi, int nb, int qp, int chunk, double cx, double cz, double c0,
    double c1, int first, int nt) {
  using A = S;
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane % L;
  const int g = (threadIdx.x / WAVE) * GE::G + lane / L;
  const RowRange rr = wg_rows(nb, chunk, GE::GWG, wg_self());
  const size_t rs = (size_t)qp * NS;
  const A acx = (A)cx, acz = (A)cz;
  const S* const X[1] = {x};

  const int qw = (nt & NT_SLICE) ? L * KQ : qp;  // columns per sweep over the rows
  for (int qs = 0; qs < qp; qs += qw) {
  RowPtr nxt = row_ptr(rp, first_row<L>(rr.r0 + g), rr.r1, nt);
  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    int i = base + g;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const RowPtr cur = nxt;
    nxt = row_ptr(rp, i + rr.stride, rr.r1, nt);
    if (i >= rr.r1) continue;
    i = cur.row;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const int p0 = cur.p0, p1 = cur.p1;
    for (int q0 = qs; q0 < qs + qw; q0 += L * KQ) {
      const size_t qo = (size_t)(q0 + l) * NS;
      A acc[1][KQ][B][NS];
#pragma unroll
      for (int k = 0; k < KQ; k++)
#pragma unroll
        for (int bi = 0; bi < B; bi++)
#pragma unroll
          for (int e = 0; e < NS; e++) acc[0][k][bi][e] = A(0);
      gather_row<S, A, CX, B, L, KQ, 1>(ci, hv, X, rs, qo, p0, p1, acc, [](int, int, auto&) {});
      // own rows of z and xi (x at the first step), all packets in one batch
      A zv[KQ][B][NS];
      double xa[KQ][B][NS];
      if (z != nullptr) {
        load_rows_nt<S, A, B, L, KQ>(z, i, rs, qo, zv, nt & NT_LOAD);
      } else {
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++) zv[k][bi][e] = A(0);
      }
      A xo[KQ][B][NS];
      if (first) {
        load_rows<S, A, B, L, KQ>(x, i, rs, qo, xo);
      } else {
        const bool ntl = nt & NT_LOAD;
#pragma unroll
        for (int bi = 0; bi < B; bi++)
#pragma unroll
          for (int k = 0; k < KQ; k++)
            ldv_nt<double, NS>(xi + (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS,
                               xa[k][bi], ntl);
      }
#pragma unroll
      for (int bi = 0; bi < B; bi++) {
#pragma unroll
        for (int k = 0; k < KQ; k++) {
          const size_t off = (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS;
          A out[NS];
#pragma unroll
          for (int e = 0; e < NS; e++) out[e] = fma(acx, acc[0][k][bi][e], acz * zv[k][bi][e]);
          stv_nt<S, NS>(y + off, out, nt & NT_STORE);
          double xn[NS];
          if (first) {
#pragma unroll
            for (int e = 0; e < NS; e++)
              xn[e] = fma(c0, (double)xo[k][bi][e], c1 * (double)out[e]);
          } else {
#pragma unroll
            for (int e = 0; e < NS; e++) xn[e] = fma(c1, (double)out[e], xa[k][bi][e]);
          }
          stv_nt<double, NS>(xi + off, xn, nt & NT_STORE);
        }
      }
    }
  }
  }
}

// The D entries one lane updates in a fused outer-product block p (the
// layouts of the D updates in the autodiff steps): with NV = B*B*HS <= 16
// scalars per block, the CNT entries at rs_base after the reduce-scatter;
// otherwise own element o*L + l of the block (o < DOWN), HS scalars each.
template <int N>
struct DVals {
  double v[N];
};

// Entry t of the prefetched D values, or *d when nothing was prefetched.
template <int N>
__device__ __forceinline__ double dpre_get(const DVals<N>& dp, int t, const double* d) {
  return dp.v[t];
}
__device__ __forceinline__ double dpre_get(const NoPre&, int, const double* d) { return *d; }

template <typename S, bool CX, int B, int L, int KQ, bool FUSED_D>
__device__ __forceinline__ auto load_dpre(const double* __restrict__ dacc, int p, int l) {
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int HS = GE::HS;
  constexpr int NV = GE::DE * HS;
  if constexpr (!FUSED_D || (NV > 16 && GE::DOWN * HS > 8)) {
    // large blocks (b = 8 on the vector units): the entries would hold too
    // many registers across the gather; the update reads them itself
    return NoPre{};
  } else if constexpr (NV <= 16) {
    constexpr int CNT = NV >= L ? NV / L : 1;
    DVals<CNT> d{};
    if (NV >= L || l < NV) {
      const double* src = dacc + (size_t)p * NV + rs_base<L, NV>(l);
#pragma unroll
      for (int t = 0; t < CNT; t++) d.v[t] = src[t];
    }
    return d;
  } else {
    DVals<GE::DOWN * HS> d{};
#pragma unroll
    for (int o = 0; o < GE::DOWN; o++) {
      const int de = o * L + l;
      if (de < GE::DE) {
#pragma unroll
        for (int c = 0; c < HS; c++) d.v[o * HS + c] = dacc[((size_t)p * GE::DE + de) * HS + c];
      }
    }
    return d;
  }
}

// ---------------------------------------------------------------------------
// Autodiff backward step m (Engine_CPU::autodiff_matrix, engine_cpu.cpp:278-294):
//   a_out = alpha_{m-1} = 2 Hs alpha_m - alpha_{m+1}
//   b_out = beta_m      = 4 c_{2m} alpha_m + 2 c_{2m+1} alpha_{m+1}
//                         + 2 Hs beta_{m+1} - beta_{m+2} + w alpha_{m-1}
//   D    += 2 alpha_m beta_{m+1}^+        (FUSED_D: on H's own pattern)
// beta_{m+1} may be virtual: bm1 == null means beta_{m+1} = sb * alpha_m
// (the step m = M/2-1, where beta_{M/2} = 2 c_{M-1} alpha_{M/2-1}); bm2 == null
// means beta_{m+2} = sb2 * alpha_{m+1} (sb2 = 0 at the first step).
// ---------------------------------------------------------------------------
template <typename S, bool CX, int B, int L, int KQ, bool FUSED_D>
__device__ __forceinline__ void step_autodiff_dev(
    WgId wid,
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  using A = S;
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS;
  constexpr int HS = GE::HS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane % L;
  const int g = (threadIdx.x / WAVE) * GE::G + lane / L;
  const RowRange rr = wg_rows(nb, chunk, GE::GWG, wid);
  const size_t rs = (size_t)qp * NS;
  const bool vb1 = (bm1 == nullptr);
  const double hb1 = vb1 ? sb : 1.0;  // scale of the beta_{m+1} gather
  // a virtual beta_{m+1} gathers alpha_m a second time (one launch per call)
  const S* const X[2] = {am, vb1 ? am : bm1};
  const A a2 = (A)2, ab2 = (A)(2.0 * hb1), ac4 = (A)c4, ac2 = (A)c2, aw = (A)w, asb2 = (A)sb2;
  const double fD = dscale * hb1;
  using DPre = decltype(load_dpre<S, CX, B, L, KQ, FUSED_D>(dacc, 0, 0));

  const int qw = (nt & NT_SLICE) ? L * KQ : qp;  // columns per sweep over the rows
  for (int qs = 0; qs < qp; qs += qw) {
  RowPtr nxt = row_ptr(rp, first_row<L>(rr.r0 + g), rr.r1, nt);
  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    int i = base + g;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const RowPtr cur = nxt;
    nxt = row_ptr(rp, i + rr.stride, rr.r1, nt);
    if (i >= rr.r1) continue;
    i = cur.row;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const int p0 = cur.p0, p1 = cur.p1;
    for (int q0 = qs; q0 < qs + qw; q0 += L * KQ) {
      const size_t qo = (size_t)(q0 + l) * NS;
      A ai[KQ][B][NS];
      A acc[2][KQ][B][NS];
      load_rows<S, A, B, L, KQ>(am, i, rs, qo, ai);
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++) acc[x][k][bi][e] = A(0);
      gather_row_pre<S, A, CX, B, L, KQ, 2, true>(
          ci, hv, X, rs, qo, p0, p1, acc, [&](int p, int, auto& xv, const DPre& dp) {
            constexpr int NV = GE::DE * HS;  // scalars of one D block
            if constexpr (FUSED_D && NV <= 16) {
              // D_p(bi, bj) += fD * sum_s alpha_m[B i + bi, s] conj(beta_{m+1}[B j + bj, s]):
              // per-lane partial sums of the whole block, then one reduce-scatter
              // over the group in the storage precision, added to D in FP64
              A v[NV];
#pragma unroll
              for (int bj = 0; bj < B; bj++)
#pragma unroll
                for (int bi = 0; bi < B; bi++) {
                  A re = A(0), im = A(0);
#pragma unroll
                  for (int k = 0; k < KQ; k++)
#pragma unroll
                    for (int e = 0; e < NS; e += HS) {
                      const A ar = ai[k][bi][e], br = xv[1][bj][k][e];
                      re = fma(ar, br, re);
                      if constexpr (CX) {
                        const A aim = ai[k][bi][e + 1], bim = xv[1][bj][k][e + 1];
                        re = fma(aim, bim, re);
                        im = fma(aim, br, im);
                        im = fma(-ar, bim, im);
                      }
                    }
                  const int de = bj * B + bi;  // column-major element index in the block
                  v[de * HS] = re;
                  if constexpr (CX) v[de * HS + 1] = im;
                }
              reduce_scatter<L, NV>(v, l);
              constexpr int CNT = NV >= L ? NV / L : 1;
              if (NV >= L || l < NV) {
                const int base = rs_base<L, NV>(l);
                double* d = dacc + (size_t)p * NV + base;
#pragma unroll
                for (int t = 0; t < CNT; t++) d[t] = fma(fD, (double)v[t], dpre_get(dp, t, d + t));
              }
            } else if constexpr (FUSED_D) {
              double own[GE::DOWN][HS];
#pragma unroll
              for (int bj = 0; bj < B; bj++)
#pragma unroll
                for (int bi = 0; bi < B; bi++) {
                  A re = A(0), im = A(0);
#pragma unroll
                  for (int k = 0; k < KQ; k++)
#pragma unroll
                    for (int e = 0; e < NS; e += HS) {
                      const A ar = ai[k][bi][e], br = xv[1][bj][k][e];
                      re = fma(ar, br, re);
                      if constexpr (CX) {
                        const A aim = ai[k][bi][e + 1], bim = xv[1][bj][k][e + 1];
                        re = fma(aim, bim, re);
                        im = fma(aim, br, im);
                        im = fma(-ar, bim, im);
                      }
                    }
                  const double sre = group_sum<L>((double)re);
                  const double sim = CX ? group_sum<L>((double)im) : 0.0;
                  const int de = bj * B + bi;  // column-major element index in the block
                  if (l == de % L) {
                    own[de / L][0] = sre;
                    if constexpr (CX) own[de / L][HS - 1] = sim;
                  }
                }
#pragma unroll
              for (int o = 0; o < GE::DOWN; o++) {
                const int de = o * L + l;
                if (de < GE::DE) {
                  double* d = dacc + ((size_t)p * GE::DE + de) * HS;
#pragma unroll
                  for (int c = 0; c < HS; c++) d[c] = fma(fD, own[o][c], dpre_get(dp, o * HS + c, d + c));
                }
              }
            }
          },
          [&](int p) { return load_dpre<S, CX, B, L, KQ, FUSED_D>(dacc, p, l); }, i, ai);
      // own rows of alpha_{m+1} and beta_{m+2}, all packets in one batch
      A a1[KQ][B][NS], b2v[KQ][B][NS];
      load_rows_nt<S, A, B, L, KQ>(am1, i, rs, qo, a1, nt & NT_LOAD);
      if (bm2 != nullptr) {
        load_rows_nt<S, A, B, L, KQ>(bm2, i, rs, qo, b2v, nt & NT_LOAD);
      } else {
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++) b2v[k][bi][e] = asb2 * a1[k][bi][e];
      }
#pragma unroll
      for (int bi = 0; bi < B; bi++) {
#pragma unroll
        for (int k = 0; k < KQ; k++) {
          const size_t off = (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS;
          A an[NS], bo[NS];
#pragma unroll
          for (int e = 0; e < NS; e++) an[e] = fma(a2, acc[0][k][bi][e], -a1[k][bi][e]);
          stv_nt<S, NS>(a_out + off, an, nt & NT_STORE);
#pragma unroll
          for (int e = 0; e < NS; e++) {
            A v = fma(ab2, acc[1][k][bi][e], -b2v[k][bi][e]);
            v = fma(ac4, ai[k][bi][e], v);
            v = fma(ac2, a1[k][bi][e], v);
            v = fma(aw, an[e], v);
            bo[e] = v;
          }
          stv_nt<S, NS>(b_out + off, bo, nt & NT_STORE);
        }
      }
    }
  }
  }
}

// ---------------------------------------------------------------------------
// Backward step with the forward vectors kept: when moments(M) left every
// alpha_k in HBM (a table of M/2 vectors), alpha_{m-1} is read instead of
// recomputed, so the step gathers beta_{m+1} only:
//   b_out = beta_m = 4 c_{2m} alpha_m + 2 c_{2m+1} alpha_{m+1}
//                    + 2 Hs beta_{m+1} - beta_{m+2} + w alpha_{m-1}
//   D    += 2 alpha_m beta_{m+1}^+        (FUSED_D: on H's own pattern)
// with am, am1, amm = alpha_m, alpha_{m+1}, alpha_{m-1} (own rows only).
// The same algorithmic bytes as k_step_autodiff (five vectors read, one
// written) but one gathered vector instead of two, so half the neighbour
// rows compete for the L2.  bm1 == null: beta_{m+1} = sb * alpha_m (its
// gather reads alpha_m); bm2 == null: beta_{m+2} = sb2 * alpha_{m+1}.
// ---------------------------------------------------------------------------
template <typename S, bool CX, int B, int L, int KQ, bool FUSED_D>
__device__ __forceinline__ void step_autodiff_stored_dev(
    WgId wid,
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, const S* __restrict__ amm, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  using A = S;
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS;
  constexpr int HS = GE::HS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane % L;
  const int g = (threadIdx.x / WAVE) * GE::G + lane / L;
  const RowRange rr = wg_rows(nb, chunk, GE::GWG, wid);
  const size_t rs = (size_t)qp * NS;
  const bool vb1 = (bm1 == nullptr);
  const double hb1 = vb1 ? sb : 1.0;  // scale of the beta_{m+1} gather
  const S* const X[1] = {vb1 ? am : bm1};
  const A ab2 = (A)(2.0 * hb1), ac4 = (A)c4, ac2 = (A)c2, aw = (A)w, asb2 = (A)sb2;
  const double fD = dscale * hb1;
  using DPre = decltype(load_dpre<S, CX, B, L, KQ, FUSED_D>(dacc, 0, 0));
  const A none[KQ][B][NS] = {};

  const int qw = (nt & NT_SLICE) ? L * KQ : qp;  // columns per sweep over the rows
  for (int qs = 0; qs < qp; qs += qw) {
  RowPtr nxt = row_ptr(rp, first_row<L>(rr.r0 + g), rr.r1, nt);
  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    int i = base + g;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const RowPtr cur = nxt;
    nxt = row_ptr(rp, i + rr.stride, rr.r1, nt);
    if (i >= rr.r1) continue;
    i = cur.row;
    if constexpr (L == WAVE) i = __builtin_amdgcn_readfirstlane(i);
    const int p0 = cur.p0, p1 = cur.p1;
    for (int q0 = qs; q0 < qs + qw; q0 += L * KQ) {
      const size_t qo = (size_t)(q0 + l) * NS;
      A ai[KQ][B][NS];
      A acc[1][KQ][B][NS];
      load_rows<S, A, B, L, KQ>(am, i, rs, qo, ai);
#pragma unroll
      for (int k = 0; k < KQ; k++)
#pragma unroll
        for (int bi = 0; bi < B; bi++)
#pragma unroll
          for (int e = 0; e < NS; e++) acc[0][k][bi][e] = A(0);
      gather_row_pre<S, A, CX, B, L, KQ, 1, false>(ci, hv, X, rs, qo, p0, p1, acc, [&](int p, int, auto& xv, const DPre& dp) {
        constexpr int NV = GE::DE * HS;  // scalars of one D block
        if constexpr (!FUSED_D) {
          (void)p;
          (void)xv;
        } else if constexpr (NV <= 16) {
          A v[NV];  // reduced in the storage precision, added to D in FP64
#pragma unroll
          for (int bj = 0; bj < B; bj++)
#pragma unroll
            for (int bi = 0; bi < B; bi++) {
              A re = A(0), im = A(0);
#pragma unroll
              for (int k = 0; k < KQ; k++)
#pragma unroll
                for (int e = 0; e < NS; e += HS) {
                  const A ar = ai[k][bi][e], br = xv[0][bj][k][e];
                  re = fma(ar, br, re);
                  if constexpr (CX) {
                    const A aim = ai[k][bi][e + 1], bim = xv[0][bj][k][e + 1];
                    re = fma(aim, bim, re);
                    im = fma(aim, br, im);
                    im = fma(-ar, bim, im);
                  }
                }
              const int de = bj * B + bi;  // column-major element index in the block
              v[de * HS] = re;
              if constexpr (CX) v[de * HS + 1] = im;
            }
          reduce_scatter<L, NV>(v, l);
          constexpr int CNT = NV >= L ? NV / L : 1;
          if (NV >= L || l < NV) {
            const int base = rs_base<L, NV>(l);
            double* d = dacc + (size_t)p * NV + base;
#pragma unroll
            for (int t = 0; t < CNT; t++) d[t] = fma(fD, (double)v[t], dpre_get(dp, t, d + t));
          }
        } else {
          double own[GE::DOWN][HS];
#pragma unroll
          for (int bj = 0; bj < B; bj++)
#pragma unroll
            for (int bi = 0; bi < B; bi++) {
              A re = A(0), im = A(0);
#pragma unroll
              for (int k = 0; k < KQ; k++)
#pragma unroll
                for (int e = 0; e < NS; e += HS) {
                  const A ar = ai[k][bi][e], br = xv[0][bj][k][e];
                  re = fma(ar, br, re);
                  if constexpr (CX) {
                    const A aim = ai[k][bi][e + 1], bim = xv[0][bj][k][e + 1];
                    re = fma(aim, bim, re);
                    im = fma(aim, br, im);
                    im = fma(-ar, bim, im);
                  }
                }
              const double sre = group_sum<L>((double)re);
              const double sim = CX ? group_sum<L>((double)im) : 0.0;
              const int de = bj * B + bi;
              if (l == de % L) {
                own[de / L][0] = sre;
                if constexpr (CX) own[de / L][HS - 1] = sim;
              }
            }
#pragma unroll
          for (int o = 0; o < GE::DOWN; o++) {
            const int de = o * L + l;
            if (de < GE::DE) {
              double* d = dacc + ((size_t)p * GE::DE + de) * HS;
#pragma unroll
              for (int c = 0; c < HS; c++) d[c] = fma(fD, own[o][c], dpre_get(dp, o * HS + c, d + c));
            }
          }
        }
      }, [&](int p) { return load_dpre<S, CX, B, L, KQ, FUSED_D>(dacc, p, l); }, -1, none);
      // own rows of alpha_{m+1}, alpha_{m-1} and beta_{m+2}, all packets in one batch
      A a1[KQ][B][NS], an[KQ][B][NS], b2v[KQ][B][NS];
      load_rows_nt<S, A, B, L, KQ>(am1, i, rs, qo, a1, nt & NT_LOAD);
      load_rows_nt<S, A, B, L, KQ>(amm, i, rs, qo, an, nt & NT_LOAD);
      if (bm2 != nullptr) {
        load_rows_nt<S, A, B, L, KQ>(bm2, i, rs, qo, b2v, nt & NT_LOAD);
      } else {
#pragma unroll
        for (int k = 0; k < KQ; k++)
#pragma unroll
          for (int bi = 0; bi < B; bi++)
#pragma unroll
            for (int e = 0; e < NS; e++) b2v[k][bi][e] = asb2 * a1[k][bi][e];
      }
#pragma unroll
      for (int bi = 0; bi < B; bi++) {
#pragma unroll
        for (int k = 0; k < KQ; k++) {
          const size_t off = (size_t)(B * i + bi) * rs + qo + (size_t)k * L * NS;
          A bo[NS];
#pragma unroll
          for (int e = 0; e < NS; e++) {
            A v = fma(ab2, acc[0][k][bi][e], -b2v[k][bi][e]);
            v = fma(ac4, ai[k][bi][e], v);
            v = fma(ac2, a1[k][bi][e], v);
            v = fma(aw, an[k][bi][e], v);
            bo[e] = v;
          }
          stv_nt<S, NS>(b_out + off, bo, nt & NT_STORE);
        }
      }
    }
  }
  }
}

template <typename S, bool CX, int B, int L, int KQ, bool FUSED_D>
__global__ __launch_bounds__(WG) void k_step_autodiff_stored(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, const S* __restrict__ amm, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  step_autodiff_stored_dev<S, CX, B, L, KQ, FUSED_D>(wg_self(), rp, ci, hv, am, am1, bm1, bm2,
                                                     amm, b_out, dacc, nb, qp, chunk, c4, c2, w,
                                                     sb, sb2, dscale, nt);
}

template <typename S, bool CX, int B, int L, int KQ, bool FUSED_D>
__global__ __launch_bounds__(WG) void k_step_autodiff(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  step_autodiff_dev<S, CX, B, L, KQ, FUSED_D>(wg_self(), rp, ci, hv, am, am1, bm1, bm2, a_out, b_out, dacc, nb, qp, chunk, c4, c2, w, sb, sb2, dscale, nt);
}

// ---------------------------------------------------------------------------
// Sparse outer product on D's block pattern:
//   D_p(bi, bj) += f * sum_s A[B i + bi, s] conj(Bm[B j + bj, s])
// The operation of Engine_CPU::outer_product (engine_cpu.cpp:213-226) and of
// the reference CUDA kernel (outer_product.cu:38-60); here one lane group per
// D block row, KQ packets per lane, lane butterflies for the S reduction.
// TA / TB are the scalar types of A and Bm (storage type or double for xi).
// ---------------------------------------------------------------------------
template <typename S, typename TA, typename TB, bool CX, int B, int L, int KQ>
__device__ __forceinline__ void outer_dev(
    WgId wid,
    const int* __restrict__ drp, const int* __restrict__ dci, const TA* __restrict__ A,
    const TB* __restrict__ Bm, double* __restrict__ dacc, int nb, int qp, int chunk, double f) {
  using GE = Geo<S, CX, B, L, KQ>;
  constexpr int NS = GE::NS;  // scalars of one storage packet (defines the column chunk)
  constexpr int HS = GE::HS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane % L;
  const int g = (threadIdx.x / WAVE) * GE::G + lane / L;
  const RowRange rr = wg_rows(nb, chunk, GE::GWG, wid);
  const size_t rs = (size_t)qp * NS;  // scalars per row (same count for every type)

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = base + g;
    const bool active = i < rr.r1;
    const int p0 = active ? drp[i] : 0;
    const int p1 = active ? drp[i + 1] : 0;
    int nblk = p1 - p0;
#pragma unroll
    for (int m = L; m < WAVE; m <<= 1) nblk = max(nblk, __shfl_xor(nblk, m, WAVE));
    for (int t = 0; t < nblk; t++) {
      const int p = p0 + t;
      const bool live = p < p1;
      const int j = live ? dci[p] : 0;
      double own[GE::DOWN][HS];
#pragma unroll
      for (int o = 0; o < GE::DOWN; o++)
#pragma unroll
        for (int c = 0; c < HS; c++) own[o][c] = 0.0;
      // Rows bi of the block are taken BI at a time to bound the registers of
      // the partial sums (B*B*HS doubles would not fit for B = 8).
      constexpr int BI = B <= 4 ? B : 2;
#pragma unroll
      for (int bi0 = 0; bi0 < B; bi0 += BI) {
        double part[BI][B][HS];
#pragma unroll
        for (int bi = 0; bi < BI; bi++)
#pragma unroll
          for (int bj = 0; bj < B; bj++)
#pragma unroll
            for (int c = 0; c < HS; c++) part[bi][bj][c] = 0.0;
        for (int q0 = 0; q0 < qp; q0 += L * KQ) {
          if (!live) break;
#pragma unroll
          for (int k = 0; k < KQ; k++) {
            const size_t q = (size_t)(q0 + k * L + l) * NS;
            double av[BI][NS], bv[B][NS];
#pragma unroll
            for (int bi = 0; bi < BI; bi++)
              ldv<TA, NS>(A + (size_t)(B * i + bi0 + bi) * rs + q, av[bi]);
#pragma unroll
            for (int bj = 0; bj < B; bj++) ldv<TB, NS>(Bm + (size_t)(B * j + bj) * rs + q, bv[bj]);
#pragma unroll
            for (int bi = 0; bi < BI; bi++)
#pragma unroll
              for (int bj = 0; bj < B; bj++)
#pragma unroll
                for (int e = 0; e < NS; e += HS) {
                  part[bi][bj][0] = fma(av[bi][e], bv[bj][e], part[bi][bj][0]);
                  if constexpr (CX) {
                    part[bi][bj][0] = fma(av[bi][e + 1], bv[bj][e + 1], part[bi][bj][0]);
                    part[bi][bj][1] = fma(av[bi][e + 1], bv[bj][e], part[bi][bj][1]);
                    part[bi][bj][1] = fma(-av[bi][e], bv[bj][e + 1], part[bi][bj][1]);
                  }
                }
          }
        }
#pragma unroll
        for (int bj = 0; bj < B; bj++)
#pragma unroll
          for (int bi = 0; bi < BI; bi++) {
            const int de = bj * B + bi0 + bi;
#pragma unroll
            for (int c = 0; c < HS; c++) {
              const double s = group_sum<L>(part[bi][bj][c]);
              if (l == de % L) own[de / L][c] = s;
            }
          }
      }
      if (live) {
#pragma unroll
        for (int o = 0; o < GE::DOWN; o++) {
          const int de = o * L + l;
          if (de < GE::DE) {
            double* d = dacc + ((size_t)p * GE::DE + de) * HS;
#pragma unroll
            for (int c = 0; c < HS; c++) d[c] = fma(f, own[o][c], d[c]);
          }
        }
      }
    }
  }
}

template <typename S, typename TA, typename TB, bool CX, int B, int L, int KQ>
__global__ __launch_bounds__(WG) void k_outer(
    const int* __restrict__ drp, const int* __restrict__ dci, const TA* __restrict__ A,
    const TB* __restrict__ Bm, double* __restrict__ dacc, int nb, int qp, int chunk, double f) {
  outer_dev<S, TA, TB, CX, B, L, KQ>(wg_self(), drp, dci, A, Bm, dacc, nb, qp, chunk, f);
}

// ---------------------------------------------------------------------------
// Small helper kernels.
// ---------------------------------------------------------------------------

#ifdef FKD_DEFINE_HELPERS
// Deterministic reduction of per-workgroup moment partials: out[2 s + c] =
// sum_w partial[s][w][c], one workgroup per step s, fixed order.
__global__ __launch_bounds__(WG) void k_reduce_partials(const double* __restrict__ partial,
                                                        int nwg, double* __restrict__ out) {
  const double* p = partial + (size_t)blockIdx.x * nwg * 2;
  double a = 0.0, b = 0.0;
  for (int w = threadIdx.x; w < nwg; w += WG) {
    a += p[2 * w];
    b += p[2 * w + 1];
  }
  wg_sum2(a, b, out + 2 * (size_t)blockIdx.x);
}

// Plain streaming copy of n 16-byte elements (the bandwidth probe).
__global__ __launch_bounds__(WG) void k_copy16(const double2* __restrict__ a,
                                               double2* __restrict__ b, size_t n) {
  for (size_t k = (size_t)blockIdx.x * WG + threadIdx.x; k < n; k += (size_t)gridDim.x * WG)
    b[k] = a[k];
}

// Weighted sums of the kept forward vectors for stoch_matrix:
//   v = sum_{i=0..K} d_i alpha_i,   u = sum_{i=0..K-1} e_i alpha_i
// (alpha_0 = R, alpha_i = table slot i-1, slots `slot` scalars apart), one
// streaming pass over the table with FP64 sums; n scalars per vector, a
// multiple of the 16-byte packet.  Each thread keeps U slots' loads in flight.
template <typename S>
__global__ __launch_bounds__(WG) void k_table_sums(const S* __restrict__ R,
                                                   const S* __restrict__ tab, size_t slot, int K,
                                                   const double* __restrict__ d,
                                                   const double* __restrict__ e,
                                                   S* __restrict__ u, S* __restrict__ v, size_t n) {
  typedef typename Vec16<S>::V VT;
  constexpr int NS = Vec16<S>::N, U = 8;
  const size_t np = n / NS;
  for (size_t q = (size_t)blockIdx.x * WG + threadIdx.x; q < np; q += (size_t)gridDim.x * WG) {
    double au[NS], av[NS];
    {
      const VT r = __builtin_nontemporal_load(reinterpret_cast<const VT*>(R) + q);
#pragma unroll
      for (int t = 0; t < NS; t++) {
        av[t] = d[0] * (double)r[t];
        au[t] = K > 0 ? e[0] * (double)r[t] : 0.0;
      }
    }
    for (int i0 = 1; i0 <= K; i0 += U) {
      VT x[U];
#pragma unroll
      for (int k = 0; k < U; k++)
        if (i0 + k <= K)
          x[k] = __builtin_nontemporal_load(reinterpret_cast<const VT*>(tab + (size_t)(i0 + k - 1) * slot) + q);
#pragma unroll
      for (int k = 0; k < U; k++) {
        const int i = i0 + k;
        if (i <= K) {
          const double di = d[i], ei = i < K ? e[i] : 0.0;
#pragma unroll
          for (int t = 0; t < NS; t++) {
            av[t] = fma(di, (double)x[k][t], av[t]);
            au[t] = fma(ei, (double)x[k][t], au[t]);
          }
        }
      }
    }
    VT ou, ov;
#pragma unroll
    for (int t = 0; t < NS; t++) {
      ou[t] = (S)au[t];
      ov[t] = (S)av[t];
    }
    reinterpret_cast<VT*>(u)[q] = ou;
    reinterpret_cast<VT*>(v)[q] = ov;
  }
}

// The fastest copy found on this device (about 6% above k_copy16): four
// 16-byte packets per thread in flight, nontemporal loads and stores.
__global__ __launch_bounds__(WG) void k_copy16_nt(const double2* __restrict__ a,
                                                  double2* __restrict__ b, size_t n) {
  typedef Vec16<double>::V VT;
  constexpr int U = 4;
  const VT* src = reinterpret_cast<const VT*>(a);
  VT* dst = reinterpret_cast<VT*>(b);
  const size_t step = (size_t)gridDim.x * WG * U;
  for (size_t base = (size_t)blockIdx.x * WG * U; base < n; base += step) {
    VT t[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t k = base + (size_t)u * WG + threadIdx.x;
      if (k < n) t[u] = __builtin_nontemporal_load(src + k);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t k = base + (size_t)u * WG + threadIdx.x;
      if (k < n) __builtin_nontemporal_store(t[u], dst + k);
    }
  }
}

// Fill n doubles with zero (graph-capturable, unlike hipMemsetAsync).
__global__ void k_zero(double* __restrict__ p, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

#endif  // FKD_DEFINE_HELPERS

// R (host type TH, column-major n x s) -> storage S, row-major n x sp, zero
// padded.  Tiled through LDS: 64 rows x 16 columns per workgroup.
template <typename TH, typename S, bool CX>
__global__ __launch_bounds__(WG) void k_pack_R(const TH* __restrict__ src, S* __restrict__ dst,
                                               int n, int s, int sp) {
  constexpr int HS = CX ? 2 : 1;
  __shared__ double tile[16][64 * HS + 1];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 16;
  // load: thread t reads row r0 + t % 64 of columns c0 + t / 64 + 4 k
  for (int k = threadIdx.x; k < 64 * 16; k += WG) {
    const int rr = k % 64, cc = k / 64;
    const int r = r0 + rr, c = c0 + cc;
#pragma unroll
    for (int h = 0; h < HS; h++)
      tile[cc][rr * HS + h] = (r < n && c < s) ? (double)src[((size_t)c * n + r) * HS + h] : 0.0;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 16; k += WG) {
    const int cc = k % 16, rr = k / 16;
    const int r = r0 + rr, c = c0 + cc;
    if (r < n && c < sp) {
#pragma unroll
      for (int h = 0; h < HS; h++) dst[((size_t)r * sp + c) * HS + h] = (S)tile[cc][rr * HS + h];
    }
  }
}

// Compact form of R when every row has at most one nonzero (the probing
// colours of set_R_correlated): rcol[r] = its column (-1 for a zero row),
// rval[r] = its value; *flag is set when some row has two or more nonzeros.
// One thread per row; consecutive threads read consecutive rows of a column.
template <typename TH, typename S, bool CX>
__global__ __launch_bounds__(WG) void k_compact_R(const TH* __restrict__ src, int n, int s,
                                                  int* __restrict__ rcol, S* __restrict__ rval,
                                                  int* __restrict__ flag) {
  constexpr int HS = CX ? 2 : 1;
  const int r = blockIdx.x * WG + threadIdx.x;
  if (r >= n) return;
  int cnt = 0, col = -1;
  TH v[HS];
#pragma unroll
  for (int h = 0; h < HS; h++) v[h] = TH(0);
  for (int c = 0; c < s; c++) {
    const TH* e = src + ((size_t)c * n + r) * HS;
    bool nz = false;
#pragma unroll
    for (int h = 0; h < HS; h++) nz |= (e[h] != TH(0));
    if (nz) {
      cnt++;
      col = c;
#pragma unroll
      for (int h = 0; h < HS; h++) v[h] = e[h];
    }
  }
  rcol[r] = cnt == 1 ? col : -1;
#pragma unroll
  for (int h = 0; h < HS; h++) rval[(size_t)r * HS + h] = (S)v[h];
  if (cnt > 1) atomicOr(flag, 1);
}

// Hs values: host type TH -> S with (H - avg I)/mag applied, as
// Engine_cuSPARSE::set_H does (engine_cusparse.cpp:711-727): divide every value
// by mag, then subtract avg/mag on the diagonal of diagonal blocks.  Computed in
// FP64 and rounded once to the storage type.
template <typename TH, typename S, bool CX>
__global__ void k_scale_H(const TH* __restrict__ src, S* __restrict__ dst,
                          const int* __restrict__ row_of_block, const int* __restrict__ ci,
                          int nnzb, int b, double mag, double shift) {
  constexpr int HS = CX ? 2 : 1;
  const size_t total = (size_t)nnzb * b * b;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t p = e / ((size_t)b * b);
    const int r = (int)(e % ((size_t)b * b));
    const int bj = r / b, bi = r % b;
#pragma unroll
    for (int h = 0; h < HS; h++) {
      double v = (double)src[e * HS + h] / mag;
      if (h == 0 && bi == bj && row_of_block[p] == ci[p]) v -= shift;
      dst[e * HS + h] = (S)v;
    }
  }
}

// D_out_k(bi, bj) = ((D_k(bi, bj) + conj(D_t(k)(bj, bi))) / 2) * inv_mag, the
// combined SpMatBsr::symmetrize + scale(1/mag) (spmat.cpp:171-197) written out
// of place; t(k) is the index of block (j, i), or -1 for an upper block
// without a partner, which symmetrize leaves unchanged.
template <bool CX>
__global__ void k_symmetrize_scale(const double* __restrict__ din, double* __restrict__ dout,
                                   const int* __restrict__ tpos, int nnzb, int b, double inv_mag) {
  constexpr int HS = CX ? 2 : 1;
  const size_t bb = (size_t)b * b;
  const size_t total = (size_t)nnzb * bb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t k = e / bb;
    const int r = (int)(e % bb);
    const int bj = r / b, bi = r % b;
    if (tpos[k] < 0) {
#pragma unroll
      for (int h = 0; h < HS; h++) dout[e * HS + h] = din[e * HS + h] * inv_mag;
      continue;
    }
    const size_t e2 = (size_t)tpos[k] * bb + (size_t)bi * b + bj;
    const double re = (din[e * HS] + din[e2 * HS]) / 2.0;
    dout[e * HS] = re * inv_mag;
    if constexpr (CX) {
      const double im = (din[e * HS + 1] - din[e2 * HS + 1]) / 2.0;
      dout[e * HS + 1] = im * inv_mag;
    }
  }
}

#ifdef FKD_DEFINE_HELPERS
// ---------------------------------------------------------------------------
// Lanczos helpers (FP64, one column) with the coefficients kept on the
// device, so the loop runs without a host round trip per iteration.
// coef = [alpha_0 .. alpha_{it-1}, beta_0 .. beta_{it-1}], beta_0 = 0; ps =
// the two reduced sums of the previous kernel.  Step j (1 <= j < it):
//   orth:  alpha_{j-1} = ps[1] = <H v1, v1>;  w -= alpha_{j-1} v1 + beta_{j-1} v0;
//          partial sums of |w|^2
//   next:  beta_j = sqrt(ps[0]); v0 = v1; v1 = w / beta_j (0 if beta_j = 0);
//          used = j at the first beta_j <= 1e-300 (an invariant subspace: the
//          j x j tridiagonal matrix is exact)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_lanczos_orth(double* __restrict__ w,
                                                     const double* __restrict__ v1,
                                                     const double* __restrict__ v0, size_t n,
                                                     const double* __restrict__ ps,
                                                     double* __restrict__ coef, int it, int j,
                                                     double* __restrict__ partial) {
  const double a = ps[1], b = coef[it + j - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) coef[j - 1] = a;
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * WG + threadIdx.x; k < n; k += (size_t)gridDim.x * WG) {
    const double v = w[k] - a * v1[k] - b * v0[k];
    w[k] = v;
    s = fma(v, v, s);
  }
  wg_sum2(s, 0.0, partial + 2 * (size_t)blockIdx.x);
}

__global__ void k_lanczos_next(double* __restrict__ v0, double* __restrict__ v1,
                               const double* __restrict__ w, size_t n,
                               const double* __restrict__ ps, double* __restrict__ coef, int it,
                               int j, int* __restrict__ used) {
  const double beta = sqrt(ps[0]);
  const bool alive = beta > 1e-300;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coef[it + j] = beta;
    if (!alive && *used == it) *used = j;
  }
  const double inv = alive ? 1.0 / beta : 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    v0[k] = v1[k];
    v1[k] = w[k] * inv;
  }
}

// last = 0: the start of the table (beta_0 = 0, used = it); last = 1: the
// final alpha_{it-1} = ps[1].
__global__ void k_lanczos_edge(const double* __restrict__ ps, double* __restrict__ coef, int it,
                               int* __restrict__ used, int last) {
  if (threadIdx.x != 0) return;
  if (last) {
    coef[it - 1] = ps[1];
  } else {
    coef[it] = 0.0;
    *used = it;
  }
}

#endif  // FKD_DEFINE_HELPERS

}  // namespace fkd
