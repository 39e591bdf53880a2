// kpm_mfma.h -- matrix-core (MFMA) step kernels for complex 8x8 BSR blocks.
//
// With 8 orbitals per block the block SpMM Y_i = sum_j Hs_ij X_j is a dense
// contraction of 8x8 complex blocks against 8 x S complex column strips, at
// about 17 flop per HBM byte in FP32: vector FMAs would make it ALU-bound, so
// the contraction runs on the CDNA4 matrix cores in its real 16x16 form
//   [Y_re]   [H_re  -H_im] [X_re]
//   [Y_im] = [H_im   H_re] [X_im]
// with v_mfma_{f32,f64}_16x16x4: four MFMAs (K = 16) per neighbour block and
// 16-column tile.  One wave owns one block row; lane (q, c) = (lane / 16,
// lane % 16) holds orbitals q and q + 4 of column c of a tile, both as an
// MFMA B operand (X_j) and as accumulator rows (Y_i), so the own-row operands
// of the axpby and of the moment dots sit in the same lane as the result.
// The rows of the 16x16 operator are permuted to make that so (mfma_lane).
//
// The autodiff step's outer product D_ij += f alpha_i beta_j^+ on H's pattern
// is the 16x16 real product G = P Q^T over the S columns (P, Q = [re; im] of
// the two 8 x S strips), also on the matrix cores; D follows from
//   D_re = G[a][b] + G[8+a][8+b],  D_im = G[8+a][b] - G[a][8+b].
// G accumulates per tile (per pair of tiles in the FP32 pair-of-tiles
// autodiff step) in the storage type and across tiles in FP64.
//
// MFMA 16x16x4 lane layout (64 lanes, l = lane, register i = 0..3):
//   A (16 x 4): lane l holds A[l % 16][l / 16]
//   B (4 x 16): lane l holds B[l / 16][l % 16]
//   C (16 x 16), f32: lane l holds C[4 (l / 16) + i][l % 16]
//                f64: lane l holds C[l / 16 + 4 i][l % 16]
// (checked on gfx950).  The operator's row permutation differs between the
// two so that in both, register i of lane (q, c) is (re, im, re, im)[i] of
// orbital (q, q, q + 4, q + 4)[i] in column c.
#pragma once

#include "kpm_kernels.h"

namespace fkd {

template <typename S>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float2 c2;
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double2 c2;
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

typedef double dv4 __attribute__((ext_vector_type(4)));

// Lane-constant part of the operator row this lane feeds as MFMA A operand.
struct MfmaLane {
  int q, c;    // lane / 16, lane % 16
  int h0, h1;  // scalar offsets in an H block of Hs(O, q) and Hs(O, q + 4)
  bool im;     // this A row is an imaginary-part row
};

template <typename S>
__device__ __forceinline__ MfmaLane mfma_lane() {
  MfmaLane m;
  const int lane = threadIdx.x & (WAVE - 1);
  m.q = lane >> 4;
  m.c = lane & 15;
  // A row r = c feeds accumulator row r: register r % 4 of lane group r / 4
  // (f32) or register r / 4 of lane group r % 4 (f64)
  const int grp = sizeof(S) == 4 ? (m.c >> 2) : (m.c & 3);
  const int reg = sizeof(S) == 4 ? (m.c & 3) : (m.c >> 2);
  const int O = grp + 4 * (reg >> 1);
  m.im = (reg & 1) != 0;
  m.h0 = (m.q * 8 + O) * 2;  // blocks are column-major: (col * 8 + row) * 2
  m.h1 = ((m.q + 4) * 8 + O) * 2;
  return m;
}


template <typename S>
__device__ __forceinline__ typename Mfma<S>::v4 mfma_block(const S (&a)[4],
                                                          const typename Mfma<S>::c2& x0,
                                                          const typename Mfma<S>::c2& x1,
                                                          typename Mfma<S>::v4 acc) {
  using MF = Mfma<S>;
  acc = MF::mac(a[0], x0.x, acc);
  acc = MF::mac(a[1], x0.y, acc);
  acc = MF::mac(a[2], x1.x, acc);
  acc = MF::mac(a[3], x1.y, acc);
  return acc;
}

template <typename S>
__device__ __forceinline__ typename Mfma<S>::c2 ldc(const S* __restrict__ p) {
  return *reinterpret_cast<const typename Mfma<S>::c2*>(p);
}
template <typename S>
__device__ __forceinline__ void stc(S* __restrict__ p, typename Mfma<S>::c2 v) {
  *reinterpret_cast<typename Mfma<S>::c2*>(p) = v;
}

// This lane's raw values of one Hs block: Hs(O, q) and Hs(O, q + 4).
template <typename S>
struct HRaw {
  typename Mfma<S>::c2 e0, e1;
};

template <typename S>
__device__ __forceinline__ HRaw<S> mfma_hraw(const S* __restrict__ blk, const MfmaLane& m) {
  return HRaw<S>{ldc<S>(blk + m.h0), ldc<S>(blk + m.h1)};
}

// The four A-operand values (K chunks 0..3: re/im of orbital q, re/im of
// orbital q + 4) of a block, times w (0 for a padding slot of a batch).  Kept
// apart from the loads so that a batch issues all its loads before any use.
template <typename S>
__device__ __forceinline__ void mfma_a_of(const HRaw<S>& h, const MfmaLane& m, S w, S (&a)[4]) {
  a[0] = w * (m.im ? h.e0.y : h.e0.x);
  a[1] = w * (m.im ? h.e0.x : -h.e0.y);
  a[2] = w * (m.im ? h.e1.y : h.e1.x);
  a[3] = w * (m.im ? h.e1.x : -h.e1.y);
}

// Neighbour blocks gathered per batch (MF_U) and 16-column tiles per pass
// (MF_TU) of the forward step.
constexpr int MF_U = 4;
constexpr int MF_TU = 2;

// Y tiles t0 .. t0 + TU - 1 of block row i: acc[tt] += sum_p Hs_p X_{j_p}
// over the row's blocks [p0, p1) (p1 > p0), X gathered at column offset
// (scalars) co[tt], which the caller keeps inside the row.  The loads of a
// batch have no branches between them: a padding slot past the row's end
// reloads the row's last block and enters the products with weight 0.
template <typename S, int MF_U = fkd::MF_U, int MF_TU = fkd::MF_TU>
__device__ __forceinline__ void mfma_gather(const int* __restrict__ ci, const S* __restrict__ hv,
                                            const S* __restrict__ X, size_t rs, const MfmaLane& m,
                                            int p0, int p1, const size_t (&co)[MF_TU],
                                            typename Mfma<S>::v4 (&acc)[MF_TU]) {
  using C2 = typename Mfma<S>::c2;
  for (int p = p0; p < p1; p += MF_U) {
    HRaw<S> hr[MF_U];
    C2 x0[MF_U][MF_TU], x1[MF_U][MF_TU];
#pragma unroll
    for (int u = 0; u < MF_U; u++) {
      const int pu = min(p + u, p1 - 1);
      const int j = ci[pu];
      hr[u] = mfma_hraw<S>(hv + (size_t)pu * 128, m);
      const size_t r0 = (size_t)(8 * j + m.q) * rs, r1 = (size_t)(8 * j + 4 + m.q) * rs;
#pragma unroll
      for (int tt = 0; tt < MF_TU; tt++) {
        x0[u][tt] = ldc<S>(X + r0 + co[tt]);
        x1[u][tt] = ldc<S>(X + r1 + co[tt]);
      }
    }
#pragma unroll
    for (int u = 0; u < MF_U; u++) {
      S a[4];
      mfma_a_of<S>(hr[u], m, (p + u < p1) ? S(1) : S(0), a);
#pragma unroll
      for (int tt = 0; tt < MF_TU; tt++) acc[tt] = mfma_block<S>(a, x0[u][tt], x1[u][tt], acc[tt]);
    }
  }
}

// Clenshaw source term cr * R for one lane's element (row, col) pair.
template <typename S>
__device__ __forceinline__ void mfma_rsource(const RSource<S>& rs_, size_t row, int col,
                                             size_t off, typename Mfma<S>::c2& y) {
  const S cr = (S)rs_.cr;
  if (rs_.rd != nullptr) {
    const typename Mfma<S>::c2 r = ldc<S>(rs_.rd + off);
    y.x = fma(cr, r.x, y.x);
    y.y = fma(cr, r.y, y.y);
  } else if (rs_.rcol[row] == col) {
    y.x = fma(cr, rs_.rval[2 * row], y.x);
    y.y = fma(cr, rs_.rval[2 * row + 1], y.y);
  }
}

// ---------------------------------------------------------------------------
// Moments / Clenshaw step on the matrix cores (complex 8x8 blocks):
//   y = cx (Hs x) + cz z [+ cr R], dots sum |x_i|^2 and Re<y_i, x_i>.
// Same contract as k_step_moments (engine_cpu.cpp:63-84 fused with both cdot's).
// ---------------------------------------------------------------------------
template <typename S, int MF_U = fkd::MF_U, int MF_TU = fkd::MF_TU>
__device__ __forceinline__ void mfma_step_dev(
    WgId wid,
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  using C2 = typename Mfma<S>::c2;
  using V4 = typename Mfma<S>::v4;
  constexpr int NS = Vec16<S>::N;
  (void)nt;
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const size_t rs = (size_t)qp * NS;  // scalars per orbital row
  const int ntile = (int)(rs / 32);   // 16 complex columns per tile
  const S acx = (S)cx, acz = (S)cz;
  double dmm = 0.0, dm1 = 0.0;

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], p1 = rp[i + 1];
    const size_t o0 = (size_t)(8 * i + m.q) * rs, o1 = (size_t)(8 * i + 4 + m.q) * rs;
    for (int t0 = 0; t0 < ntile; t0 += MF_TU) {
      size_t co[MF_TU];
      bool tv[MF_TU];
      V4 acc[MF_TU];
#pragma unroll
      for (int tt = 0; tt < MF_TU; tt++) {
        // a tile past the row's end recomputes the last one; its result is dropped
        co[tt] = (size_t)min(t0 + tt, ntile - 1) * 32 + 2 * m.c;
        tv[tt] = (t0 + tt) < ntile;
        acc[tt] = V4{S(0), S(0), S(0), S(0)};
      }
      if (p1 > p0) mfma_gather<S, MF_U, MF_TU>(ci, hv, x, rs, m, p0, p1, co, acc);
#pragma unroll
      for (int tt = 0; tt < MF_TU; tt++) {
        if (!tv[tt]) continue;
        const size_t f0 = o0 + co[tt], f1 = o1 + co[tt];
        const C2 x0 = ldc<S>(x + f0), x1 = ldc<S>(x + f1);
        C2 z0{S(0), S(0)}, z1{S(0), S(0)};
        if (z != nullptr) {
          z0 = ldc<S>(z + f0);
          z1 = ldc<S>(z + f1);
        }
        C2 y0, y1;
        y0.x = fma(acx, acc[tt].x, acz * z0.x);
        y0.y = fma(acx, acc[tt].y, acz * z0.y);
        y1.x = fma(acx, acc[tt].z, acz * z1.x);
        y1.y = fma(acx, acc[tt].w, acz * z1.y);
        if (rsrc.cr != 0.0) {
          const int col = (t0 + tt) * 16 + m.c;
          mfma_rsource<S>(rsrc, (size_t)(8 * i + m.q), col, f0, y0);
          mfma_rsource<S>(rsrc, (size_t)(8 * i + 4 + m.q), col, f1, y1);
        }
        stc<S>(y + f0, y0);
        stc<S>(y + f1, y1);
        const double a0 = x0.x, a1 = x0.y, a2 = x1.x, a3 = x1.y;
        dmm = fma(a0, a0, fma(a1, a1, fma(a2, a2, fma(a3, a3, dmm))));
        dm1 = fma((double)y0.x, a0,
                  fma((double)y0.y, a1, fma((double)y1.x, a2, fma((double)y1.y, a3, dm1))));
      }
    }
  }
  if (partial != nullptr) wg_sum2(dmm, dm1, partial + 2 * (size_t)wid.k);
}

template <typename S>
__global__ __launch_bounds__(WG) void k_mfma_step(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  mfma_step_dev<S>(wg_self(), rp, ci, hv, x, z, y, nb, qp, chunk, cx, cz, partial, nt, rsrc);
}

// ---------------------------------------------------------------------------
// Row-resident forward step for rows of at most DM blocks.  A block row's A
// operands (selected and signed Hs values) and its neighbours' row offsets
// are computed once per row; the tile loop then issues only buffer loads
// (SGPR-based resource, 32-bit per-lane offsets, the tile's offset in an
// SGPR), MFMAs and the epilogue.  Same contract as mfma_step_dev; the host
// uses it when every row has at most DM blocks and a vector spans less than
// 4 GiB.
// ---------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const void* p, uint32_t bytes) {
  // raw buffer (stride 0, dword 3 of the gfx9 resource); reads past `bytes` return 0
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// Raw buffer loads / stores of one complex element, bound to the LLVM
// intrinsics (one dwordx2 / dwordx4 access; voffset per lane, soffset uniform).
typedef float bf32x2 __attribute__((ext_vector_type(2)));
typedef double bf64x2 __attribute__((ext_vector_type(2)));
__device__ bf32x2 fkd_raw_ld_f32x2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ bf64x2 fkd_raw_ld_f64x2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ void fkd_raw_st_f32x2(bf32x2 v, __amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                 int aux) __asm("llvm.amdgcn.raw.ptr.buffer.store.v2f32");
__device__ void fkd_raw_st_f64x2(bf64x2 v, __amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                 int aux) __asm("llvm.amdgcn.raw.ptr.buffer.store.v2f64");

typedef float bf32x4 __attribute__((ext_vector_type(4)));
__device__ bf32x4 fkd_raw_ld_f32x4(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ void fkd_raw_st_f32x4(bf32x4 v, __amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                 int aux) __asm("llvm.amdgcn.raw.ptr.buffer.store.v4f32");

__device__ float fkd_raw_ld_f32(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                               int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.f32");
__device__ double fkd_raw_ld_f64(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.f64");

template <typename S>
struct BufC2;
template <>
struct BufC2<float> {
  static __device__ __forceinline__ float2 ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    const bf32x2 v = fkd_raw_ld_f32x2(r, (int)vo, (int)so, 0);
    return float2{v.x, v.y};
  }
  static __device__ __forceinline__ void st(float2 d, __amdgpu_buffer_rsrc_t r, uint32_t vo,
                                            uint32_t so) {
    fkd_raw_st_f32x2(bf32x2{d.x, d.y}, r, (int)vo, (int)so, 0);
  }
  static __device__ __forceinline__ float ld1(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    return fkd_raw_ld_f32(r, (int)vo, (int)so, 0);
  }
};
template <>
struct BufC2<double> {
  static __device__ __forceinline__ double2 ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    const bf64x2 v = fkd_raw_ld_f64x2(r, (int)vo, (int)so, 0);
    return double2{v.x, v.y};
  }
  static __device__ __forceinline__ void st(double2 d, __amdgpu_buffer_rsrc_t r, uint32_t vo,
                                            uint32_t so) {
    fkd_raw_st_f64x2(bf64x2{d.x, d.y}, r, (int)vo, (int)so, 0);
  }
  static __device__ __forceinline__ double ld1(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    return fkd_raw_ld_f64(r, (int)vo, (int)so, 0);
  }
};

constexpr int MF_DM = 8;  // most blocks per row of the row-resident kernels

template <typename S, int DM = MF_DM, int G = 4>
__device__ __forceinline__ void mfma_step_rr_dev(
    WgId wid, const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  using C2 = typename Mfma<S>::c2;
  using V4 = typename Mfma<S>::v4;
  using IO = BufC2<S>;
  constexpr int NS = Vec16<S>::N;
  constexpr uint32_t ES = sizeof(S);
  (void)nt;
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const uint32_t rs = (uint32_t)qp * NS;  // scalars per orbital row
  const int ntile = (int)(rs / 32);
  const uint32_t vbytes = 8u * (uint32_t)nb * rs * ES;
  const __amdgpu_buffer_rsrc_t bx = vec_rsrc(x, vbytes), by = vec_rsrc(y, vbytes);
  const __amdgpu_buffer_rsrc_t bz = vec_rsrc(z != nullptr ? z : x, z != nullptr ? vbytes : 0u);
  const S acx = (S)cx, acz = (S)cz;
  const uint32_t half = 4u * rs * ES;  // orbital q -> q + 4
  // a neighbour's offset: this lane's constant part (VGPR) plus the block's
  // wave-uniform 8 j rs ES (the buffer load's SGPR offset)
  const uint32_t vlo = (m.q * rs + 2u * m.c) * ES;
  double dmm = 0.0, dm1 = 0.0;

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], deg = rp[i + 1] - p0;
    S a[DM][4];
    uint32_t sj[DM];
    if (deg > 0) {
      // every block's index and H values in one batch of loads (padding
      // slots: the last block, weight 0)
      int jj[DM];
      HRaw<S> hr[DM];
#pragma unroll
      for (int u = 0; u < DM; u++) {
        const int pu = p0 + min(u, deg - 1);
        jj[u] = ci[pu];
        hr[u] = mfma_hraw<S>(hv + (size_t)pu * 128, m);
      }
#pragma unroll
      for (int u = 0; u < DM; u++) {
        mfma_a_of<S>(hr[u], m, u < deg ? S(1) : S(0), a[u]);
        sj[u] = (uint32_t)__builtin_amdgcn_readfirstlane(8 * jj[u]) * rs * ES;
      }
    } else {
#pragma unroll
      for (int u = 0; u < DM; u++) {
#pragma unroll
        for (int k = 0; k < 4; k++) a[u][k] = S(0);
        sj[u] = 0u;
      }
    }
    const uint32_t f0 = ((uint32_t)(8 * i + m.q) * rs + 2u * m.c) * ES;
    for (int t = 0; t < ntile; t++) {
      const uint32_t so = (uint32_t)t * 32u * ES;
      V4 acc = V4{S(0), S(0), S(0), S(0)};
#pragma unroll
      for (int u0 = 0; u0 < DM; u0 += G) {
        C2 x0[G], x1[G];
#pragma unroll
        for (int v = 0; v < G; v++) {
          if (u0 + v >= DM) continue;  // DM % G != 0: the last group is short
          x0[v] = IO::ld(bx, vlo, sj[u0 + v] + so);
          x1[v] = IO::ld(bx, vlo, sj[u0 + v] + so + half);
        }
#pragma unroll
        for (int v = 0; v < G; v++)
          if (u0 + v < DM) acc = mfma_block<S>(a[u0 + v], x0[v], x1[v], acc);
      }
      const C2 xo0 = IO::ld(bx, f0, so), xo1 = IO::ld(bx, f0 + half, so);
      const C2 z0 = IO::ld(bz, f0, so), z1 = IO::ld(bz, f0 + half, so);  // 0 without z
      C2 y0, y1;
      y0.x = fma(acx, acc.x, acz * z0.x);
      y0.y = fma(acx, acc.y, acz * z0.y);
      y1.x = fma(acx, acc.z, acz * z1.x);
      y1.y = fma(acx, acc.w, acz * z1.y);
      if (rsrc.cr != 0.0) {
        const int col = t * 16 + m.c;
        mfma_rsource<S>(rsrc, (size_t)(8 * i + m.q), col, (f0 + so) / ES, y0);
        mfma_rsource<S>(rsrc, (size_t)(8 * i + 4 + m.q), col, (f0 + half + so) / ES, y1);
      }
      IO::st(y0, by, f0, so);
      IO::st(y1, by, f0 + half, so);
      const double a0 = xo0.x, a1 = xo0.y, a2 = xo1.x, a3 = xo1.y;
      dmm = fma(a0, a0, fma(a1, a1, fma(a2, a2, fma(a3, a3, dmm))));
      dm1 = fma((double)y0.x, a0,
                fma((double)y0.y, a1, fma((double)y1.x, a2, fma((double)y1.y, a3, dm1))));
    }
  }
  if (partial != nullptr) wg_sum2(dmm, dm1, partial + 2 * (size_t)wid.k);
}

// FP32 row-resident forward / Clenshaw step on pairs of column tiles (same
// contract as mfma_step_rr_dev<float>; qp * 4 a multiple of 64).  Lane (q, c)
// takes columns 2c and 2c + 1 of a pair of 16-column tiles with one 16-byte
// access per orbital row (q and q + 4) and feeds them to two accumulators, so
// each group of neighbour loads serves 32 columns.
template <int DM, int G>
__device__ __forceinline__ void mfma_step_rr2_dev(
    WgId wid, const int* __restrict__ rp, const int* __restrict__ ci,
    const float* __restrict__ hv, const float* __restrict__ x, const float* __restrict__ z,
    float* __restrict__ y, int nb, int qp, int chunk, double cx, double cz,
    double* __restrict__ partial, RSource<float> rsrc) {
  using S = float;
  using C2 = Mfma<float>::c2;
  using V4 = Mfma<float>::v4;
  constexpr int NS = Vec16<S>::N;
  constexpr uint32_t ES = sizeof(S);
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const uint32_t rs = (uint32_t)qp * NS;
  const int npair = (int)(rs / 64);
  const uint32_t vbytes = 8u * (uint32_t)nb * rs * ES;
  const __amdgpu_buffer_rsrc_t bx = vec_rsrc(x, vbytes), by = vec_rsrc(y, vbytes);
  const __amdgpu_buffer_rsrc_t bz = vec_rsrc(z != nullptr ? z : x, z != nullptr ? vbytes : 0u);
  const S acx = (S)cx, acz = (S)cz;
  const uint32_t half = 4u * rs * ES;
  const uint32_t vlo = (m.q * rs + 4u * m.c) * ES;  // columns 2c, 2c + 1 of orbital q
  double dmm = 0.0, dm1 = 0.0;

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], deg = rp[i + 1] - p0;
    S a[DM][4];
    uint32_t sj[DM];
    if (deg > 0) {
      int jj[DM];
      HRaw<S> hr[DM];
#pragma unroll
      for (int u = 0; u < DM; u++) {
        const int pu = p0 + min(u, deg - 1);
        jj[u] = ci[pu];
        hr[u] = mfma_hraw<S>(hv + (size_t)pu * 128, m);
      }
#pragma unroll
      for (int u = 0; u < DM; u++) {
        mfma_a_of<S>(hr[u], m, u < deg ? S(1) : S(0), a[u]);
        sj[u] = (uint32_t)__builtin_amdgcn_readfirstlane(8 * jj[u]) * rs * ES;
      }
    } else {
#pragma unroll
      for (int u = 0; u < DM; u++) {
#pragma unroll
        for (int k = 0; k < 4; k++) a[u][k] = S(0);
        sj[u] = 0u;
      }
    }
    const uint32_t f0 = ((uint32_t)(8 * i + m.q) * rs + 4u * m.c) * ES;
    // compact R (one nonzero per orbital row): the lane's two rows' column and
    // value, loaded once per block row rather than per tile
    const bool rcomp = rsrc.cr != 0.0 && rsrc.rd == nullptr;
    int rc[2] = {-1, -1};
    S rv[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    if (rcomp) {
#pragma unroll
      for (int o = 0; o < 2; o++) {
        const size_t row = (size_t)(8 * i + 4 * o + m.q);
        rc[o] = rsrc.rcol[row];
        rv[o][0] = rsrc.rval[2 * row];
        rv[o][1] = rsrc.rval[2 * row + 1];
      }
    }
    const S acr = (S)rsrc.cr;
    for (int t = 0; t < npair; t++) {
      const uint32_t so = (uint32_t)t * 64u * ES;
      V4 acc[2] = {V4{0.f, 0.f, 0.f, 0.f}, V4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int u0 = 0; u0 < DM; u0 += G) {
        bf32x4 x0[G], x1[G];
#pragma unroll
        for (int v = 0; v < G; v++) {
          if (u0 + v >= DM) continue;  // DM % G != 0: the last group is short
          x0[v] = fkd_raw_ld_f32x4(bx, (int)vlo, (int)(sj[u0 + v] + so), 0);
          x1[v] = fkd_raw_ld_f32x4(bx, (int)vlo, (int)(sj[u0 + v] + so + half), 0);
        }
#pragma unroll
        for (int v = 0; v < G; v++) {
          if (u0 + v >= DM) continue;
          acc[0] = mfma_block<S>(a[u0 + v], C2{x0[v].x, x0[v].y}, C2{x1[v].x, x1[v].y}, acc[0]);
          acc[1] = mfma_block<S>(a[u0 + v], C2{x0[v].z, x0[v].w}, C2{x1[v].z, x1[v].w}, acc[1]);
        }
      }
      const bf32x4 xo0 = fkd_raw_ld_f32x4(bx, (int)f0, (int)so, 0);
      const bf32x4 xo1 = fkd_raw_ld_f32x4(bx, (int)(f0 + half), (int)so, 0);
      const bf32x4 z0 = fkd_raw_ld_f32x4(bz, (int)f0, (int)so, 0);  // 0 without z
      const bf32x4 z1 = fkd_raw_ld_f32x4(bz, (int)(f0 + half), (int)so, 0);
      C2 y0[2], y1[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        y0[h].x = fma(acx, acc[h].x, acz * z0[2 * h]);
        y0[h].y = fma(acx, acc[h].y, acz * z0[2 * h + 1]);
        y1[h].x = fma(acx, acc[h].z, acz * z1[2 * h]);
        y1[h].y = fma(acx, acc[h].w, acz * z1[2 * h + 1]);
        const int col = t * 32 + 2 * m.c + h;
        if (rcomp) {
          if (rc[0] == col) {
            y0[h].x = fma(acr, rv[0][0], y0[h].x);
            y0[h].y = fma(acr, rv[0][1], y0[h].y);
          }
          if (rc[1] == col) {
            y1[h].x = fma(acr, rv[1][0], y1[h].x);
            y1[h].y = fma(acr, rv[1][1], y1[h].y);
          }
        } else if (rsrc.cr != 0.0) {  // dense R
          mfma_rsource<S>(rsrc, (size_t)(8 * i + m.q), col, (f0 + so) / ES + 2 * h, y0[h]);
          mfma_rsource<S>(rsrc, (size_t)(8 * i + 4 + m.q), col, (f0 + half + so) / ES + 2 * h,
                          y1[h]);
        }
      }
      fkd_raw_st_f32x4(bf32x4{y0[0].x, y0[0].y, y0[1].x, y0[1].y}, by, (int)f0, (int)so, 0);
      fkd_raw_st_f32x4(bf32x4{y1[0].x, y1[0].y, y1[1].x, y1[1].y}, by, (int)(f0 + half), (int)so, 0);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const double a0 = xo0[2 * h], a1 = xo0[2 * h + 1], a2 = xo1[2 * h], a3 = xo1[2 * h + 1];
        dmm = fma(a0, a0, fma(a1, a1, fma(a2, a2, fma(a3, a3, dmm))));
        dm1 = fma((double)y0[h].x, a0,
                  fma((double)y0[h].y, a1, fma((double)y1[h].x, a2, fma((double)y1[h].y, a3, dm1))));
      }
    }
  }
  if (partial != nullptr) wg_sum2(dmm, dm1, partial + 2 * (size_t)wid.k);
}

// D(a, b) += fD * (G[a][b] + G[8+a][8+b], G[8+a][b] - G[a][8+b]) of block p from
// the accumulator G in the MFMA output layout of S (see the file comment).
template <typename S>
__device__ __forceinline__ void mfma_flush_D(const dv4& G, const MfmaLane& m, double fD,
                                             double* __restrict__ dblk) {
  if constexpr (sizeof(S) == 4) {
    // lane (q, c) holds G[4q + i][c]; lanes q < 2, c < 8 own D(a = 4q + i, b = c)
    double dre[4], dim[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const double gv = G[e];
      const double g40 = __shfl_xor(gv, 40);  // G[8 + a][8 + b]
      const double g32 = __shfl_xor(gv, 32);  // G[8 + a][b]
      const double g08 = __shfl_xor(gv, 8);   // G[a][8 + b]
      dre[e] = gv + g40;
      dim[e] = g32 - g08;
    }
    if (m.q < 2 && m.c < 8) {
      double* d = dblk + (m.c * 8 + 4 * m.q) * 2;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        d[2 * e] = fma(fD, dre[e], d[2 * e]);
        d[2 * e + 1] = fma(fD, dim[e], d[2 * e + 1]);
      }
    }
  } else {
    // lane (q, c) holds G[q + 4i][c]: G[a][c] (i = 0, 1), G[8 + a][c] (i = 2, 3)
    double dre[2], dim[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const double gr = G[e], gi = G[e + 2];
      const double xr = __shfl_xor(gr, 8);  // G[a][8 + b]
      const double xi = __shfl_xor(gi, 8);  // G[8 + a][8 + b]
      dre[e] = gr + xi;
      dim[e] = gi - xr;
    }
    if (m.c < 8) {
      double* d = dblk + (m.c * 8 + m.q) * 2;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        d[8 * e] = fma(fD, dre[e], d[8 * e]);
        d[8 * e + 1] = fma(fD, dim[e], d[8 * e + 1]);
      }
    }
  }
}

// Neighbour blocks whose loads the row-resident autodiff step issues together
// before their products: with the A operands in LDS, four fit the registers
// of 2 waves per SIMD in FP64 on 7 block slots (232 VGPRs + 24 AGPRs), two on
// 8 slots (four would leave one wave), two in FP32 (3 waves).
template <typename S, int DM>
constexpr int mf_rr_group() {
  return sizeof(S) == 8 && DM == 7 ? 4 : 2;
}

// ---------------------------------------------------------------------------
// Row-resident autodiff step for rows of at most DM blocks (same contract as
// mfma_autodiff_dev; the host uses it when every row has at most DM blocks
// and a vector spans less than 4 GiB): A operands and neighbour offsets once
// per row, buffer loads in the tile loop.
// ---------------------------------------------------------------------------
template <typename S, bool FUSED_D, int DM = MF_DM, int GR = 2>
__device__ __forceinline__ void mfma_autodiff_rr_dev(
    WgId wid, const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  using C2 = typename Mfma<S>::c2;
  using V4 = typename Mfma<S>::v4;
  using MF = Mfma<S>;
  using IO = BufC2<S>;
  constexpr int NS = Vec16<S>::N;
  constexpr uint32_t ES = sizeof(S);
  (void)nt;
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const uint32_t rs = (uint32_t)qp * NS;
  const int ntile = (int)(rs / 32);
  const uint32_t vbytes = 8u * (uint32_t)nb * rs * ES;
  const bool vb1 = (bm1 == nullptr);
  const double hb1 = vb1 ? sb : 1.0;
  const S* xbp = vb1 ? am : bm1;  // gathered beta_{m+1} (or alpha_m, scaled by sb)
  const __amdgpu_buffer_rsrc_t ba = vec_rsrc(am, vbytes), bb = vec_rsrc(xbp, vbytes);
  const __amdgpu_buffer_rsrc_t ba1 = vec_rsrc(am1, vbytes);
  const __amdgpu_buffer_rsrc_t bb2 = vec_rsrc(bm2 != nullptr ? bm2 : am1, vbytes);
  const __amdgpu_buffer_rsrc_t bao = vec_rsrc(a_out, vbytes), bbo = vec_rsrc(b_out, vbytes);
  const S a2 = (S)2, ab2 = (S)(2.0 * hb1), ac4 = (S)c4, ac2 = (S)c2, aw = (S)w;
  const S asb2 = bm2 != nullptr ? S(1) : (S)sb2;  // beta_{m+2} = asb2 * (bm2 or alpha_{m+1})
  const double fD = dscale * hb1;
  const int dorb = m.c & 7, dpart = m.c >> 3;
  const uint32_t half = 4u * rs * ES;

  // This wave's slice of LDS holds the row's A operands across the tile loop
  // (lane-major, four values a lane), and a neighbour's offset splits into a
  // lane constant (VGPR) and the block's wave-uniform part 8 j rs ES (the
  // buffer load's SGPR offset): the registers go to loads in flight instead.
  __shared__ S lds_a[WG / WAVE][DM][WAVE][4];
  const int lane = threadIdx.x & (WAVE - 1);
  const uint32_t vlo = (m.q * rs + 2u * m.c) * ES;                    // MFMA operand layout
  const uint32_t vld = (dorb * rs + (uint32_t)dpart + 2u * m.q) * ES;  // D layout
  S (&la)[DM][WAVE][4] = lds_a[g];

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], deg = rp[i + 1] - p0;
    uint32_t sj[DM];  // 8 j rs ES of block u's column block j (wave-uniform)
    if (deg > 0) {
      // every block's index and H values in one batch of loads
      int jj[DM];
      HRaw<S> hr[DM];
#pragma unroll
      for (int u = 0; u < DM; u++) {
        const int pu = p0 + min(u, deg - 1);
        jj[u] = ci[pu];
        hr[u] = mfma_hraw<S>(hv + (size_t)pu * 128, m);
      }
#pragma unroll
      for (int u = 0; u < DM; u++) {
        S a[4];
        mfma_a_of<S>(hr[u], m, u < deg ? S(1) : S(0), a);
#pragma unroll
        for (int q = 0; q < 4; q++) la[u][lane][q] = a[q];
        sj[u] = (uint32_t)__builtin_amdgcn_readfirstlane(8 * jj[u]) * rs * ES;
      }
    } else {
#pragma unroll
      for (int u = 0; u < DM; u++) {
#pragma unroll
        for (int q = 0; q < 4; q++) la[u][lane][q] = S(0);
        sj[u] = 0u;
      }
    }
    const uint32_t si = (uint32_t)(8 * i) * rs * ES;  // the row's own block (wave-uniform)
    dv4 G[DM];
#pragma unroll
    for (int u = 0; u < DM; u++) G[u] = dv4{0.0, 0.0, 0.0, 0.0};
    for (int t = 0; t < ntile; t++) {
      const uint32_t so = (uint32_t)t * 32u * ES;
      V4 accA = V4{S(0), S(0), S(0), S(0)}, accB = accA;
      S pv[4];
      if constexpr (FUSED_D) {
#pragma unroll
        for (int k = 0; k < 4; k++) pv[k] = IO::ld1(ba, vld + 8u * k * ES, si + so);
      }
#pragma unroll
      for (int u0 = 0; u0 < DM; u0 += GR) {
        C2 xa0[GR], xa1[GR], xb0[GR], xb1[GR];
        S qv[GR][4];
#pragma unroll
        for (int v = 0; v < GR; v++) {
          if (u0 + v >= DM) continue;  // DM % GR != 0: the last group is short
          const uint32_t sb_ = sj[u0 + v] + so;
          xa0[v] = IO::ld(ba, vlo, sb_);
          xa1[v] = IO::ld(ba, vlo, sb_ + half);
          xb0[v] = IO::ld(bb, vlo, sb_);
          xb1[v] = IO::ld(bb, vlo, sb_ + half);
          if constexpr (FUSED_D) {
#pragma unroll
            for (int k = 0; k < 4; k++) qv[v][k] = IO::ld1(bb, vld + 8u * k * ES, sb_);
          }
        }
#pragma unroll
        for (int v = 0; v < GR; v++) {
          if (u0 + v >= DM) continue;
          S au[4];
#pragma unroll
          for (int q = 0; q < 4; q++) au[q] = la[u0 + v][lane][q];
          accA = mfma_block<S>(au, xa0[v], xa1[v], accA);
          accB = mfma_block<S>(au, xb0[v], xb1[v], accB);
          if constexpr (FUSED_D) {
            V4 gt = V4{S(0), S(0), S(0), S(0)};
#pragma unroll
            for (int k = 0; k < 4; k++) gt = MF::mac(pv[k], qv[v][k], gt);
            G[u0 + v] += dv4{(double)gt.x, (double)gt.y, (double)gt.z, (double)gt.w};
          }
        }
      }
      const C2 ai0 = IO::ld(ba, vlo, si + so), ai1 = IO::ld(ba, vlo, si + so + half);
      const C2 a10 = IO::ld(ba1, vlo, si + so), a11 = IO::ld(ba1, vlo, si + so + half);
      const C2 b20 = IO::ld(bb2, vlo, si + so), b21 = IO::ld(bb2, vlo, si + so + half);
      const S hx[4] = {accA.x, accA.y, accA.z, accA.w};
      const S hb[4] = {accB.x, accB.y, accB.z, accB.w};
      const S ai[4] = {ai0.x, ai0.y, ai1.x, ai1.y};
      const S a1[4] = {a10.x, a10.y, a11.x, a11.y};
      const S b2[4] = {asb2 * b20.x, asb2 * b20.y, asb2 * b21.x, asb2 * b21.y};
      S an[4], bo[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        an[e] = fma(a2, hx[e], -a1[e]);
        S v = fma(ab2, hb[e], -b2[e]);
        v = fma(ac4, ai[e], v);
        v = fma(ac2, a1[e], v);
        v = fma(aw, an[e], v);
        bo[e] = v;
      }
      IO::st(C2{an[0], an[1]}, bao, vlo, si + so);
      IO::st(C2{an[2], an[3]}, bao, vlo, si + so + half);
      IO::st(C2{bo[0], bo[1]}, bbo, vlo, si + so);
      IO::st(C2{bo[2], bo[3]}, bbo, vlo, si + so + half);
    }
    if constexpr (FUSED_D) {
#pragma unroll
      for (int u = 0; u < DM; u++) {
        if (u >= deg) break;
        mfma_flush_D<S>(G[u], m, fD, dacc + (size_t)(p0 + u) * 128);
      }
    }
  }
}

// FP32 row-resident autodiff step on pairs of column tiles (same contract
// and results as mfma_autodiff_rr_dev<float, true>; qp * 4 a multiple of 64,
// i.e. an even number of 16-column tiles).  A lane's two columns of a pair,
// 2c and 2c + 1 of orbital q (and q + 4), are four adjacent floats: one
// 16-byte load feeds the even and the odd MFMA column sets, so each round trip
// of a neighbour group's loads serves 32 columns instead of 16.  The D
// operands pair columns 2n and 2n + 1 (n = 4k + q, k = 0..3), each lane's
// part (re or im) of both.
// Neighbour blocks whose loads the pair-of-tiles step issues together: two
// (24 loads) keep three waves per SIMD; four leave two (C4 FP32: 1561 against
// 1654 us per step).
constexpr int MF_RR2_GROUP = 2;

template <int DM, int GR>
__device__ __forceinline__ void mfma_autodiff_rr2_dev(
    WgId wid, const int* __restrict__ rp, const int* __restrict__ ci,
    const float* __restrict__ hv, const float* __restrict__ am, const float* __restrict__ am1,
    const float* __restrict__ bm1, const float* __restrict__ bm2, float* __restrict__ a_out,
    float* __restrict__ b_out, double* __restrict__ dacc, int nb, int qp, int chunk, double c4,
    double c2, double w, double sb, double sb2, double dscale) {
  using S = float;
  using C2 = Mfma<float>::c2;
  using V4 = Mfma<float>::v4;
  using MF = Mfma<float>;
  using IO = BufC2<float>;
  constexpr int NS = Vec16<S>::N;
  constexpr uint32_t ES = sizeof(S);
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const uint32_t rs = (uint32_t)qp * NS;
  const int npair = (int)(rs / 64);
  const uint32_t vbytes = 8u * (uint32_t)nb * rs * ES;
  const bool vb1 = (bm1 == nullptr);
  const double hb1 = vb1 ? sb : 1.0;
  const S* xbp = vb1 ? am : bm1;  // gathered beta_{m+1} (or alpha_m, scaled by sb)
  const __amdgpu_buffer_rsrc_t ba = vec_rsrc(am, vbytes), bb = vec_rsrc(xbp, vbytes);
  const __amdgpu_buffer_rsrc_t ba1 = vec_rsrc(am1, vbytes);
  const __amdgpu_buffer_rsrc_t bb2 = vec_rsrc(bm2 != nullptr ? bm2 : am1, vbytes);
  const __amdgpu_buffer_rsrc_t bao = vec_rsrc(a_out, vbytes), bbo = vec_rsrc(b_out, vbytes);
  const S a2 = 2.0f, ab2 = (S)(2.0 * hb1), ac4 = (S)c4, ac2 = (S)c2, aw = (S)w;
  const S asb2 = bm2 != nullptr ? 1.0f : (S)sb2;  // beta_{m+2} = asb2 * (bm2 or alpha_{m+1})
  const double fD = dscale * hb1;
  const int dorb = m.c & 7, dpart = m.c >> 3;
  const uint32_t half = 4u * rs * ES;

  __shared__ S lds_a[WG / WAVE][DM][WAVE][4];
  const int lane = threadIdx.x & (WAVE - 1);
  const uint32_t vlo = (m.q * rs + 4u * m.c) * ES;  // columns 2c, 2c + 1 of orbital q
  // D operand of column 2n + h, n = 4k + q: scalar 4n + 2h + dpart of orbital dorb
  const uint32_t vld = (dorb * rs + 4u * m.q + (uint32_t)dpart) * ES;
  S (&la)[DM][WAVE][4] = lds_a[g];

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], deg = rp[i + 1] - p0;
    uint32_t sj[DM];
    if (deg > 0) {
      int jj[DM];
      HRaw<S> hr[DM];
#pragma unroll
      for (int u = 0; u < DM; u++) {
        const int pu = p0 + min(u, deg - 1);
        jj[u] = ci[pu];
        hr[u] = mfma_hraw<S>(hv + (size_t)pu * 128, m);
      }
#pragma unroll
      for (int u = 0; u < DM; u++) {
        S a[4];
        mfma_a_of<S>(hr[u], m, u < deg ? S(1) : S(0), a);
#pragma unroll
        for (int q = 0; q < 4; q++) la[u][lane][q] = a[q];
        sj[u] = (uint32_t)__builtin_amdgcn_readfirstlane(8 * jj[u]) * rs * ES;
      }
    } else {
#pragma unroll
      for (int u = 0; u < DM; u++) {
#pragma unroll
        for (int q = 0; q < 4; q++) la[u][lane][q] = S(0);
        sj[u] = 0u;
      }
    }
    const uint32_t si = (uint32_t)(8 * i) * rs * ES;
    dv4 G[DM];
#pragma unroll
    for (int u = 0; u < DM; u++) G[u] = dv4{0.0, 0.0, 0.0, 0.0};
    for (int t = 0; t < npair; t++) {
      const uint32_t so = (uint32_t)t * 64u * ES;
      V4 accAe = V4{0.f, 0.f, 0.f, 0.f}, accAo = accAe, accBe = accAe, accBo = accAe;
      S pv[4][2];
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int h = 0; h < 2; h++) pv[k][h] = IO::ld1(ba, vld + (16u * k + 2u * h) * ES, si + so);
#pragma unroll
      for (int u0 = 0; u0 < DM; u0 += GR) {
        bf32x4 xa0[GR], xa1[GR], xb0[GR], xb1[GR];
        S qv[GR][4][2];
#pragma unroll
        for (int v = 0; v < GR; v++) {
          if (u0 + v >= DM) continue;  // DM % GR != 0: the last group is short
          const uint32_t sb_ = sj[u0 + v] + so;
          xa0[v] = fkd_raw_ld_f32x4(ba, (int)vlo, (int)sb_, 0);
          xa1[v] = fkd_raw_ld_f32x4(ba, (int)vlo, (int)(sb_ + half), 0);
          xb0[v] = fkd_raw_ld_f32x4(bb, (int)vlo, (int)sb_, 0);
          xb1[v] = fkd_raw_ld_f32x4(bb, (int)vlo, (int)(sb_ + half), 0);
#pragma unroll
          for (int k = 0; k < 4; k++)
#pragma unroll
            for (int h = 0; h < 2; h++) qv[v][k][h] = IO::ld1(bb, vld + (16u * k + 2u * h) * ES, sb_);
        }
        __builtin_amdgcn_sched_barrier(0);  // the group's loads all issued before its products
#pragma unroll
        for (int v = 0; v < GR; v++) {
          if (u0 + v >= DM) continue;
          S au[4];
#pragma unroll
          for (int q = 0; q < 4; q++) au[q] = la[u0 + v][lane][q];
          accAe = mfma_block<S>(au, C2{xa0[v].x, xa0[v].y}, C2{xa1[v].x, xa1[v].y}, accAe);
          accAo = mfma_block<S>(au, C2{xa0[v].z, xa0[v].w}, C2{xa1[v].z, xa1[v].w}, accAo);
          accBe = mfma_block<S>(au, C2{xb0[v].x, xb0[v].y}, C2{xb1[v].x, xb1[v].y}, accBe);
          accBo = mfma_block<S>(au, C2{xb0[v].z, xb0[v].w}, C2{xb1[v].z, xb1[v].w}, accBo);
          V4 gt = V4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int k = 0; k < 4; k++)
#pragma unroll
            for (int h = 0; h < 2; h++) gt = MF::mac(pv[k][h], qv[v][k][h], gt);
          G[u0 + v] += dv4{(double)gt.x, (double)gt.y, (double)gt.z, (double)gt.w};
        }
        // one group's operands live at a time (the scheduler would otherwise
        // hoist every group's loads and leave one wave per SIMD)
        __builtin_amdgcn_sched_barrier(0);
      }
      const bf32x4 ai0 = fkd_raw_ld_f32x4(ba, (int)vlo, (int)(si + so), 0);
      const bf32x4 ai1 = fkd_raw_ld_f32x4(ba, (int)vlo, (int)(si + so + half), 0);
      const bf32x4 a10 = fkd_raw_ld_f32x4(ba1, (int)vlo, (int)(si + so), 0);
      const bf32x4 a11 = fkd_raw_ld_f32x4(ba1, (int)vlo, (int)(si + so + half), 0);
      const bf32x4 b20 = fkd_raw_ld_f32x4(bb2, (int)vlo, (int)(si + so), 0);
      const bf32x4 b21 = fkd_raw_ld_f32x4(bb2, (int)vlo, (int)(si + so + half), 0);
      // column 2c (h = 0) and 2c + 1 (h = 1): (re, im) of orbital q, then q + 4
      const S hx[2][4] = {{accAe.x, accAe.y, accAe.z, accAe.w}, {accAo.x, accAo.y, accAo.z, accAo.w}};
      const S hb[2][4] = {{accBe.x, accBe.y, accBe.z, accBe.w}, {accBo.x, accBo.y, accBo.z, accBo.w}};
      const S ai[2][4] = {{ai0.x, ai0.y, ai1.x, ai1.y}, {ai0.z, ai0.w, ai1.z, ai1.w}};
      const S a1[2][4] = {{a10.x, a10.y, a11.x, a11.y}, {a10.z, a10.w, a11.z, a11.w}};
      const S b2[2][4] = {{b20.x, b20.y, b21.x, b21.y}, {b20.z, b20.w, b21.z, b21.w}};
      S an[2][4], bo[2][4];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          an[h][e] = fma(a2, hx[h][e], -a1[h][e]);
          S v = fma(ab2, hb[h][e], -(asb2 * b2[h][e]));
          v = fma(ac4, ai[h][e], v);
          v = fma(ac2, a1[h][e], v);
          v = fma(aw, an[h][e], v);
          bo[h][e] = v;
        }
      fkd_raw_st_f32x4(bf32x4{an[0][0], an[0][1], an[1][0], an[1][1]}, bao, (int)vlo, (int)(si + so), 0);
      fkd_raw_st_f32x4(bf32x4{an[0][2], an[0][3], an[1][2], an[1][3]}, bao, (int)vlo,
                       (int)(si + so + half), 0);
      fkd_raw_st_f32x4(bf32x4{bo[0][0], bo[0][1], bo[1][0], bo[1][1]}, bbo, (int)vlo, (int)(si + so), 0);
      fkd_raw_st_f32x4(bf32x4{bo[0][2], bo[0][3], bo[1][2], bo[1][3]}, bbo, (int)vlo,
                       (int)(si + so + half), 0);
    }
#pragma unroll
    for (int u = 0; u < DM; u++) {
      if (u >= deg) break;
      mfma_flush_D<S>(G[u], m, fD, dacc + (size_t)(p0 + u) * 128);
    }
  }
}

// Neighbour blocks of one pass of the autodiff kernel: the FP64 G
// accumulators of the fused outer product are held per block of the pass.
constexpr int MF_DMAX = 8;

// ---------------------------------------------------------------------------
// Autodiff backward step on the matrix cores (complex 8x8 blocks); same
// contract as k_step_autodiff (engine_cpu.cpp:278-294):
//   a_out = 2 Hs alpha_m - alpha_{m+1}
//   b_out = 4 c_{2m} alpha_m + 2 c_{2m+1} alpha_{m+1} + 2 Hs beta_{m+1}
//           - beta_{m+2} + w alpha_{m-1}
//   D    += dscale alpha_m beta_{m+1}^+ on H's pattern (FUSED_D)
// bm1 == null: beta_{m+1} = sb alpha_m; bm2 == null: beta_{m+2} = sb2 alpha_{m+1}.
// A block row with more than MF_DMAX blocks takes several passes over its
// tiles; between passes the partial products Hs alpha_m, Hs beta_{m+1} are
// parked in a_out / b_out, which only this wave touches.
// ---------------------------------------------------------------------------
template <typename S, bool FUSED_D>
__device__ __forceinline__ void mfma_autodiff_dev(
    WgId wid,
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  using C2 = typename Mfma<S>::c2;
  using V4 = typename Mfma<S>::v4;
  using MF = Mfma<S>;
  constexpr int NS = Vec16<S>::N;
  (void)nt;
  const MfmaLane m = mfma_lane<S>();
  const int g = threadIdx.x / WAVE;
  const RowRange rr = wg_rows(nb, chunk, WG / WAVE, wid);
  const size_t rs = (size_t)qp * NS;
  const int ntile = (int)(rs / 32);
  const bool vb1 = (bm1 == nullptr);
  const double hb1 = vb1 ? sb : 1.0;
  const S* __restrict__ xb = vb1 ? am : bm1;  // gathered beta_{m+1} (or alpha_m, scaled by sb)
  const S a2 = (S)2, ab2 = (S)(2.0 * hb1), ac4 = (S)c4, ac2 = (S)c2, aw = (S)w, asb2 = (S)sb2;
  const double fD = dscale * hb1;
  // D-operand layout: lane l holds row l % 16 (orbital c & 7, part c >> 3)
  // of column 4 kappa + q of a tile
  const int dorb = m.c & 7, dpart = m.c >> 3;

  for (int base = rr.r0; base < rr.r1; base += rr.stride) {
    const int i = __builtin_amdgcn_readfirstlane(base + g);
    if (i >= rr.r1) continue;
    const int p0 = rp[i], p1 = rp[i + 1];
    const size_t o0 = (size_t)(8 * i + m.q) * rs, o1 = (size_t)(8 * i + 4 + m.q) * rs;
    const size_t od = (size_t)(8 * i + dorb) * rs + dpart;
    for (int pc = p0; pc < p1 || pc == p0; pc += MF_DMAX) {
      const int pe = min(pc + MF_DMAX, p1);
      const bool first = (pc == p0), last = (pe == p1);
      dv4 G[MF_DMAX];
#pragma unroll
      for (int u = 0; u < MF_DMAX; u++) G[u] = dv4{0.0, 0.0, 0.0, 0.0};
      for (int t = 0; t < ntile; t++) {
        const size_t co = (size_t)t * 32 + 2 * m.c;
        const size_t f0 = o0 + co, f1 = o1 + co;
        V4 accA = V4{S(0), S(0), S(0), S(0)}, accB = accA;
        if (!first) {  // partial products parked by the previous pass
          const C2 pa0 = ldc<S>(a_out + f0), pa1 = ldc<S>(a_out + f1);
          const C2 pb0 = ldc<S>(b_out + f0), pb1 = ldc<S>(b_out + f1);
          accA = V4{pa0.x, pa0.y, pa1.x, pa1.y};
          accB = V4{pb0.x, pb0.y, pb1.x, pb1.y};
        }
        S pv[4];
        if constexpr (FUSED_D) {
#pragma unroll
          for (int k = 0; k < 4; k++) pv[k] = am[od + 2 * ((size_t)t * 16 + 4 * k + m.q)];
        }
#pragma unroll
        for (int u0 = 0; u0 < MF_DMAX; u0 += 2) {
          if (pc + u0 >= pe) break;
          S a[2][4];
          C2 xa0[2], xa1[2], xb0[2], xb1[2];
          S qv[2][4];
#pragma unroll
          for (int v = 0; v < 2; v++) {
            const int p = pc + u0 + v;
            if (p < pe) {
              const int j = ci[p];
              mfma_a_of<S>(mfma_hraw<S>(hv + (size_t)p * 128, m), m, S(1), a[v]);
              const size_t r0 = (size_t)(8 * j + m.q) * rs + co, r1 = (size_t)(8 * j + 4 + m.q) * rs + co;
              xa0[v] = ldc<S>(am + r0);
              xa1[v] = ldc<S>(am + r1);
              xb0[v] = ldc<S>(xb + r0);
              xb1[v] = ldc<S>(xb + r1);
              if constexpr (FUSED_D) {
                const size_t jd = (size_t)(8 * j + dorb) * rs + dpart;
#pragma unroll
                for (int k = 0; k < 4; k++) qv[v][k] = xb[jd + 2 * ((size_t)t * 16 + 4 * k + m.q)];
              }
            } else {
#pragma unroll
              for (int k = 0; k < 4; k++) {
                a[v][k] = S(0);
                qv[v][k] = S(0);
              }
              xa0[v] = xa1[v] = xb0[v] = xb1[v] = C2{S(0), S(0)};
            }
          }
#pragma unroll
          for (int v = 0; v < 2; v++) {
            if (pc + u0 + v >= pe) continue;
            accA = mfma_block<S>(a[v], xa0[v], xa1[v], accA);
            accB = mfma_block<S>(a[v], xb0[v], xb1[v], accB);
            if constexpr (FUSED_D) {
              V4 gt = V4{S(0), S(0), S(0), S(0)};
#pragma unroll
              for (int k = 0; k < 4; k++) gt = MF::mac(pv[k], qv[v][k], gt);
              G[u0 + v] += dv4{(double)gt.x, (double)gt.y, (double)gt.z, (double)gt.w};
            }
          }
        }
        if (!last) {
          stc<S>(a_out + f0, C2{accA.x, accA.y});
          stc<S>(a_out + f1, C2{accA.z, accA.w});
          stc<S>(b_out + f0, C2{accB.x, accB.y});
          stc<S>(b_out + f1, C2{accB.z, accB.w});
          continue;
        }
        const C2 ai0 = ldc<S>(am + f0), ai1 = ldc<S>(am + f1);
        const C2 a10 = ldc<S>(am1 + f0), a11 = ldc<S>(am1 + f1);
        C2 b20, b21;
        if (bm2 != nullptr) {
          b20 = ldc<S>(bm2 + f0);
          b21 = ldc<S>(bm2 + f1);
        } else {
          b20 = C2{asb2 * a10.x, asb2 * a10.y};
          b21 = C2{asb2 * a11.x, asb2 * a11.y};
        }
        const S hx[4] = {accA.x, accA.y, accA.z, accA.w};
        const S hb[4] = {accB.x, accB.y, accB.z, accB.w};
        const S ai[4] = {ai0.x, ai0.y, ai1.x, ai1.y};
        const S a1[4] = {a10.x, a10.y, a11.x, a11.y};
        const S b2[4] = {b20.x, b20.y, b21.x, b21.y};
        S an[4], bo[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          an[e] = fma(a2, hx[e], -a1[e]);
          S v = fma(ab2, hb[e], -b2[e]);
          v = fma(ac4, ai[e], v);
          v = fma(ac2, a1[e], v);
          v = fma(aw, an[e], v);
          bo[e] = v;
        }
        stc<S>(a_out + f0, C2{an[0], an[1]});
        stc<S>(a_out + f1, C2{an[2], an[3]});
        stc<S>(b_out + f0, C2{bo[0], bo[1]});
        stc<S>(b_out + f1, C2{bo[2], bo[3]});
      }
      if constexpr (FUSED_D && sizeof(S) == 4) {
        // f32 layout: lane (q, c) holds G[4q + i][c]; lanes q < 2, c < 8 own
        // D(a = 4q + i, b = c)
        const bool owner = (m.q < 2) && (m.c < 8);
#pragma unroll
        for (int u = 0; u < MF_DMAX; u++) {
          if (pc + u >= pe) break;
          double dre[4], dim[4];
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const double gv = G[u][e];
            const double g40 = __shfl_xor(gv, 40);  // G[8 + a][8 + b]
            const double g32 = __shfl_xor(gv, 32);  // G[8 + a][b]
            const double g08 = __shfl_xor(gv, 8);   // G[a][8 + b]
            dre[e] = gv + g40;
            dim[e] = g32 - g08;
          }
          if (owner) {
            double* d = dacc + ((size_t)(pc + u) * 64 + m.c * 8 + 4 * m.q) * 2;
#pragma unroll
            for (int e = 0; e < 4; e++) {
              d[2 * e] = fma(fD, dre[e], d[2 * e]);
              d[2 * e + 1] = fma(fD, dim[e], d[2 * e + 1]);
            }
          }
        }
      } else if constexpr (FUSED_D) {
        // f64 layout: lane (q, c) holds G[q + 4i][c]: G[a][c] (i = 0, 1) and
        // G[8 + a][c] (i = 2, 3) for a = q + 4i; lanes c < 8 own D(a, b = c)
        const bool owner = m.c < 8;
#pragma unroll
        for (int u = 0; u < MF_DMAX; u++) {
          if (pc + u >= pe) break;
          double dre[2], dim[2];
#pragma unroll
          for (int e = 0; e < 2; e++) {
            const double gr = G[u][e], gi = G[u][e + 2];
            const double xr = __shfl_xor(gr, 8);  // G[a][8 + b]
            const double xi = __shfl_xor(gi, 8);  // G[8 + a][8 + b]
            dre[e] = gr + xi;
            dim[e] = gi - xr;
          }
          if (owner) {
            double* d = dacc + ((size_t)(pc + u) * 64 + m.c * 8 + m.q) * 2;
#pragma unroll
            for (int e = 0; e < 2; e++) {
              d[8 * e] = fma(fD, dre[e], d[8 * e]);
              d[8 * e + 1] = fma(fD, dim[e], d[8 * e + 1]);
            }
          }
        }
      }
      if (p1 == p0) break;
    }
  }
}

template <typename S, bool FUSED_D>
__global__ __launch_bounds__(WG) void k_mfma_autodiff(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  mfma_autodiff_dev<S, FUSED_D>(wg_self(), rp, ci, hv, am, am1, bm1, bm2, a_out, b_out, dacc, nb, qp, chunk, c4, c2, w, sb, sb2, dscale, nt);
}

// DM = 8 (MF_DM), or 7 for rows of at most 7 blocks (a cubic lattice's six
// neighbours and the site: no padding slot in the matrix-core work).
template <typename S, int DM>
__global__ __launch_bounds__(WG) void k_mfma_step_rr(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ x, const S* __restrict__ z, S* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<S> rsrc) {
  mfma_step_rr_dev<S, DM, 4>(wg_self(), rp, ci, hv, x, z, y, nb, qp, chunk, cx, cz, partial, nt,
                             rsrc);
}

template <typename S, int DM>
__global__ __launch_bounds__(WG) void k_mfma_autodiff_rr(
    const int* __restrict__ rp, const int* __restrict__ ci, const S* __restrict__ hv,
    const S* __restrict__ am, const S* __restrict__ am1, const S* __restrict__ bm1,
    const S* __restrict__ bm2, S* __restrict__ a_out, S* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  mfma_autodiff_rr_dev<S, true, DM, mf_rr_group<S, DM>()>(wg_self(), rp, ci, hv, am, am1, bm1, bm2,
                                                         a_out, b_out, dacc, nb, qp, chunk, c4, c2,
                                                         w, sb, sb2, dscale, nt);
}

// FP32 forward / Clenshaw step on pairs of column tiles (mfma_step_rr2_dev);
// the same arguments as k_mfma_step_rr.
template <int DM>
__global__ __launch_bounds__(WG) void k_mfma_step_rr2(
    const int* __restrict__ rp, const int* __restrict__ ci, const float* __restrict__ hv,
    const float* __restrict__ x, const float* __restrict__ z, float* __restrict__ y, int nb, int qp,
    int chunk, double cx, double cz, double* __restrict__ partial, int nt, RSource<float> rsrc) {
  (void)nt;
  mfma_step_rr2_dev<DM, 4>(wg_self(), rp, ci, hv, x, z, y, nb, qp, chunk, cx, cz, partial, rsrc);
}

// FP32 on pairs of column tiles (mfma_autodiff_rr2_dev; qp * 4 a multiple of
// 64); the same arguments as k_mfma_autodiff_rr.  Three waves per SIMD: the
// register budget that keeps them (168) also keeps the G accumulators and the
// group's operands out of the AGPRs.
template <int DM, int GR>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) void k_mfma_autodiff_rr2(
    const int* __restrict__ rp, const int* __restrict__ ci, const float* __restrict__ hv,
    const float* __restrict__ am, const float* __restrict__ am1, const float* __restrict__ bm1,
    const float* __restrict__ bm2, float* __restrict__ a_out, float* __restrict__ b_out,
    double* __restrict__ dacc, int nb, int qp, int chunk, double c4, double c2, double w,
    double sb, double sb2, double dscale, int nt) {
  (void)nt;
  mfma_autodiff_rr2_dev<DM, GR>(wg_self(), rp, ci, hv, am, am1, bm1, bm2, a_out, b_out, dacc, nb,
                                qp, chunk, c4, c2, w, sb, sb2, dscale);
}

}  // namespace fkd
