// kpm_dots.h -- the M x M block of dots C = A^H B of two K x M matrices on
// the matrix cores.
//
// The two-current moments mu[m1][m2] = sum_k <alpha_m1(k), j2 atild_m2(k)>
// keep, for a pass over probe columns, all M vectors alpha_m and all M
// vectors j2 atild_m as the columns of two K x M matrices A and B (column m
// contiguous, K = rows x padded columns of one vector).  C = A^H B is tall and
// skinny in K (K ~ 1e7, M ~ 1e2): a library GEMM tiles only the M x M output
// and leaves most of the chip idle, so here the K range is split over the
// grid as well.  Each workgroup owns a 64 x 64 output block and every
// gridDim.y-th slice of K;
// its four waves own 32 x 32 quarters, i.e. 2 x 2 tiles of 16 x 16, each
// accumulated with v_mfma_{f64,f32}_16x16x4:
//   Re C += Ar^T Br + Ai^T Bi,   Im C += Ar^T Bi - Ai^T Br.
// A lane (i, g) = (lane % 16, lane / 16) reads elements k + 2g, k + 2g + 1
// of column i of each tile (the contraction index k of an MFMA is only
// summed over, so any assignment of k to lane groups works as long as A and
// B agree).  Grid rows write FP64 partial blocks that a second kernel adds in
// a fixed order (deterministic).
#pragma once

#include "kpm_mfma.h"

namespace fkd {

constexpr int DOT_BM = 64;  // output block of one workgroup (DOT_BM x DOT_BM)
constexpr int DOT_KU = 8;   // k per loop trip: two per lane, four lane groups

// The K range [kbeg, kend) is cut into slices of kslice; the workgroups with
// grid row y take slices y, y + ny, y + 2 ny, ... (ny = gridDim.y), so the
// workgroups resident at one time work on one narrow window of K: each column
// is touched within a few MB, which keeps the address translations of the
// 2 x 64 columns of a block in the TLBs.  A, B: column m at A + m * ld * HS
// (ld >= kend).  part: ny blocks of M * M (* 2 if complex) doubles,
// row-major in (m1, m2).  FP32 sums restart at every slice and are added to
// FP64 registers, so FP32 accumulation never spans more than kslice terms.
template <typename S, bool CX>
__global__ __launch_bounds__(WG) void k_dot_block(const S* __restrict__ A,
                                                   const S* __restrict__ B, long long ld,
                                                   long long kbeg, long long kend,
                                                   long long kslice, int M,
                                                   double* __restrict__ part) {
  using MF = Mfma<S>;
  using v4 = typename MF::v4;
  constexpr int HS = CX ? 2 : 1;
  constexpr bool F32 = sizeof(S) == 4;
  const int nbm = (M + DOT_BM - 1) / DOT_BM;
  const int bm1 = blockIdx.x / nbm, bm2 = blockIdx.x % nbm;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int i = lane & 15, g = lane >> 4;
  const int m1w = bm1 * DOT_BM + (wave >> 1) * 32;
  const int m2w = bm2 * DOT_BM + (wave & 1) * 32;
  const S* pa[2];
  const S* pb[2];
  bool va[2], vb[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int ca = m1w + 16 * t + i, cb = m2w + 16 * t + i;
    va[t] = ca < M;
    vb[t] = cb < M;
    pa[t] = A + (size_t)(va[t] ? ca : 0) * (size_t)ld * HS;
    pb[t] = B + (size_t)(vb[t] ? cb : 0) * (size_t)ld * HS;
  }
  v4 cre[2][2], cim[2][2];
  dv4 sre[2][2], sim[2][2];  // FP64 sums of the FP32 slices
  auto zero = [&]() {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          cre[a][b][v] = S(0);
          cim[a][b][v] = S(0);
        }
  };
  zero();
  if constexpr (F32) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          sre[a][b][v] = 0.0;
          sim[a][b][v] = 0.0;
        }
  }
  // One trip: elements k = kb + 2 g + e (e = 0, 1) of the lane's columns.
  // The loads are unconditional (the index clamped into the column, a
  // column past M read as column 0) and the values past k1 or M replaced by
  // zeros, so the trip count and the control flow are the same in every lane
  // (MFMAs need the whole wave).  The next trip's loads are issued before the
  // current trip's MFMAs.
  struct Trip {
    S ar[2][2], ai[2][2], br[2][2], bi[2][2];  // [tile][element]
  };
  auto load = [&](long long kb, long long k1, Trip& d) {
    const long long ke = kb + 2 * g;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const bool ok = ke + e < k1;
      const size_t o = (size_t)(ke + e < ld ? ke + e : ld - 1) * HS;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        if constexpr (CX) {
          const typename MF::c2 x = *reinterpret_cast<const typename MF::c2*>(pa[t] + o);
          const typename MF::c2 y = *reinterpret_cast<const typename MF::c2*>(pb[t] + o);
          d.ar[t][e] = (va[t] && ok) ? x.x : S(0);
          d.ai[t][e] = (va[t] && ok) ? x.y : S(0);
          d.br[t][e] = (vb[t] && ok) ? y.x : S(0);
          d.bi[t][e] = (vb[t] && ok) ? y.y : S(0);
        } else {
          const S x = pa[t][o], y = pb[t][o];
          d.ar[t][e] = (va[t] && ok) ? x : S(0);
          d.br[t][e] = (vb[t] && ok) ? y : S(0);
        }
      }
    }
  };
  for (long long k0 = kbeg + (long long)blockIdx.y * kslice; k0 < kend;
       k0 += (long long)gridDim.y * kslice) {
    const long long k1 = k0 + kslice < kend ? k0 + kslice : kend;
    Trip cur, nxt;
    load(k0, k1, cur);
    for (long long kb = k0; kb < k1; kb += DOT_KU) {
      load(kb + DOT_KU, k1, nxt);  // past k1 on the last trip: zeros, never used
#pragma unroll
      for (int e = 0; e < 2; e++)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            cre[a][b] = MF::mac(cur.ar[a][e], cur.br[b][e], cre[a][b]);
            if constexpr (CX) {
              cre[a][b] = MF::mac(cur.ai[a][e], cur.bi[b][e], cre[a][b]);
              cim[a][b] = MF::mac(cur.ar[a][e], cur.bi[b][e], cim[a][b]);
              cim[a][b] = MF::mac(-cur.ai[a][e], cur.br[b][e], cim[a][b]);
            }
          }
      cur = nxt;
    }
    if constexpr (F32) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            sre[a][b][v] += (double)cre[a][b][v];
            sim[a][b][v] += (double)cim[a][b][v];
          }
      zero();
    }
  }
  double* out = part + (size_t)blockIdx.y * M * M * HS;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        // accumulator row of register v: f64 C[g + 4 v][i], f32 C[4 g + v][i]
        const int r = F32 ? 4 * g + v : g + 4 * v;
        const int m1 = m1w + 16 * a + r, m2 = m2w + 16 * b + i;
        if (m1 < M && m2 < M) {
          const size_t o = ((size_t)m1 * M + m2) * HS;
          out[o] = F32 ? sre[a][b][v] : (double)cre[a][b][v];
          if constexpr (CX) out[o + 1] = F32 ? sim[a][b][v] : (double)cim[a][b][v];
        }
      }
}

// acc[x] = (first ? 0 : acc[x]) + sum over y of part[y * cnt + x], y in order.
__global__ __launch_bounds__(WG) void k_dot_reduce(const double* __restrict__ part, int ny,
                                                    size_t cnt, double* __restrict__ acc,
                                                    int first) {
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < cnt;
       x += (size_t)gridDim.x * blockDim.x) {
    double s = first ? 0.0 : acc[x];
    for (int y = 0; y < ny; y++) s += part[(size_t)y * cnt + x];
    acc[x] = s;
  }
}

}  // namespace fkd
