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
// gridDim.y-th slice of K; its four waves own 32 x 32 quarters, i.e. 2 x 2
// tiles of 16 x 16, each accumulated with v_mfma_{f64,f32}_16x16x4:
//   Re C += Ar^T Br + Ai^T Bi,   Im C += Ar^T Bi - Ai^T Br.
// A lane (i, g) = (lane % 16, lane / 16) feeds element 4 kk + g of column i
// of each tile to the MFMA of k-group kk.  Grid rows write FP64 partial blocks
// that a second kernel adds in a fixed order (deterministic).
#pragma once

#include <type_traits>

#include "kpm_mfma.h"

namespace fkd {

constexpr int DOT_BM = 64;  // output block of one workgroup (DOT_BM x DOT_BM)

// The K range [kbeg, kend) is cut into slices of kslice; the workgroups with
// grid row y take slices y, y + ny, y + 2 ny, ... (ny = gridDim.y), so the
// workgroups resident at one time work on one narrow window of K.  A, B:
// column m at A + m * ld * HS (ld >= kend).  part: ny blocks of M * M (* 2 if
// complex) doubles, row-major in (m1, m2).  FP32 sums restart at every slice
// and are added to FP64 registers, so FP32 accumulation never spans more than
// kslice terms.
//
// The operands are staged through LDS.  Per step of DOT_KT = 16 elements of K
// the workgroup copies the 64 columns of its A and B blocks into LDS: a load
// instruction covers 4 columns x 16 consecutive elements (256 B runs: few
// columns, and so few address translations, per instruction), thread t
// holding element t % 16 of columns t / 16 + 16 u, u = 0..3.  LDS rows are k-major, [k][column] with 65 entries per row (the
// 16-byte stores of consecutive k then fall on distinct banks, the MFMA
// operand reads of 8 consecutive columns are contiguous).  The next step's
// global loads are issued before the current step's MFMAs.
constexpr int DOT_KT = 16;   // K elements per LDS step
constexpr int DOT_LDW = 65;  // LDS row: 64 columns + 1 pad

template <typename S, bool CX>
__global__ __launch_bounds__(WG) void k_dot_block(const S* __restrict__ A,
                                                   const S* __restrict__ B, long long ld,
                                                   long long kbeg, long long kend,
                                                   long long kslice, int M,
                                                   double* __restrict__ part) {
  using MF = Mfma<S>;
  using v4 = typename MF::v4;
  using E = typename std::conditional<CX, typename MF::c2, S>::type;  // one element
  constexpr int HS = CX ? 2 : 1;
  constexpr bool F32 = sizeof(S) == 4;
  __shared__ E la[DOT_KT][DOT_LDW], lb[DOT_KT][DOT_LDW];
  const int nbm = (M + DOT_BM - 1) / DOT_BM;
  const int bm1 = blockIdx.x / nbm, bm2 = blockIdx.x % nbm;
  const int t = threadIdx.x, wave = t / WAVE, lane = t % WAVE;
  const int i = lane & 15, g = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // this wave's quarter
  // loader: element t % 16 of columns t / 16 + 16 u of the A and B blocks
  const int le = t & 15;
  const S* pa[4];
  const S* pb[4];
  bool va[4], vb[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int c = (t >> 4) + 16 * u;
    const int ca = bm1 * DOT_BM + c, cb = bm2 * DOT_BM + c;
    va[u] = ca < M;
    vb[u] = cb < M;
    pa[u] = A + (size_t)(va[u] ? ca : 0) * (size_t)ld * HS;
    pb[u] = B + (size_t)(vb[u] ? cb : 0) * (size_t)ld * HS;
  }
  // raw loads (always from a valid address) and their validity; the values
  // past k1 or M are replaced by zeros only when stored to LDS, so nothing
  // waits for the loads before then
  E ra[4], rb[4];
  bool rok = false;
  auto gload = [&](long long kt, long long k1) {
    const long long k = kt + le;
    rok = k < k1;
    const size_t o = (size_t)(k < ld ? k : ld - 1) * HS;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      ra[u] = *reinterpret_cast<const E*>(pa[u] + o);
      rb[u] = *reinterpret_cast<const E*>(pb[u] + o);
    }
  };
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
  for (long long k0 = kbeg + (long long)blockIdx.y * kslice; k0 < kend;
       k0 += (long long)gridDim.y * kslice) {
    const long long k1 = k0 + kslice < kend ? k0 + kslice : kend;
    gload(k0, k1);
    for (long long kt = k0; kt < k1; kt += DOT_KT) {
      __syncthreads();  // every wave has read the previous step's tiles
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int c = (t >> 4) + 16 * u;
        la[le][c] = (va[u] && rok) ? ra[u] : E{};
        lb[le][c] = (vb[u] && rok) ? rb[u] : E{};
      }
      __syncthreads();
      gload(kt + DOT_KT, k1);  // past k1 on the last step: zeros, never stored
      // keep those loads here, ahead of the MFMAs they overlap
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int kk = 0; kk < DOT_KT / 4; kk++) {
        const int kr = kk * 4 + g;
        E xa[2], xb[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
          xa[q] = la[kr][wm + 16 * q + i];
          xb[q] = lb[kr][wn + 16 * q + i];
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            if constexpr (CX) {
              cre[a][b] = MF::mac(xa[a].x, xb[b].x, cre[a][b]);
              cre[a][b] = MF::mac(xa[a].y, xb[b].y, cre[a][b]);
              cim[a][b] = MF::mac(xa[a].x, xb[b].y, cim[a][b]);
              cim[a][b] = MF::mac(-xa[a].y, xb[b].x, cim[a][b]);
            } else {
              cre[a][b] = MF::mac(xa[a], xb[b], cre[a][b]);
            }
          }
      }
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
        const int m1 = bm1 * DOT_BM + wm + 16 * a + r, m2 = bm2 * DOT_BM + wn + 16 * b + i;
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
