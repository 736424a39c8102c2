// Site-local twisted clover block of one chirality, shared by the block kernels
// (clover.hip) and the clover epilogues of the hop kernels (tm_hop.hip), so that the
// fused and the separate applies run the same arithmetic.
//
// Clover field of one parity: 36 complex components per site (G, H, I, J 3x3 blocks
// of nissa's clover_term_t, clover_term.cpp:22-55, k = 9*b + 3*row + col); inverse
// field: 72 complex per site (k = 36 h + 6 row + col, row/col = 3 spin + colour,
// inv_clover_term_t[h][id][ic][x_id][x_ic]); component k of device site e at
// B[k*stride + e] in FP64 or FP32 complex.
#pragma once

#include "nq_device.hpp"

namespace nq {

// One chirality h of the site-local block: MODE 0 applies (1/2kappa +- i mu) +
// the clover term, z + [[C0, C1^+], [C1, -C0]]; MODE 1 the inverse block (the
// dagger uses the daggered block of chirality 0 and the plain one of chirality 1).
// The block of a chirality is loaded once (ClovBlk) and can be applied to several
// vectors; clover_half does both for one vector.
template <int MODE, class R>
struct ClovBlk {
  cx<R> m[MODE == 0 ? 18 : 36];  // MODE 0: C0 (9) then C1 (9); MODE 1: the 6x6 inverse
};

template <int MODE, class R>
__device__ __forceinline__ void clover_load(const cx<R>* B, int64_t st, uint32_t e, int h, ClovBlk<MODE, R>& k)
{
  if constexpr (MODE == 0) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
      k.m[q] = B[(18 * h + q) * st + e];
      k.m[9 + q] = B[(18 * h + 9 + q) * st + e];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 36; q++) k.m[q] = B[(36 * h + q) * st + e];
  }
}

template <int MODE, int DAG, class R>
__device__ __forceinline__ void clover_apply_blk(const ClovBlk<MODE, R>& k, int h, double kappa, double mu,
                                                 const cx<R> iv[6], cx<R> ov[6])
{
  typedef cx<R> CR;
  if constexpr (MODE == 0) {
    const double m = (DAG ? -mu : mu) * (h == 0 ? 1.0 : -1.0);
    const CR z = mk<R>((R)(1 / (2 * kappa)), (R)m);
    const CR* c0 = k.m;
    const CR* c1 = k.m + 9;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      CR s0 = z * iv[a], s1 = z * iv[3 + a];
#pragma unroll
      for (int b = 0; b < 3; b++) {
        s0 = s0 + c0[3 * a + b] * iv[b] + conjg(c1[3 * b + a]) * iv[3 + b];
        s1 = s1 + c1[3 * a + b] * iv[b] - c0[3 * a + b] * iv[3 + b];
      }
      ov[a] = s0;
      ov[3 + a] = s1;
    }
  } else {
    constexpr bool hd0 = DAG != 0;  // chirality 0 uses the daggered block for the dagger
    const bool hd = h == 0 ? hd0 : !hd0;
    const CR* M = k.m;
#pragma unroll
    for (int r = 0; r < 6; r++) {
      CR sum = mk<R>(0, 0);
#pragma unroll
      for (int c = 0; c < 6; c++) sum = sum + (hd ? conjg(M[6 * c + r]) : M[6 * r + c]) * iv[c];
      ov[r] = sum;
    }
  }
}

template <int MODE, int DAG, class R>
__device__ __forceinline__ void clover_half(const cx<R>* B, int64_t st, uint32_t e, int h, double kappa, double mu,
                                            const cx<R> iv[6], cx<R> ov[6])
{
  ClovBlk<MODE, R> k;
  clover_load<MODE, R>(B, st, e, h, k);
  clover_apply_blk<MODE, DAG, R>(k, h, kappa, mu, iv, ov);
}

}  // namespace nq
