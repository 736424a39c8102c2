// Site-local twisted clover block of one chirality, shared by the block kernels
// (clover.hip) and the clover epilogues of the hop kernels (tm_hop.hip), so that the
// fused and the separate applies run the same arithmetic.
//
// Clover field of one parity: 36 complex components per site (G, H, I, J 3x3 blocks
// of nissa's clover_term_t, clover_term.cpp:22-55, k = 9*b + 3*row + col); inverse
// field: 72 complex per site (k = 36 h + 6 row + col, row/col = 3 spin + colour,
// inv_clover_term_t[h][id][ic][x_id][x_ic]).  Storage: chirality h (N = 18 or 36
// components) is a packed array of 16-byte records, [record][site], of 16/EB
// components each (Pack<EB, N>; EB = 16 FP64, 8 FP32, 4 FP16) starting at byte
// h N EB stride, so that a site's block takes N EB / 16 record loads.  For FP64
// that is component k of device site e at byte (k stride + e) 16.
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

// the N components of chirality h of site e (64-bit addressing: a block field
// can exceed the 4 GB reach of a buffer offset)
template <int MODE, class R, int EB>
__device__ __forceinline__ void clover_load(const void* B, int64_t st, uint32_t e, int h, ClovBlk<MODE, R>& k)
{
  constexpr int N = MODE == 0 ? 18 : 36;
  using P = Pack<EB, N>;
  static_assert(P::NL == 0 || (EB == 4 && P::NL == 2), "block layout");
  const char* base = (const char*)B + (int64_t)h * N * EB * st;
#pragma unroll
  for (int q = 0; q < P::NG; q++)
    decode_record<EB, R>(*(const nq_u4*)(base + ((int64_t)q * st + e) * 16), k.m + q * P::G);
  if constexpr (P::NL == 2) {
    const nq_hv4 t = *(const nq_hv4*)(base + (int64_t)P::NG * st * 16 + (int64_t)e * 8);
    k.m[N - 2] = mk<R>((R)(float)t[0], (R)(float)t[1]);
    k.m[N - 1] = mk<R>((R)(float)t[2], (R)(float)t[3]);
  }
}

// chirality h of an FP64 block field ([k][e] complex, N per chirality) stored in the
// packed layout above with element bytes EB (the sloppy copies); false when an FP16
// copy would overflow (a real part or imaginary part above 65000 in magnitude)
template <int MODE, int EB>
__device__ __forceinline__ bool clover_pack(const cx<double>* in, void* out, int64_t st, uint32_t e, int h)
{
  constexpr int N = MODE == 0 ? 18 : 36;
  cx<double> v[N];
  double big = 0;
#pragma unroll
  for (int q = 0; q < N; q++) {
    v[q] = in[(int64_t)(N * h + q) * st + e];
    big = fmax(big, fmax(fabs(v[q].re), fabs(v[q].im)));
  }
  store_packed<EB, N, 0, double>((char*)out + (int64_t)h * N * EB * st, st, e, v);
  return EB != 4 || big <= 65000.0;
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

template <int MODE, int DAG, class R, int EB>
__device__ __forceinline__ void clover_half(const void* B, int64_t st, uint32_t e, int h, double kappa, double mu,
                                            const cx<R> iv[6], cx<R> ov[6])
{
  ClovBlk<MODE, R> k;
  clover_load<MODE, R, EB>(B, st, e, h, k);
  clover_apply_blk<MODE, DAG, R>(k, h, kappa, mu, iv, ov);
}

}  // namespace nq
