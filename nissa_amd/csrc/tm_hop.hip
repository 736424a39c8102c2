// Twisted-mass Wilson hopping term of one parity with fused epilogues, for gfx950.
//
// Reference: tmn2Deo_or_tmn2Doe_eos (src/dirac_operators/tmD_eoprec/
// dirac_operator_tmD_eoprec_portable.hpp:36-169):
//   out(x) = sum_mu (1 - g_mu) U_mu(x) in(x+mu) + (1 + g_mu) U_mu(x-mu)^+ in(x-mu)
// in nissa's gamma basis (src/new_types/dirac.cpp:43-58).  The spin projections
// and the way the two projected colour vectors are scattered back onto the four
// spins are fixed by that basis, so the per-direction table below necessarily
// matches the reference's (same algebra, written for our accumulators).
//
// One thread per output site; links are streamed from the per-output-parity
// slot arrays (backward links pre-daggered at upload), neighbour spinors are
// gathered from the opposite-parity SoA field (L2/MALL-resident).  The epilogue
// applies the site-local pieces of the eo operator in registers:
//   out_h = c1_h hop + c2_h aux + c3_h aux2    (h = upper/lower spin pair)
// which covers tmn2D, inv_tmDee(hop), g5(Doo in - hop/4), g5(eta + hop/2),
// Dee^-1(eta + hop/2), and Q = g5 D on each parity; mode 2 fuses the CG update
// x += alpha p, r -= alpha s and |r|^2; mode 3 writes a residual b - A x in FP64
// plus its sloppy copy and |r|^2.

#include <hip/hip_ext.h>

#include <cstdlib>
#include <string>

#include "clover_block.hpp"

namespace nq {

// Hop launches: with the ledger on (ctx->hop_ev set by prof_begin_ext) the launch
// records the kernel's own start and end, so that the ledger's time per launch is
// the kernel's, as in a profiler's kernel trace, without the dispatch gap that two
// stream events around the launch would add.
#define NQ_HOP_LAUNCH(KERNEL, GRID, BLOCK, ...)                                                          \
  do {                                                                                                  \
    if (ctx->hop_ev[0])                                                                                 \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, ctx->stream, ctx->hop_ev[0], ctx->hop_ev[1], 0, __VA_ARGS__); \
    else                                                                                                \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, ctx->stream, __VA_ARGS__);                             \
  } while (0)

template <class R>
struct Half {
  cx<R> h0[3], h1[3];
};

// spin projection of the neighbour spinor for slot (mu, fwd/bwd)
template <class R, int MU, bool FWD>
__device__ __forceinline__ void project(const cx<R> s[12], cx<R> h0[3], cx<R> h1[3])
{
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<R> s0 = s[c], s1 = s[3 + c], s2 = s[6 + c], s3 = s[9 + c];
    if constexpr (FWD) {  // (1 - gamma_mu)
      if constexpr (MU == 0) { h0[c] = s0 + s2; h1[c] = s1 + s3; }
      if constexpr (MU == 1) { h0[c] = iadd(s0, s3); h1[c] = iadd(s1, s2); }
      if constexpr (MU == 2) { h0[c] = s0 + s3; h1[c] = s1 - s2; }
      if constexpr (MU == 3) { h0[c] = iadd(s0, s2); h1[c] = isub(s1, s3); }
    } else {  // (1 + gamma_mu)
      if constexpr (MU == 0) { h0[c] = s0 - s2; h1[c] = s1 - s3; }
      if constexpr (MU == 1) { h0[c] = isub(s0, s3); h1[c] = isub(s1, s2); }
      if constexpr (MU == 2) { h0[c] = s0 - s3; h1[c] = s1 + s2; }
      if constexpr (MU == 3) { h0[c] = isub(s0, s2); h1[c] = iadd(s1, s3); }
    }
  }
}

// reconstruct the four spins from the two multiplied half spinors
template <class R, int MU, bool FWD>
__device__ __forceinline__ void scatter(cx<R> o[12], const cx<R> g0[3], const cx<R> g1[3])
{
#pragma unroll
  for (int c = 0; c < 3; c++) {
    o[c] = o[c] + g0[c];
    o[3 + c] = o[3 + c] + g1[c];
    if constexpr (FWD) {
      if constexpr (MU == 0) { o[6 + c] = o[6 + c] + g0[c]; o[9 + c] = o[9 + c] + g1[c]; }
      if constexpr (MU == 1) { o[6 + c] = isub(o[6 + c], g1[c]); o[9 + c] = isub(o[9 + c], g0[c]); }
      if constexpr (MU == 2) { o[6 + c] = o[6 + c] - g1[c]; o[9 + c] = o[9 + c] + g0[c]; }
      if constexpr (MU == 3) { o[6 + c] = isub(o[6 + c], g0[c]); o[9 + c] = iadd(o[9 + c], g1[c]); }
    } else {
      if constexpr (MU == 0) { o[6 + c] = o[6 + c] - g0[c]; o[9 + c] = o[9 + c] - g1[c]; }
      if constexpr (MU == 1) { o[6 + c] = iadd(o[6 + c], g1[c]); o[9 + c] = iadd(o[9 + c], g0[c]); }
      if constexpr (MU == 2) { o[6 + c] = o[6 + c] + g1[c]; o[9 + c] = o[9 + c] - g0[c]; }
      if constexpr (MU == 3) { o[6 + c] = iadd(o[6 + c], g0[c]); o[9 + c] = isub(o[9 + c], g1[c]); }
    }
  }
}

// lds: staged records of sites [lds_e0, lds_e0 + kStageSites) of `in`, or null
template <int SP, int LK, class R, int MU, bool FWD, bool PIN>
__device__ __forceinline__ void hop_slot(const Geo& g, const Site& s, uint32_t e, const SpinorView& in, int r,
                                         const void* ub, const double* det, R anis, bool use_anis,
                                         cx<R> o[12], const char* lds = nullptr, uint32_t lds_e0 = 0)
{
  const uint32_t n = neigh_eo(g, s, MU, FWD);
  cx<R> sp[12];
  const uint32_t d = n - lds_e0;
  if (lds && d < (uint32_t)kStageSites) Stage<SP, 12>::get(lds, (int)d, sp);
  else SpAcc<SP, R, 12>::load(in, g.stride, n, sp, r);
  cx<R> h0[3], h1[3];
  project<R, MU, FWD>(sp, h0, h1);
  cx<R> u[9];
  constexpr int slot = 2 * MU + (FWD ? 0 : 1);
  LkAcc<LK, R>::load(ub, g.stride, slot, e, det + 2 * slot, u);
  cx<R> g0[3], g1[3];
  mat_vec(u, h0, g0);
  mat_vec(u, h1, g1);
  if (MU > 0 && use_anis) {
#pragma unroll
    for (int c = 0; c < 3; c++) { g0[c] = scal(g0[c], anis); g1[c] = scal(g1[c], anis); }
  }
  scatter<R, MU, FWD>(o, g0, g1);
  // Materialise the accumulators before the next group of slots issues its loads:
  // without this the compiler sinks the arithmetic of all eight slots below all
  // their loads and keeps ~400 values live (FP64 spills to scratch).  Groups of
  // four slots keep enough loads in flight per wait (measured best on gfx950).
  if constexpr (PIN) {
#pragma unroll
    for (int k = 0; k < 12; k++) asm volatile("" : "+v"(o[k].re), "+v"(o[k].im));
    asm volatile("" ::: "memory");
  }
}

// The eight slots of one output site, the accumulators pinned after every PK slots
// (PK = 4: two groups of four slots; the persistent FP64 kernel takes PK = 2).  The
// pins order the schedule only; the arithmetic, and so every bit of the result, is
// the same for every PK.
template <int SP, int LK, class R, int PK>
__device__ __forceinline__ void hop_site(const Geo& g, const Site& s, uint32_t e, const SpinorView& in, int r,
                                         const void* ub, const double* det, R anis, bool ua, cx<R> o[12],
                                         const char* lds = nullptr, uint32_t e0 = 0)
{
  hop_slot<SP, LK, R, 0, true, (1 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 0, false, (2 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 1, true, (3 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 1, false, (4 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 2, true, (5 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 2, false, (6 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 3, true, (7 % PK) == 0>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
  hop_slot<SP, LK, R, 3, false, true>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
}

template <int SP>
using RealOf = typename std::conditional<SP == SP_F64, double, float>::type;

// Twisted clover block (MODE 0: z + clover term, 1: inverse) of each chirality
// applied in place to the twelve components v of site e (the epilogue's block stage);
// the blocks are stored in the spinor precision SP
template <int BMODE, int SP, class R>
__device__ __forceinline__ void clover_site(const HopEpi& epi, const Geo& g, uint32_t e, cx<R> v[12])
{
  constexpr int EB = SP == SP_F64 ? 16 : (SP == SP_F32 ? 8 : 4);
  const void* B = epi.blk.data;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    cx<R> w[6];
    if (epi.blk_dag) clover_half<BMODE, 1, R, EB>(B, g.stride, e, h, epi.blk_kappa, epi.blk_mu, v + 6 * h, w);
    else clover_half<BMODE, 0, R, EB>(B, g.stride, e, h, epi.blk_kappa, epi.blk_mu, v + 6 * h, w);
#pragma unroll
    for (int q = 0; q < 6; q++) v[6 * h + q] = w[q];
  }
}

// Site-local epilogue of one right-hand side r at output site e (shared by the
// one-RHS and the register-blocked kernels, so both give the same bits):
//   out_h = c1_h hop + c2_h aux + c3_h aux2, then the mode's tail; returns the
// site's contribution to the reduced norm (modes 1-3).  CLV 1: the inverse clover
// blocks act on c1 hop; CLV 2: the clover blocks act on aux (the twisted clover K).
template <int SP, int MODE, int XP, class R, int CLV = 0>
__device__ __forceinline__ double hop_epilogue(const Geo& g, const HopEpi& epi, const SpinorView& out, uint32_t e,
                                               int r, CgState* st, cx<R> o[12])
{
  using RX = RealOf<XP>;
  double nrm = 0;
  double cf[12];
  if (epi.coef) {
#pragma unroll
    for (int k = 0; k < 12; k++) cf[k] = epi.coef[12 * r + k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      cf[k] = (&epi.c1[0][0])[k];
      cf[4 + k] = (&epi.c2[0][0])[k];
      cf[8 + k] = (&epi.c3[0][0])[k];
    }
  }
  const cx<R> c1u = mk<R>((R)cf[0], (R)cf[1]), c1l = mk<R>((R)cf[2], (R)cf[3]);
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = (k < 6 ? c1u : c1l) * o[k];
  if constexpr (CLV == 1) clover_site<1, SP, R>(epi, g, e, o);
  if (epi.aux.data) {
    cx<R> a[12];
    SpAcc<SP, R, 12, kNT>::load(epi.aux, g.stride, e, a, r);
    if constexpr (CLV == 2) clover_site<0, SP, R>(epi, g, e, a);
    const cx<R> cu = mk<R>((R)cf[4], (R)cf[5]), cl = mk<R>((R)cf[6], (R)cf[7]);
#pragma unroll
    for (int k = 0; k < 12; k++) fma_to(o[k], k < 6 ? cu : cl, a[k]);
  }
  if (epi.aux2.data) {
    cx<R> a[12];
    SpAcc<SP, R, 12, kNT>::load(epi.aux2, g.stride, e, a, r);
    const cx<R> cu = mk<R>((R)cf[8], (R)cf[9]), cl = mk<R>((R)cf[10], (R)cf[11]);
#pragma unroll
    for (int k = 0; k < 12; k++) fma_to(o[k], k < 6 ? cu : cl, a[k]);
  }
  if constexpr (MODE == 0) {
    SpAcc<SP, R, 12, kNT>::store(out, g.stride, e, o, r);
  } else if constexpr (MODE == 1) {
    if (epi.write) SpAcc<SP, R, 12, kNT>::store(out, g.stride, e, o, r);
#pragma unroll
    for (int k = 0; k < 12; k++) nrm += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
  } else if constexpr (MODE == 2) {
    // CG tail: s = o; x += alpha p (when x is given) ; r -= alpha s ; |r|^2
    const RX alpha = (RX)st->alpha;
    if (epi.x.data) {
      cx<RX> x[12];
      cx<R> p[12];
      SpAcc<XP, RX, 12, kNT>::load(epi.x, g.stride, e, x, r);
      SpAcc<SP, R, 12, kNT>::load(epi.p, g.stride, e, p, r);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        x[k].re = fma(alpha, (RX)p[k].re, x[k].re);
        x[k].im = fma(alpha, (RX)p[k].im, x[k].im);
      }
      SpAcc<XP, RX, 12, kNT>::store(epi.x, g.stride, e, x, r);
    }
    cx<RX> rv[12];
    SpAcc<XP, RX, 12, kNT>::load(epi.r, g.stride, e, rv, r);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      rv[k].re = fma(-alpha, (RX)o[k].re, rv[k].re);
      rv[k].im = fma(-alpha, (RX)o[k].im, rv[k].im);
      nrm += (double)rv[k].re * (double)rv[k].re + (double)rv[k].im * (double)rv[k].im;
    }
    SpAcc<XP, RX, 12, kNT>::store(epi.r, g.stride, e, rv, r);
  } else {
    // residual tail: r64 = o ; rs = sloppy(o) ; |o|^2
    SpAcc<SP, R, 12, kNT>::store(epi.r, g.stride, e, o, r);
    if (epi.rs.data) {
      cx<RX> q[12];
#pragma unroll
      for (int k = 0; k < 12; k++) q[k] = mk<RX>((RX)o[k].re, (RX)o[k].im);
      SpAcc<XP, RX, 12, kNT>::store(epi.rs, g.stride, e, q, r);
    }
#pragma unroll
    for (int k = 0; k < 12; k++) nrm += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
  }
  return nrm;
}

__device__ __forceinline__ bool rhs_idle(const CgState* st, int pred)
{
  return st->done || (pred == 1 && !st->reliable) || (pred == 2 && st->reliable);
}

// workgroups of at most 256 threads; FP64 builds get the full 256-VGPR budget
// (one wavefront per SIMD per workgroup), the sloppy builds two workgroups per CU
template <int SP, int LK, int MODE, int XP, int CLV = 0>
__global__ void __launch_bounds__(256, (SP != SP_F64 ? 2 : 1)) k_tm_hop(Geo g, int op, SpinorView out, SpinorView in, LinkView U, double anis_d,
                                               int use_anis, HopEpi epi, Shape S, int64_t nchunks,
                                               double* partial)
{
  using R = RealOf<SP>;
  if constexpr (MODE == 0) {
    // the workgroup past the last block finalises the previous launch's reduction
    // (one right-hand side; the hop itself never reads what it writes)
    if (epi.fold_fin != FIN_NONE && xcd_block() == (int64_t)gridDim.x - 1) {
      __shared__ double vl[1024];
      CgState* fs = epi.st;
      if (!(fs->done || (epi.fold_pred == 1 && !fs->reliable) || (epi.fold_pred == 2 && fs->reliable))) {
        const double tot = block_sum_partials_256(partial, nchunks, vl);
        if (threadIdx.x == 0) finalize(fs, epi.fold_fin, tot);
      }
      return;
    }
  }
  const Map m = map_thread(S, g.Vh);
  CgState* st = epi.st ? epi.st + m.r : nullptr;
  const bool skip = st && rhs_idle(st, epi.pred_reliable);
  // a launch whose every right-hand side is idle does nothing (uniform over the grid)
  if (epi.st && S.NR == 1 && skip) return;
  double nrm = 0;
  if (m.active && !skip) {
    const uint32_t e = m.e;
    const int r = m.r;
    const Site s = site_of(g, op, e);
    const void* ub = U.data[op];
    const double* det = &U.det[op][0][0];
    const R anis = (R)anis_d;
    const bool ua = use_anis != 0;
    cx<R> o[12];
#pragma unroll
    for (int k = 0; k < 12; k++) o[k] = mk<R>(0, 0);
    hop_site<SP, LK, R, 4>(g, s, e, in, r, ub, det, anis, ua, o);
    nrm = hop_epilogue<SP, MODE, XP, R, CLV>(g, epi, out, e, r, st, o);
  }
  if constexpr (MODE != 0) store_partial(nrm, partial, nchunks, m.r, m.chunk);
}

// One right-hand side, neighbours staged in LDS (Stage): the workgroup's four
// wavefronts cover 256 consecutive output sites (Shape W = 4, NRB = 1).
template <int SP, int LK, int MODE, int XP, int CLV = 0>
__global__ void __launch_bounds__(256, (SP != SP_F64 ? 2 : 1)) k_tm_hop_lds(Geo g, int op, SpinorView out,
                                                                        SpinorView in, LinkView U, double anis_d,
                                                                        int use_anis, HopEpi epi, Shape S,
                                                                        int64_t nchunks, double* partial)
{
  using R = RealOf<SP>;
  __shared__ __attribute__((aligned(16))) char lds[Stage<SP, 12>::BYTES];
  const Map m = map_thread(S, g.Vh);
  CgState* st = epi.st ? epi.st + m.r : nullptr;
  const bool skip = st && rhs_idle(st, epi.pred_reliable);
  if (epi.st && skip) return;  // uniform: one right-hand side per launch
  const uint32_t e0 = (uint32_t)((m.chunk - (threadIdx.x >> 6)) * 64);  // first site of the workgroup
  const int t = (int)threadIdx.x;
  Stage<SP, 12>::put(lds, in, g.stride, e0 + t, e0 + t < (uint64_t)g.Vh, t, m.r);
  __syncthreads();
  double nrm = 0;
  if (m.active) {
    const uint32_t e = m.e;
    const int r = m.r;
    const Site s = site_of(g, op, e);
    const void* ub = U.data[op];
    const double* det = &U.det[op][0][0];
    const R anis = (R)anis_d;
    const bool ua = use_anis != 0;
    cx<R> o[12];
#pragma unroll
    for (int k = 0; k < 12; k++) o[k] = mk<R>(0, 0);
    hop_site<SP, LK, R, 4>(g, s, e, in, r, ub, det, anis, ua, o, lds, e0);
    nrm = hop_epilogue<SP, MODE, XP, R, CLV>(g, epi, out, e, r, st, o);
  }
  if constexpr (MODE != 0) store_partial(nrm, partial, nchunks, m.r, m.chunk);
}

// Persistent form of k_tm_hop_lds for the FP64 hop: one resident workgroup per slot
// loops over 256-site blocks, so no SIMD idles between a finished workgroup and the
// dispatch of the next.  Its accumulators are pinned after every two slots (PK = 2:
// 176-195 VGPRs, two wavefronts per SIMD; with PK = 4 the loads of four slots in
// flight take 256 VGPRs + 32 AGPRs and one wavefront per SIMD).  Measured, hop
// alone: 24^3x48 102.2 us (PK = 2) against 106.1 (PK = 4), 32^3x64 276.0 against
// 276.5; PK = 1 (136-162 VGPRs, three wavefronts) 106 and 296 us.  The site
// body is the one of k_tm_hop_lds, written out in the same form (the compiler's
// FMA contraction follows the code's shape, and the kernels must agree bit for
// bit with each other and with k_tm_hop).  The
// workgroups of XCD x (b % 8 == x, the hardware's round-robin) walk one
// contiguous eighth of the blocks together, block bi, bi + nbx, ... of it, as the
// one-shot kernel's dispatch order does.  Partials stay per 64-site chunk.
template <int SP, int LK, int MODE, int XP, int CLV = 0>
__global__ void __launch_bounds__(256, 1) k_tm_hop_lds_p(Geo g, int op, SpinorView out, SpinorView in, LinkView U,
                                                         double anis_d, int use_anis, HopEpi epi, int64_t nblk,
                                                         int64_t nchunks, double* partial)
{
  using R = RealOf<SP>;
  __shared__ __attribute__((aligned(16))) char lds[Stage<SP, 12>::BYTES];
  CgState* st = epi.st;
  if (st && rhs_idle(st, epi.pred_reliable)) return;  // uniform: one right-hand side per launch
  const uint32_t b = blockIdx.x, nb = gridDim.x;
  const uint32_t xcd = b & 7u, bi = b >> 3;
  const int64_t nbx = (int64_t)(nb >> 3) + (xcd < (nb & 7u) ? 1 : 0);  // workgroups on this XCD
  const int64_t q8 = nblk >> 3, r8 = nblk & 7;
  const int64_t lo = (int64_t)xcd * q8 + ((int64_t)xcd < r8 ? (int64_t)xcd : r8);
  const int64_t hi = lo + q8 + ((int64_t)xcd < r8 ? 1 : 0);
  const int t = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const void* ub = U.data[op];
  const double* det = &U.det[op][0][0];
  const R anis = (R)anis_d;
  const bool ua = use_anis != 0;
  if constexpr (MODE == 0) {
    // No reduction, so no 64-site chunk to respect: the XCD's sites are walked in
    // rounds of 256 per workgroup, and the sites left after the last whole round
    // are split evenly over all of the XCD's workgroups and, inside each, over its
    // four wavefronts.  With whole blocks only, 24^3x48 (162 blocks per XCD on 64
    // workgroups) ends with a third round in which 30 of 64 workgroups idle.
    const int64_t s_lo = (int64_t)g.Vh * xcd / 8, n = (int64_t)g.Vh * (xcd + 1) / 8 - s_lo;
    const int64_t per_round = nbx * kStageSites, full = n / per_round;
    const int lane = t & 63;
    for (int64_t k = 0; k <= full; k++) {  // uniform over the workgroup
      int64_t e0, len;
      if (k < full) {
        e0 = s_lo + (k * nbx + bi) * kStageSites;
        len = kStageSites;
      } else {
        const int64_t rem = n - full * per_round, q = (rem + nbx - 1) / nbx;
        e0 = s_lo + full * per_round + (int64_t)bi * q;
        len = s_lo + n - e0 < q ? s_lo + n - e0 : q;
        if (len <= 0) break;
      }
      __syncthreads();  // the previous window's LDS reads are done
      Stage<SP, 12>::put(lds, in, g.stride, (uint32_t)e0 + t, e0 + t < g.Vh, t, 0);
      __syncthreads();
      const int wl = (int)((len + 3) / 4);  // sites per wavefront (64 in a whole round)
      const int64_t eo = (int64_t)wave * wl + lane;
      if (lane < wl && eo < len) {
        const uint32_t e = (uint32_t)(e0 + eo);
        const Site s = site_of(g, op, e);
        cx<R> o[12];
#pragma unroll
        for (int k2 = 0; k2 < 12; k2++) o[k2] = mk<R>(0, 0);
        hop_site<SP, LK, R, 2>(g, s, e, in, 0, ub, det, anis, ua, o, lds, (uint32_t)e0);
        hop_epilogue<SP, MODE, XP, R, CLV>(g, epi, out, e, 0, st, o);
      }
    }
    return;
  }
  for (int64_t cb = lo + bi; cb < hi; cb += nbx) {  // uniform over the workgroup
    const uint32_t e0 = (uint32_t)(cb * kStageSites);
    __syncthreads();  // the previous block's LDS reads are done
    Stage<SP, 12>::put(lds, in, g.stride, e0 + t, e0 + t < (uint64_t)g.Vh, t, 0);
    __syncthreads();
    const uint32_t e = e0 + (uint32_t)t;
    double nrm = 0;
    if (e < (uint64_t)g.Vh) {
      const Site s = site_of(g, op, e);
      cx<R> o[12];
#pragma unroll
      for (int k = 0; k < 12; k++) o[k] = mk<R>(0, 0);
      hop_site<SP, LK, R, 2>(g, s, e, in, 0, ub, det, anis, ua, o, lds, e0);
      nrm = hop_epilogue<SP, MODE, XP, R, CLV>(g, epi, out, e, 0, st, o);
    }
    if constexpr (MODE != 0) store_partial(nrm, partial, nchunks, 0, cb * 4 + wave);
  }
}

// resident workgroups per CU of the persistent kernel (queried once per instance)
template <int SP, int LK, int MODE, int XP, int CLV>
static int persist_per_cu()
{
  static const int occ = [] {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, (const void*)k_tm_hop_lds_p<SP, LK, MODE, XP, CLV>, 256, 0) !=
            hipSuccess || o < 1)
      o = 0;
    return o;
  }();
  return occ;
}

// ---------------------------------------------------------------- register-blocked RHS
// One slot of the hop for RB right-hand sides r0..r0+nrb-1 of the same site: the
// link is loaded once and applied to each right-hand side's projected neighbour.
template <int SP, int LK, class R, int MU, bool FWD, int RB>
__device__ __forceinline__ void hop_slot_rb(const Geo& g, const Site& s, uint32_t e, const SpinorView& in, int r0,
                                            int nrb, unsigned busy, const void* ub, const double* det, R anis,
                                            bool use_anis, cx<R> o[RB][12], const char* lds = nullptr,
                                            uint32_t lds_e0 = 0)
{
  const uint32_t n = neigh_eo(g, s, MU, FWD);
  cx<R> u[9];
  constexpr int slot = 2 * MU + (FWD ? 0 : 1);
  LkAcc<LK, R, kRbLinkCP>::load(ub, g.stride, slot, e, det + 2 * slot, u);
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (b >= nrb || !((busy >> b) & 1u)) continue;  // uniform over the wavefront
    cx<R> sp[12];
    const uint32_t d = n - lds_e0;
    if (lds && d < (uint32_t)kStageSites) Stage<SP, 12>::get(lds + b * Stage<SP, 12>::BYTES, (int)d, sp);
    else SpAcc<SP, R, 12>::load(in, g.stride, n, sp, r0 + b);
    cx<R> h0[3], h1[3];
    project<R, MU, FWD>(sp, h0, h1);
    cx<R> g0[3], g1[3];
    mat_vec(u, h0, g0);
    mat_vec(u, h1, g1);
    if (MU > 0 && use_anis) {
#pragma unroll
      for (int c = 0; c < 3; c++) { g0[c] = scal(g0[c], anis); g1[c] = scal(g1[c], anis); }
    }
    scatter<R, MU, FWD>(o[b], g0, g1);
#pragma unroll
    for (int k = 0; k < 12; k++) asm volatile("" : "+v"(o[b][k].re), "+v"(o[b][k].im));
  }
  asm volatile("" ::: "memory");
}

// Multi-RHS hop for the sloppy precisions: wavefront w of a workgroup takes chunk
// 4*cb + w and the RB right-hand sides of group rg (Shape: W = 4, NRB = 1, NR = NRG
// = number of RB-groups).  Link traffic and link loads per right-hand side drop by RB.
template <int SP, int LK, int MODE, int XP, int RB>
__global__ void __launch_bounds__(256) k_tm_hop_rb(Geo g, int op, SpinorView out, SpinorView in, LinkView U,
                                                   double anis_d, int use_anis, HopEpi epi, Shape S, int nr,
                                                   int64_t nchunks, double* partial)
{
  using R = RealOf<SP>;
  const Map m = map_thread_rb(S, g, nr);
  const int r0 = m.r * RB, nrb = min(RB, nr - r0);
  unsigned busy = 0;
#pragma unroll
  for (int b = 0; b < RB; b++)
    if (b < nrb && !(epi.st && rhs_idle(epi.st + r0 + b, epi.pred_reliable))) busy |= 1u << b;
  // right-hand sides that idle here are skipped by the finalisation as well.
  // FP16: the busy right-hand sides' records of the workgroup's 256 sites are
  // staged in LDS and the neighbours inside that range read from there (the
  // batched hop is bound by the L2->L1 path: 12 sources, 684 -> 667 us for the
  // plain hop and 987 -> 937 us for the CG tail at 32^3x64)
  constexpr bool STG = SP == SP_F16;
  __shared__ __attribute__((aligned(16))) char lds[STG ? RB * Stage<SP, 12>::BYTES : 16];
  const uint32_t e0 = (uint32_t)((m.chunk - (threadIdx.x >> 6)) * 64);
  if (!busy) {  // uniform over the workgroup (one right-hand-side group)
    if constexpr (MODE != 0) {
#pragma unroll
      for (int b = 0; b < RB; b++)
        if (b < nrb) store_partial(0.0, partial, nchunks, r0 + b, m.chunk);
    }
    return;
  }
  if constexpr (STG) {
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (b < nrb && ((busy >> b) & 1u))
        Stage<SP, 12>::put(lds + b * Stage<SP, 12>::BYTES, in, g.stride, e0 + threadIdx.x,
                           e0 + threadIdx.x < (uint64_t)g.Vh, (int)threadIdx.x, r0 + b);
    __syncthreads();
  }
  const char* lp = STG ? lds : nullptr;
  const uint32_t e = m.e;
  cx<R> o[RB][12];
#pragma unroll
  for (int b = 0; b < RB; b++)
#pragma unroll
    for (int k = 0; k < 12; k++) o[b][k] = mk<R>(0, 0);
  // sites past the end (a partial last chunk) compute nothing but take part in the
  // wavefront reductions with zero
  if (m.active) {
    const Site s = site_of(g, op, e);
    const void* ub = U.data[op];
    const double* det = &U.det[op][0][0];
    const R anis = (R)anis_d;
    const bool ua = use_anis != 0;
    hop_slot_rb<SP, LK, R, 0, true, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 0, false, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 1, true, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 1, false, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 2, true, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 2, false, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 3, true, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
    hop_slot_rb<SP, LK, R, 3, false, RB>(g, s, e, in, r0, nrb, busy, ub, det, anis, ua, o, lp, e0);
  }
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (b >= nrb) continue;
    double nrm = 0;
    if (m.active && ((busy >> b) & 1u)) nrm = hop_epilogue<SP, MODE, XP, R>(g, epi, out, e, r0 + b, epi.st ? epi.st + r0 + b : nullptr, o[b]);
    if constexpr (MODE != 0) store_partial(nrm, partial, nchunks, r0 + b, m.chunk);
  }
}

static double sp_site_bytes(int sp, int ncomp)
{
  switch (sp) {
  case SP_F64: return 16.0 * ncomp;
  case SP_F32: return 8.0 * ncomp;
  default: return 4.0 * ncomp + 4.0;
  }
}
static double lk_bytes(int lk)
{
  switch (lk) {
  case LK_F64_18: return 144;
  case LK_F32_18: return 72;
  case LK_F32_12: return 48;
  case LK_F64_12: return 96;
  default: return 24;
  }
}

static int xp_of(const HopEpi& epi, int sp)
{
  if (epi.mode == 2) return epi.r.prec;
  if (epi.mode == 3) return epi.rs.data ? epi.rs.prec : sp;
  return sp;
}

double tm_hop_bytes(int sp, int lk, const HopEpi& epi, int ncomp)
{
  // per output site: the 8 links once for all right-hand sides, everything else per RHS
  const double s = sp_site_bytes(sp, ncomp);
  double b = s;  // input spinor (each site read once)
  if (epi.aux.data) b += s;
  if (epi.aux2.data) b += s;
  if (epi.mode == 0 || (epi.mode == 1 && epi.write)) b += s;
  if (epi.mode == 2) {
    const double sx = sp_site_bytes(xp_of(epi, sp), ncomp);
    b += epi.x.data ? 4 * sx + s : 2 * sx;  // r (and x) read+write, p read
  }
  if (epi.mode == 3) {
    b += s;
    if (epi.rs.data) b += sp_site_bytes(epi.rs.prec, ncomp);
  }
  // clover blocks of the epilogue: 72 (inverse) or 36 complex per site in the spinor
  // precision, read once for all right-hand sides
  const double blk = epi.blk_mode ? (epi.blk_mode == 1 ? 72.0 : 36.0) * (sp == SP_F64 ? 16 : (sp == SP_F32 ? 8 : 4)) : 0.0;
  return 8 * lk_bytes(lk) + blk + b * epi.nrhs;
}

int64_t blocks_for(int64_t n) { return (n + kBlock - 1) / kBlock; }

// launch of the register-blocked kernel with RB right-hand sides per wavefront
// (RB = 0: the one-RHS kernel with the field-kernel shape)
template <int SP, int LK, int MODE, int XP>
static int launch_rb(nq_ctx* ctx, int RB, int op, const SpinorView& out, const SpinorView& in, const LinkView& U,
                     double anis, const HopEpi& epi, int nr)
{
  const Geo& g = ctx->geo;
  const int ua = anis != 1.0 ? 1 : 0;
  if (RB <= 0) {
    // one right-hand side per wavefront (the kernel of the unbatched operators)
    const Launch L = launch_shape(g, nr);
    NQ_HOP_LAUNCH((k_tm_hop<SP, LK, MODE, XP>), dim3((unsigned)L.blocks), dim3(L.threads), g, op,
                       out, in, U, anis, ua, epi, L.s, L.nchunks, ctx->partial);
    NQ_HIP(hipGetLastError());
    return NQ_OK;
  }
  Shape sh;
  sh.W = 4;
  sh.NRB = 1;
  sh.NRG = sh.NR = (nr + RB - 1) / RB;
  const int64_t nchunks = (g.Vh + 63) / 64;
  const dim3 grid((unsigned)((nchunks + 3) / 4 * sh.NRG)), blk(256);
  switch (RB) {
  case 1: NQ_HOP_LAUNCH((k_tm_hop_rb<SP, LK, MODE, XP, 1>), grid, blk, g, op, out, in, U, anis, ua, epi, sh, nr, nchunks, ctx->partial); break;
  case 2: NQ_HOP_LAUNCH((k_tm_hop_rb<SP, LK, MODE, XP, 2>), grid, blk, g, op, out, in, U, anis, ua, epi, sh, nr, nchunks, ctx->partial); break;
  default: NQ_HOP_LAUNCH((k_tm_hop_rb<SP, LK, MODE, XP, 4>), grid, blk, g, op, out, in, U, anis, ua, epi, sh, nr, nchunks, ctx->partial); break;
  }
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

// Autotuner of the right-hand sides per wavefront of the batched sloppy hop: on the
// first batched launch of a (precision, batch size) it times the write-only kernel
// (no solver state, output into a scratch buffer) with RB = 1, 2 and 4 right-hand
// sides per wavefront on the caller's input and links, and keeps the fastest for
// every mode of that precision and batch size.  All RB build the same per-RHS code
// (one unrolled copy per right-hand side), so the choice never changes a result; the
// one-RHS kernel (RB = 0 in launch_rb) is not a candidate: its FP16 code schedules
// the same arithmetic differently and does not give the same bits.  NQ_TUNE=0 or an
// open graph capture use RB = 2.
template <int SP, int LK, int XP>
static int tuned_rb(nq_ctx* ctx, int op, const SpinorView& in, const LinkView& U, double anis, const HopEpi& epi,
                    int nr, int* rb)
{
  const std::string key = "tm_hop_rb/" + std::to_string(SP) + "/" + std::to_string(LK) + "/" + std::to_string(nr);
  auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) {
    *rb = it->second;
    return NQ_OK;
  }
  const char* env = getenv("NQ_TUNE");
  if (ctx->capturing || (env && env[0] == '0')) {
    *rb = 2;
    return NQ_OK;
  }
  SpinorView scratch;
  NQ_TRY(alloc_spinor_nr(ctx, "tune.out" + std::to_string(SP), SP, 12, nr, &scratch));
  HopEpi e0;
  e0.nrhs = nr;
  e0.batched = 1;
  e0.coef = epi.coef;
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  hipEvent_t a, b;
  NQ_HIP(hipEventCreate(&a));
  NQ_HIP(hipEventCreate(&b));
  float best = 1e30f;
  int best_rb = 2;
  for (int cand : {1, 2, 4}) {
    if (cand > 1 && cand > nr) continue;
    NQ_TRY((launch_rb<SP, LK, 0, XP>(ctx, cand, op, scratch, in, U, anis, e0, nr)));  // warm-up
    NQ_HIP(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < 3; i++) NQ_TRY((launch_rb<SP, LK, 0, XP>(ctx, cand, op, scratch, in, U, anis, e0, nr)));
    NQ_HIP(hipEventRecord(b, ctx->stream));
    NQ_HIP(hipEventSynchronize(b));
    float ms = 0;
    NQ_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best * 0.98f) {  // prefer the smaller block unless clearly slower
      best = ms;
      best_rb = cand;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->profiling = prof;
  ctx->tuned[key] = best_rb;
  *rb = best_rb;
  return NQ_OK;
}

// Launches the persistent staged hop and returns true when it applies: FP64
// with 18-real links and a lattice that fills the resident slots at least twice
// (measured 2.8 % faster at 32^3x64 than the one-shot kernel).  The 12-real FP64 hop runs two wavefronts
// per SIMD in the one-shot kernel and measured slower in the persistent one.
template <int SP, int LK, int MODE, int XP, int CLV>
static bool launch_persistent(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U,
                              double anis, const HopEpi& epi, const Launch& L)
{
  if constexpr (SP == SP_F64 && LK == LK_F64_18) {
    const char* penv = getenv("NQ_PERSIST");
    if (penv && penv[0] == '0') return false;
    const int64_t nblk = (L.nchunks + 3) / 4, slots = (int64_t)ctx->ncu * persist_per_cu<SP, LK, MODE, XP, CLV>();
    if (slots < 1 || nblk < 2 * slots) return false;
    NQ_HOP_LAUNCH((k_tm_hop_lds_p<SP, LK, MODE, XP, CLV>), dim3((unsigned)slots), dim3(256),
                       ctx->geo, op, out, in, U, anis, anis != 1.0 ? 1 : 0, epi, nblk, L.nchunks, ctx->partial);
    return true;
  } else {
    (void)ctx; (void)op; (void)out; (void)in; (void)U; (void)anis; (void)epi; (void)L;
    return false;
  }
}

template <int SP, int LK, int MODE, int XP, int CLV = 0>
static int go(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U, double anis,
              const HopEpi& epi, const char* name)
{
  const Geo& g = ctx->geo;
  const int nr = epi.nrhs;
  if (nr < 1 || nr > 16) return fail(NQ_EINVAL, "tm hop: right-hand sides per launch must be 1..16");
  const Launch L = launch_shape(g, nr);
  if (L.nchunks * nr > ctx->max_partials) return fail(NQ_EINVAL, "tm hop: reduction scratch too small");
  if constexpr (MODE != 3 && CLV == 0) {
    if (epi.batched || nr > 1) {
      // register-blocked right-hand sides (always for batched solves, so that a
      // right-hand side runs the same code alone and in a batch); RB right-hand
      // sides per wavefront, chosen by the autotuner for this precision and batch.
      // FP64 too: the reliable updates of a batch share each link load (measured
      // 953 -> 879 us for 12 sources, the batched FP64 solve 95 -> 85 ms per source)
      int RB = 0;
      NQ_TRY((tuned_rb<SP, LK, XP>(ctx, op, in, U, anis, epi, nr, &RB)));
      prof_begin_ext(ctx, name, tm_hop_bytes(SP, LK, epi) * (double)g.Vh, &ctx->hop_ev[0], &ctx->hop_ev[1]);
      const int rc = launch_rb<SP, LK, MODE, XP>(ctx, RB, op, out, in, U, anis, epi, nr);
      prof_end_ext(ctx);
      NQ_TRY(rc);
      if (epi.fold_fin != FIN_NONE) NQ_TRY(launch_finalize(ctx, nr, epi.st, epi.fold_fin, epi.fold_pred));
      if (MODE != 0 && epi.fin != FIN_PARTIAL) NQ_TRY(launch_finalize(ctx, nr, epi.st, epi.fin, epi.pred_reliable));
      return NQ_OK;
    }
  }
  prof_begin_ext(ctx, name, tm_hop_bytes(SP, LK, epi) * (double)g.Vh, &ctx->hop_ev[0], &ctx->hop_ev[1]);
  bool fold_here = false;
  // FP64: neighbours staged in LDS (measured 4 % faster at 32^3x64; the sloppy
  // kernels measured slower with staging and read their neighbours from memory)
  const char* lenv = getenv("NQ_LDS");
  if (SP == SP_F64 && nr == 1 && L.s.W == 4 && !(lenv && lenv[0] == '0')) {
    if (!launch_persistent<SP, LK, MODE, XP, CLV>(ctx, op, out, in, U, anis, epi, L))
      NQ_HOP_LAUNCH((k_tm_hop_lds<SP, LK, MODE, XP, CLV>), dim3((unsigned)L.blocks), dim3(L.threads), g,
                         op, out, in, U, anis, anis != 1.0 ? 1 : 0, epi, L.s, L.nchunks, ctx->partial);
  }
  else {
    // one extra workgroup when it carries the finalisation of the previous reduction
    // (256 threads: the 1024-lane summation order); the kernel sees fold_fin only then
    fold_here = MODE == 0 && epi.fold_fin != FIN_NONE && nr == 1 && epi.st && L.threads == 256;
    HopEpi ke = epi;
    if (!fold_here) ke.fold_fin = FIN_NONE;
    NQ_HOP_LAUNCH((k_tm_hop<SP, LK, MODE, XP, CLV>), dim3((unsigned)(L.blocks + (fold_here ? 1 : 0))),
                  dim3(L.threads), g, op, out, in, U, anis, anis != 1.0 ? 1 : 0, ke, L.s, L.nchunks, ctx->partial);
  }
  prof_end_ext(ctx);
  NQ_HIP(hipGetLastError());
  if (epi.fold_fin != FIN_NONE && !fold_here) NQ_TRY(launch_finalize(ctx, 1, epi.st, epi.fold_fin, epi.fold_pred));
  if (MODE != 0 && epi.fin != FIN_PARTIAL) NQ_TRY(launch_finalize(ctx, nr, epi.st, epi.fin, epi.pred_reliable));
  return NQ_OK;
}

#define NQ_MODES(SP, LK, XP, TAG)                                                            \
  switch (epi.mode) {                                                                       \
  case 0: return go<SP, LK, 0, SP>(ctx, op, out, in, U, anis, epi, "tm_hop_" TAG);            \
  case 1: return go<SP, LK, 1, SP>(ctx, op, out, in, U, anis, epi, "tm_hop_norm_" TAG);       \
  case 2: return go<SP, LK, 2, XP>(ctx, op, out, in, U, anis, epi, "tm_hop_cgtail_" TAG);     \
  default: break;                                                                           \
  }

// Twisted clover K: the first hop with the inverse blocks (blk_mode 1, mode 0 only)
// and the second with the clover blocks on aux (blk_mode 2, modes 0-2), for the
// operator precisions of the clover solves: FP64 spinors with FP64 18-real links and
// FP64 blocks; FP32 spinors with FP32 18-real links and FP32 blocks; FP16 spinors
// with FP32 or FP16 12-real links and FP16 blocks.
template <int SP, int LK, int XP>
static int clover_modes(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U,
                        double anis, const HopEpi& epi, const char* tag)
{
  const std::string n1 = std::string("tm_hop_clinv_") + tag, n0 = std::string("tm_hop_cldee_") + tag,
                    nn = std::string("tm_hop_norm_cldee_") + tag, nt = std::string("tm_hop_cgtail_cldee_") + tag;
  if (epi.blk_mode == 1 && epi.mode == 0) return go<SP, LK, 0, SP, 1>(ctx, op, out, in, U, anis, epi, n1.c_str());
  if (epi.blk_mode == 2) {
    if (!epi.aux.data) return fail(NQ_EINVAL, "tm hop: clover blocks on aux without aux");
    switch (epi.mode) {
    case 0: return go<SP, LK, 0, SP, 2>(ctx, op, out, in, U, anis, epi, n0.c_str());
    case 1: return go<SP, LK, 1, SP, 2>(ctx, op, out, in, U, anis, epi, nn.c_str());
    case 2: return go<SP, LK, 2, XP, 2>(ctx, op, out, in, U, anis, epi, nt.c_str());
    default: break;
    }
  }
  return fail(NQ_EINVAL, "tm hop: clover epilogue %d with mode %d", epi.blk_mode, epi.mode);
}

static int launch_tm_hop_clover(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U,
                                double anis, const HopEpi& epi)
{
  const int sp = in.prec, lk = U.prec;
  if (!epi.blk.data || epi.blk.prec != sp || epi.blk.ncomp != (epi.blk_mode == 1 ? 72 : 36))
    return fail(NQ_EINVAL, "tm hop: clover blocks of the wrong kind or precision");
  if (sp == SP_F64 && lk == LK_F64_18) return clover_modes<SP_F64, LK_F64_18, SP_F64>(ctx, op, out, in, U, anis, epi, "f64");
  if (sp == SP_F32 && lk == LK_F32_18) return clover_modes<SP_F32, LK_F32_18, SP_F32>(ctx, op, out, in, U, anis, epi, "f32");
  if (sp == SP_F16 && lk == LK_F32_12)
    return clover_modes<SP_F16, LK_F32_12, SP_F32>(ctx, op, out, in, U, anis, epi, "f16_l32c");
  if (sp == SP_F16 && lk == LK_F16_12)
    return clover_modes<SP_F16, LK_F16_12, SP_F32>(ctx, op, out, in, U, anis, epi, "f16_l16c");
  return fail(NQ_EINVAL, "tm hop: no clover epilogue for spinor precision %d with links %d", sp, lk);
}

int launch_tm_hop(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U, double anis,
                  const HopEpi& epi)
{
  if (in.ncomp != 12) return fail(NQ_EINVAL, "tm hop needs spincolor fields");
  if (epi.mode != 0 && !epi.st) return fail(NQ_EINVAL, "reducing epilogue without state");
  const int sp = in.prec, lk = U.prec;
  if ((epi.mode == 0 || (epi.mode == 1 && epi.write)) && out.prec != sp)
    return fail(NQ_EINVAL, "tm hop: output precision differs from input");
  if (epi.aux.data && epi.aux.prec != sp) return fail(NQ_EINVAL, "tm hop: aux precision mismatch");
  if (epi.aux2.data && epi.aux2.prec != sp) return fail(NQ_EINVAL, "tm hop: aux2 precision mismatch");
  if (epi.mode == 2 && (epi.p.prec != sp || (epi.x.data && epi.x.prec != epi.r.prec)))
    return fail(NQ_EINVAL, "tm hop: CG tail precision mismatch");
  if (epi.blk_mode) return launch_tm_hop_clover(ctx, op, out, in, U, anis, epi);
  if (sp == SP_F64 && lk == LK_F64_18) {
    if (epi.mode == 2 && epi.r.prec != SP_F64) return fail(NQ_EINVAL, "FP64 CG tail needs FP64 x,r");
    NQ_MODES(SP_F64, LK_F64_18, SP_F64, "f64")
    if (epi.mode == 3) {
      if (epi.r.prec != SP_F64) return fail(NQ_EINVAL, "residual tail needs FP64 r");
      if (!epi.rs.data) return go<SP_F64, LK_F64_18, 3, SP_F64>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64");
      if (epi.rs.prec == SP_F32) return go<SP_F64, LK_F64_18, 3, SP_F32>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64_s32");
      if (epi.rs.prec == SP_F16) return go<SP_F64, LK_F64_18, 3, SP_F16>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64_s16");
    }
  } else if (sp == SP_F64 && lk == LK_F64_12) {
    if (epi.mode == 2 && epi.r.prec != SP_F64) return fail(NQ_EINVAL, "FP64 CG tail needs FP64 x,r");
    NQ_MODES(SP_F64, LK_F64_12, SP_F64, "f64_l64c")
    if (epi.mode == 3) {
      if (epi.r.prec != SP_F64) return fail(NQ_EINVAL, "residual tail needs FP64 r");
      if (!epi.rs.data) return go<SP_F64, LK_F64_12, 3, SP_F64>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64_l64c");
      if (epi.rs.prec == SP_F32) return go<SP_F64, LK_F64_12, 3, SP_F32>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64_l64c_s32");
      if (epi.rs.prec == SP_F16) return go<SP_F64, LK_F64_12, 3, SP_F16>(ctx, op, out, in, U, anis, epi, "tm_hop_res_f64_l64c_s16");
    }
  } else if (sp == SP_F32 && lk == LK_F32_18) {
    if (epi.mode == 2 && epi.r.prec != SP_F32) return fail(NQ_EINVAL, "FP32 CG tail needs FP32 x,r");
    NQ_MODES(SP_F32, LK_F32_18, SP_F32, "f32")
  } else if (sp == SP_F16 && lk == LK_F32_12) {
    if (epi.mode == 2 && epi.r.prec != SP_F32) return fail(NQ_EINVAL, "FP16 CG tail needs FP32 x,r");
    NQ_MODES(SP_F16, LK_F32_12, SP_F32, "f16_l32c")
  } else if (sp == SP_F16 && lk == LK_F16_12) {
    if (epi.mode == 2 && epi.r.prec != SP_F32) return fail(NQ_EINVAL, "FP16 CG tail needs FP32 x,r");
    NQ_MODES(SP_F16, LK_F16_12, SP_F32, "f16_l16c")
  }
  return fail(NQ_EINVAL, "tm hop: unsupported precision combination (spinor %d, links %d, mode %d)", sp, lk, epi.mode);
}

}  // namespace nq
