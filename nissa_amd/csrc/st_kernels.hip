// Staggered operator and multi-shift CG kernels for gfx950.
//
// Reference: apply_st2Doe / apply_stDeo_half / apply_stD2ee_m2
// (src/dirac_operators/stD/dirac_operator_stD.hpp:47-225) and cgm_invert
// (src/inverters/templates/cgm_invert_template_threaded.hpp:13-295).
//   hop(x) = sum_mu U_mu(x) in(x+mu) - U_mu(x-mu)^+ in(x-mu)   (staggered phases in U)
//   A      = m2 - 1/4 Deo Doe  on even sites
// Links come from the same per-output-parity slot arrays as the twisted-mass
// hop (slot 2mu+1 already holds U_mu(x-mu)^+).  Colour fields are SoA with three
// complex components; all kernels take the (64-site chunk x right-hand side)
// mapping so one launch serves every source of a batch.
//
// CGM iteration (per source, shifts k in the running set):
//   C1: t = Doe p,                |t|^2 -> pap = |t|^2/4   (= <p, A(0) p>, Doe^+ = -Deo)
//   C2: s = -1/4 Deo t; sol_k -= beta_k ps_k; r += betaa s;  |r|^2
//   C3: p = r + alpha p; ps_k = zeta_k r + alpha_k ps_k
// The ps_k update of C3 is carried out by the next iteration's C2, which reads ps_k
// and the unchanged r anyway: each running shift then costs one ps_k read and write
// fewer per iteration (C2 applies the pending update first, with the same operands
// and operations, so every bit is the one C3 would write).
// The scalar recurrences of the reference (zeta/beta/alpha per shift, the
// per-shift residual rr*zeta_k^2/|b|^2 < res_k stop) run in the finalisation
// kernels on the device.

#include "nq_device.hpp"

namespace nq {

template <int SP>
using RealOfS = typename std::conditional<SP == SP_F64, double, float>::type;

// LDS staging of colour fields (three complex per site): the records of the 256
// sites of a workgroup for each of its right-hand sides, [record][site].  FP64:
// three 16-byte records; FP32: one 16-byte record (components 0, 1) and the
// 8-byte leftover (component 2).  The decoded values are those of a memory load.
template <int SP>
struct StStage {
  static constexpr int BYTES = SP == SP_F64 ? 3 * kStageSites * 16 : kStageSites * 24;
  __device__ __forceinline__ static void put(char* lds, const SpinorView& v, int64_t stride, uint32_t e, bool valid,
                                             int t, int r)
  {
    constexpr int EB = SP == SP_F64 ? 16 : 8;
    const char* base = (const char*)v.data + (int64_t)r * v.rstride * EB;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, (int64_t)3 * stride * EB);
    if constexpr (SP == SP_F64) {
#pragma unroll
      for (int q = 0; q < 3; q++) {
        nq_u4 u = {0u, 0u, 0u, 0u};
        if (valid) u = __builtin_amdgcn_raw_buffer_load_b128(rs, e * 16u, (uint32_t)(q * stride * 16), 0);
        *(nq_u4*)(lds + (q * kStageSites + t) * 16) = u;
      }
    } else {
      nq_u4 u = {0u, 0u, 0u, 0u};
      nq_fv2 f = {0.f, 0.f};
      if (valid) {
        u = __builtin_amdgcn_raw_buffer_load_b128(rs, e * 16u, 0u, 0);
        f = __builtin_bit_cast(nq_fv2, __builtin_amdgcn_raw_buffer_load_b64(rs, e * 8u, (uint32_t)(stride * 16), 0));
      }
      *(nq_u4*)(lds + t * 16) = u;
      *(nq_fv2*)(lds + kStageSites * 16 + t * 8) = f;
    }
  }
  template <class R>
  __device__ __forceinline__ static void get(const char* lds, int d, cx<R> s[3])
  {
    if constexpr (SP == SP_F64) {
#pragma unroll
      for (int q = 0; q < 3; q++) decode_record<16, R>(*(const nq_u4*)(lds + (q * kStageSites + d) * 16), s + q);
    } else {
      decode_record<8, R>(*(const nq_u4*)(lds + d * 16), s);
      const nq_fv2 f = *(const nq_fv2*)(lds + kStageSites * 16 + d * 8);
      s[2] = mk<R>((R)f.x, (R)f.y);
    }
  }
};

// One link slot of the hop applied to the nrb <= RB right-hand sides r0.. of one
// site: the link is loaded once and multiplies every right-hand side's
// neighbour.  With a full group the neighbour loads are issued together, with no
// branch between them; right-hand sides that idle are computed and dropped by
// the caller.  Each right-hand side runs the same unrolled code for every RB, so
// a result does not depend on how many right-hand sides share the wavefront.
template <int SP, int LK, class R, int MU, bool FWD, bool PIN, int RB>
__device__ __forceinline__ void st_slot(const Geo& g, const Site& s, uint32_t e, const SpinorView& in, int r0,
                                        int nrb, const void* ub, cx<R> o[RB][3], const char* lds, uint32_t lds_e0)
{
  const uint32_t n = neigh_eo(g, s, MU, FWD);
  cx<R> u[9];
  constexpr int slot = 2 * MU + (FWD ? 0 : 1);
  const double det[2] = {1, 0};
  LkAcc<LK, R, (RB > 1 ? kRbLinkCP : kLinkCP)>::load(ub, g.stride, slot, e, det, u);
  cx<R> v[RB][3];
  const uint32_t d = n - lds_e0;
  if (d < (uint32_t)kStageSites) {  // neighbour staged in LDS
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (RB == 1 || b < nrb) StStage<SP>::get(lds + b * StStage<SP>::BYTES, (int)d, v[b]);
  } else if (RB == 1 || nrb == RB) {
#pragma unroll
    for (int b = 0; b < RB; b++) SpAcc<SP, R, 3>::load(in, g.stride, n, v[b], r0 + b);
  } else {
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (b < nrb) SpAcc<SP, R, 3>::load(in, g.stride, n, v[b], r0 + b);  // uniform over the wavefront
  }
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (RB > 1 && b >= nrb) break;
    cx<R> w[3];
    mat_vec(u, v[b], w);
#pragma unroll
    for (int c = 0; c < 3; c++) o[b][c] = FWD ? o[b][c] + w[c] : o[b][c] - w[c];
    if constexpr (PIN) {
#pragma unroll
      for (int c = 0; c < 3; c++) asm volatile("" : "+v"(o[b][c].re), "+v"(o[b][c].im));
    }
  }
  if constexpr (PIN) asm volatile("" ::: "memory");
}

template <int SP, int LK, class R, int RB>
__device__ __forceinline__ void st_hop_site(const Geo& g, int op, uint32_t e, const SpinorView& in, int r0, int nrb,
                                            const LinkView& U, cx<R> o[RB][3], const char* lds, uint32_t lds_e0)
{
  const Site s = site_of(g, op, e);
  const void* ub = U.data[op];
#pragma unroll
  for (int b = 0; b < RB; b++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[b][c] = mk<R>(0, 0);
  st_slot<SP, LK, R, 0, true, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 0, false, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 1, true, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 1, false, true, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 2, true, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 2, false, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 3, true, false, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
  st_slot<SP, LK, R, 3, false, true, RB>(g, s, e, in, r0, nrb, ub, o, lds, lds_e0);
}

// right-hand side r is skipped when skip_flags[r*skip_stride] (or, with skip_n > 1,
// the next flag) is set
__device__ __forceinline__ bool st_skipped(const int* skip_flags, int skip_stride, int skip_n, int r)
{
  return skip_flags && (skip_flags[r * skip_stride] || (skip_n > 1 && skip_flags[r * skip_stride + 1]));
}

// Stage the records of the workgroup's 256 sites of the nrb right-hand sides r0..
// of `in` (every thread of the workgroup takes part); returns the first site.
template <int SP, int RB>
__device__ __forceinline__ uint32_t stage_sources(char* lds, const Geo& g, const SpinorView& in, const Map& m, int r0,
                                                  int nrb)
{
  const int t = (int)threadIdx.x;
  const uint32_t e0 = (uint32_t)((m.chunk - (t >> 6)) * 64);
#pragma unroll
  for (int b = 0; b < RB; b++)
    if (b < nrb) StStage<SP>::put(lds + b * StStage<SP>::BYTES, in, g.stride, e0 + t, e0 + t < (uint64_t)g.Vh, t, r0 + b);
  __syncthreads();
  return e0;
}

// out = c1 hop + c2 aux (real coefficients, per rhs from coef[2r], coef[2r+1] when
// tabulated). MODE 0: write; 1: write (if write) + |out|^2; 4: write (if write) + Re<aux,out>;
// 6 (first half of the batched CG matvec, out = c1 hop): write (if write) +
// sigma_r |in|^2 + 1/4 |out|^2 = <in, A(sigma_r) in> for A = sigma - 1/4 Deo Doe with
// sigma_r = coef[2r+1], |in|^2 taken from the staged input sites of the lanes' own
// indices (each input site is staged by exactly one workgroup).
// Wavefront w of a workgroup takes chunk 4*cb + w and the RB right-hand sides of
// group m.r (Shape W = 4, NRB = 1, NR = NRG = number of RB groups).
// (one logical block lb; lds: RB staging areas; shared by k_st_hop and the persistent CGM)
template <int SP, int LK, int MODE, int RB>
__device__ __forceinline__ void st_hop_block(int64_t lb, char* lds, const Geo& g, int op, const SpinorView& out,
                                             const SpinorView& in, const LinkView& U, double c1, double c2,
                                             const double* coef, const SpinorView& aux, int write, const Shape& S,
                                             int nr, int64_t nchunks, double* partial, const int* skip_flags,
                                             int skip_stride, int skip_n)
{
  using R = RealOfS<SP>;
  const Map m = map_thread_lb(S, g.Vh, lb);
  const int r0 = m.r * RB, nrb = min(RB, nr - r0);
  unsigned busy = 0;
#pragma unroll
  for (int b = 0; b < RB; b++)
    if (b < nrb && !st_skipped(skip_flags, skip_stride, skip_n, r0 + b)) busy |= 1u << b;
  if (!busy) {  // uniform over the workgroup (one group of right-hand sides)
    if constexpr (MODE != 0) {
#pragma unroll
      for (int b = 0; b < RB; b++)
        if (b < nrb) store_partial(0.0, partial, nchunks, r0 + b, m.chunk);
    }
    return;
  }
  const uint32_t e0 = stage_sources<SP, RB>(lds, g, in, m, r0, nrb);
  // sites past the end (a partial last chunk) compute nothing but take part in
  // the wavefront reductions with zero
  cx<R> o[RB][3];
  if (m.active) st_hop_site<SP, LK, R, RB>(g, op, m.e, in, r0, nrb, U, o, lds, e0);
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (b >= nrb) continue;
    const int r = r0 + b;
    double red = 0;
    if (m.active && ((busy >> b) & 1u)) {
      const bool tab = coef && MODE != 6;
      const double a = tab ? coef[2 * r] : c1, bb = tab ? coef[2 * r + 1] : c2;
#pragma unroll
      for (int c = 0; c < 3; c++) o[b][c] = scal(o[b][c], (R)a);
      if (aux.data) {
        cx<R> x[3];
        SpAcc<SP, R, 3, kNT>::load(aux, g.stride, m.e, x, r);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          o[b][c].re = fma((R)bb, x[c].re, o[b][c].re);
          o[b][c].im = fma((R)bb, x[c].im, o[b][c].im);
          if constexpr (MODE == 4) red += (double)x[c].re * o[b][c].re + (double)x[c].im * o[b][c].im;
        }
      }
      if (MODE == 0 || write) SpAcc<SP, R, 3, kNT>::store(out, g.stride, m.e, o[b], r);
      if constexpr (MODE == 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) red += (double)o[b][c].re * o[b][c].re + (double)o[b][c].im * o[b][c].im;
      }
      if constexpr (MODE == 6) {
        cx<R> pv[3];
        StStage<SP>::get(lds + b * StStage<SP>::BYTES, (int)(m.e - e0), pv);
        double p2 = 0, t2 = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          p2 += (double)pv[c].re * pv[c].re + (double)pv[c].im * pv[c].im;
          t2 += (double)o[b][c].re * o[b][c].re + (double)o[b][c].im * o[b][c].im;
        }
        red = coef[2 * r + 1] * p2 + 0.25 * t2;
      }
    }
    if constexpr (MODE != 0) store_partial(red, partial, nchunks, r, m.chunk);
  }
}

template <int SP, int LK, int MODE, int RB>
__global__ void __launch_bounds__(256) k_st_hop(Geo g, int op, SpinorView out, SpinorView in, LinkView U, double c1,
                                                 double c2, const double* coef, SpinorView aux, int write, Shape S,
                                                 int nr, int64_t nchunks, double* partial, const int* skip_flags,
                                                 int skip_stride, int skip_n)
{
  __shared__ __attribute__((aligned(16))) char lds[RB * StStage<SP>::BYTES];
  st_hop_block<SP, LK, MODE, RB>(xcd_block(), lds, g, op, out, in, U, c1, c2, coef, aux, write, S, nr, nchunks,
                                 partial, skip_flags, skip_stride, skip_n);
}

// Second half of the batched CG matvec fused with the CG update: at even site e of
// system r, s = coef[2r] hop(t) + coef[2r+1] p (-1/4 Deo t + sigma_r p), then
// x += alpha_r p ; r -= alpha_r s ; |r|^2 partials, the update in FP64 on the
// loaded values as the BLAS update kernel does.  Systems that are done or wait for
// a reliable update idle.
template <int SP, int LK, int RB>
__global__ void __launch_bounds__(256) k_st_cgtail(Geo g, SpinorView t, LinkView U, const double* coef, SpinorView p,
                                                    SpinorView x, SpinorView rv, const CgState* st, Shape S, int nr,
                                                    int64_t nchunks, double* partial)
{
  using R = RealOfS<SP>;
  const Map m = map_thread(S, g.Vh);
  const int r0 = m.r * RB, nrb = min(RB, nr - r0);
  unsigned busy = 0;
#pragma unroll
  for (int b = 0; b < RB; b++)
    if (b < nrb && !st[r0 + b].done && !st[r0 + b].reliable) busy |= 1u << b;
  if (!busy) {  // uniform over the workgroup
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (b < nrb) store_partial(0.0, partial, nchunks, r0 + b, m.chunk);
    return;
  }
  __shared__ __attribute__((aligned(16))) char lds[RB * StStage<SP>::BYTES];
  const uint32_t e0 = stage_sources<SP, RB>(lds, g, t, m, r0, nrb);
  cx<R> o[RB][3];
  if (m.active) st_hop_site<SP, LK, R, RB>(g, 0, m.e, t, r0, nrb, U, o, lds, e0);
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (b >= nrb) continue;
    const int r = r0 + b;
    double red = 0;
    if (m.active && ((busy >> b) & 1u)) {
      const double a = coef[2 * r], bb = coef[2 * r + 1];
      cx<R> pv[3];
      SpAcc<SP, R, 3, kNT>::load(p, g.stride, m.e, pv, r);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        o[b][c] = scal(o[b][c], (R)a);
        o[b][c].re = fma((R)bb, pv[c].re, o[b][c].re);
        o[b][c].im = fma((R)bb, pv[c].im, o[b][c].im);
      }
      const double alpha = st[r].alpha;
      cx<R> xv[3], rr[3];
      SpAcc<SP, R, 3, kNT>::load(x, g.stride, m.e, xv, r);
      SpAcc<SP, R, 3, kNT>::load(rv, g.stride, m.e, rr, r);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double xr = fma(alpha, (double)pv[c].re, (double)xv[c].re);
        const double xi = fma(alpha, (double)pv[c].im, (double)xv[c].im);
        const double rre = fma(-alpha, (double)o[b][c].re, (double)rr[c].re);
        const double rim = fma(-alpha, (double)o[b][c].im, (double)rr[c].im);
        red += rre * rre + rim * rim;
        xv[c] = mk<R>((R)xr, (R)xi);
        rr[c] = mk<R>((R)rre, (R)rim);
      }
      SpAcc<SP, R, 3, kNT>::store(x, g.stride, m.e, xv, r);
      SpAcc<SP, R, 3, kNT>::store(rv, g.stride, m.e, rr, r);
    }
    store_partial(red, partial, nchunks, r, m.chunk);
  }
}

// ------------------------------------------------------------------ CGM
// C2: s = -1/4 Deo t (even output, never stored) ; sol_k -= betas_k ps_k ; r += betaa s ; |r|^2
// sol/ps of shift k of rhs r: base + r*rstride + k*sstride (elements)
// (RB right-hand sides per wavefront as in k_st_hop)
template <int SP, int LK, int RB>
__device__ __forceinline__ void cgm_hop2_block(int64_t lb, char* lds, const Geo& g, const SpinorView& t,
                                               const LinkView& U, const SpinorView& rv, const SpinorView& sol,
                                               const SpinorView& ps, int64_t sstride, const CgmState* st,
                                               const Shape& SH, int nr, int64_t nchunks, double* partial)
{
  using R = RealOfS<SP>;
  const Map m = map_thread_lb(SH, g.Vh, lb);
  const int r0 = m.r * RB, nrb = min(RB, nr - r0);
  unsigned busy = 0;
#pragma unroll
  for (int b = 0; b < RB; b++)
    if (b < nrb && !st[r0 + b].done && !st[r0 + b].reliable) busy |= 1u << b;
  if (!busy) {  // uniform over the workgroup
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (b < nrb) store_partial(0.0, partial, nchunks, r0 + b, m.chunk);
    return;
  }
  const uint32_t e0 = stage_sources<SP, RB>(lds, g, t, m, r0, nrb);
  cx<R> o[RB][3];
  if (m.active) st_hop_site<SP, LK, R, RB>(g, 0, m.e, t, r0, nrb, U, o, lds, e0);
#pragma unroll
  for (int b = 0; b < RB; b++) {
    if (b >= nrb) continue;
    const int r = r0 + b;
    const CgmState* S = st + r;
    double red = 0;
    if (m.active && ((busy >> b) & 1u)) {
      const R betaa = (R)S->betaa;
      cx<R> rr[3];
      SpAcc<SP, R, 3, kNT>::load(rv, g.stride, m.e, rr, r);
      const bool pend = S->pend != 0;
      // lag: the last hop2 left its terms betas_prev_k ps_k(previous) to this one
      // (then pend is set too, so the stored ps_k is still the previous iteration's)
      const bool lag = S->lag != 0;
      // shifts in groups of G, every load of a group issued before its first store
      // (one memory wait per group instead of one per shift)
      constexpr int G = SP == SP_F64 ? 2 : 4;
      const int64_t kb = sstride * (SP == SP_F64 ? 16 : 8);  // bytes between shifts
      for (int k0 = 0; k0 < S->nshift; k0 += G) {
        bool on[G];
        bool any = false;
#pragma unroll
        for (int j = 0; j < G; j++) {
          on[j] = k0 + j < S->nshift && S->run[k0 + j];
          any = any || on[j];
        }
        if (!any) continue;  // uniform over the wavefront
        cx<R> q[G][3], qo[G][3], x[G][3];
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (!on[j]) continue;
          SpinorView pk = ps;
          pk.data = (char*)ps.data + (k0 + j) * kb;
          SpAcc<SP, R, 3, kNT>::load(pk, g.stride, m.e, q[j], r);
        }
        if (lag) {
#pragma unroll
          for (int j = 0; j < G; j++) {
            if (!on[j]) continue;
            SpinorView sk = sol;
            sk.data = (char*)sol.data + (k0 + j) * kb;
            SpAcc<SP, R, 3, kNT>::load(sk, g.stride, m.e, x[j], r);
          }
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (!on[j]) continue;
          const int k = k0 + j;
#pragma unroll
          for (int c = 0; c < 3; c++) qo[j][c] = q[j][c];
          if (pend) {
            // the last iteration's C3 update ps_k = zeta_k r + alpha_k ps_k (r not yet
            // updated; zas_k holds the zfs_k that C3 would use, alphas_k is unchanged)
            const R z = (R)S->zas[k], al = (R)S->alphas[k];
#pragma unroll
            for (int c = 0; c < 3; c++) {
              q[j][c].re = fma(al, q[j][c].re, z * rr[c].re);
              q[j][c].im = fma(al, q[j][c].im, z * rr[c].im);
            }
            SpinorView pk = ps;
            pk.data = (char*)ps.data + k * kb;
            SpAcc<SP, R, 3, kNT>::store(pk, g.stride, m.e, q[j], r);
          }
          if (!lag) continue;  // this iteration's term waits for the next hop2
          // sol_k -= betas_prev_k ps_k(previous) ; sol_k -= betas_k ps_k, in that
          // order: the bits of two separate updates
          const R bp = (R)S->betas_prev[k], bk = (R)S->betas[k];
#pragma unroll
          for (int c = 0; c < 3; c++) {
            x[j][c].re = fma(-bp, qo[j][c].re, x[j][c].re);
            x[j][c].im = fma(-bp, qo[j][c].im, x[j][c].im);
          }
#pragma unroll
          for (int c = 0; c < 3; c++) {
            x[j][c].re = fma(-bk, q[j][c].re, x[j][c].re);
            x[j][c].im = fma(-bk, q[j][c].im, x[j][c].im);
          }
          SpinorView sk = sol;
          sk.data = (char*)sol.data + k * kb;
          SpAcc<SP, R, 3, kNT>::store(sk, g.stride, m.e, x[j], r);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const cx<R> sv = scal(o[b][c], (R)-0.25);
        rr[c].re = fma(betaa, sv.re, rr[c].re);
        rr[c].im = fma(betaa, sv.im, rr[c].im);
        red += (double)rr[c].re * rr[c].re + (double)rr[c].im * rr[c].im;
      }
      SpAcc<SP, R, 3, kNT>::store(rv, g.stride, m.e, rr, r);
    }
    store_partial(red, partial, nchunks, r, m.chunk);
  }
}

template <int SP, int LK, int RB>
__global__ void __launch_bounds__(256) k_cgm_hop2(Geo g, SpinorView t, LinkView U, SpinorView rv, SpinorView sol,
                                                   SpinorView ps, int64_t sstride, CgmState* st, Shape SH, int nr,
                                                   int64_t nchunks, double* partial)
{
  __shared__ __attribute__((aligned(16))) char lds[RB * StStage<SP>::BYTES];
  cgm_hop2_block<SP, LK, RB>(xcd_block(), lds, g, t, U, rv, sol, ps, sstride, st, SH, nr, nchunks, partial);
}

// C3: p = r + alpha p ; ps_k = zfs_k r + alphas_k ps_k for running shifts (the ps_k
// part only when no update is pending for C2: before a reliable update, which
// replaces r, and never otherwise; it then applies the sol_k term that C2 left)
template <int SP>
__device__ __forceinline__ void cgm_update_block(int64_t lb, const Geo& g, const SpinorView& rv, const SpinorView& p,
                                                 const SpinorView& ps, const SpinorView& sol, int64_t sstride,
                                                 const CgmState* st, const Shape& SH)
{
  using R = RealOfS<SP>;
  const Map m = map_thread_lb(SH, g.Vh, lb);
  const CgmState* S = st + m.r;
  // reliable == 1: the update was flagged in this iteration, whose p/ps update
  // still runs; 2: waiting for the host's reliable update
  if (!m.active || S->done || S->reliable == 2) return;
  cx<R> rr[3], pp[3];
  SpAcc<SP, R, 3>::load(rv, g.stride, m.e, rr, m.r);
  SpAcc<SP, R, 3>::load(p, g.stride, m.e, pp, m.r);
  const R a = (R)S->alpha;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    pp[c].re = fma(a, pp[c].re, rr[c].re);
    pp[c].im = fma(a, pp[c].im, rr[c].im);
  }
  SpAcc<SP, R, 3>::store(p, g.stride, m.e, pp, m.r);
  if (S->pend) return;  // the ps_k update runs in the next iteration's C2
  for (int k = 0; k < S->nshift; k++) {
    if (!S->run[k]) continue;
    SpinorView pk = ps;
    pk.data = (char*)ps.data + k * sstride * (SP == SP_F64 ? 16 : 8);
    cx<R> q[3];
    SpAcc<SP, R, 3, kNT>::load(pk, g.stride, m.e, q, m.r);
    if (S->owe[k] == 2) {
      // a reliable update follows: the sol_k term of this iteration's C2 first
      SpinorView sk = sol;
      sk.data = (char*)sol.data + k * sstride * (SP == SP_F64 ? 16 : 8);
      cx<R> x[3];
      SpAcc<SP, R, 3, kNT>::load(sk, g.stride, m.e, x, m.r);
      const R bk = (R)S->betas[k];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        x[c].re = fma(-bk, q[c].re, x[c].re);
        x[c].im = fma(-bk, q[c].im, x[c].im);
      }
      SpAcc<SP, R, 3, kNT>::store(sk, g.stride, m.e, x, m.r);
    }
    const R z = (R)S->zfs[k], al = (R)S->alphas[k];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      q[c].re = fma(al, q[c].re, z * rr[c].re);
      q[c].im = fma(al, q[c].im, z * rr[c].im);
    }
    SpAcc<SP, R, 3, kNT>::store(pk, g.stride, m.e, q, m.r);
  }
}

template <int SP>
__global__ void __launch_bounds__(256) k_cgm_update(Geo g, SpinorView rv, SpinorView p, SpinorView ps,
                                                     SpinorView sol, int64_t sstride, CgmState* st, Shape SH)
{
  cgm_update_block<SP>(xcd_block(), g, rv, p, ps, sol, sstride, st, SH);
}

enum { FIN_CGM_PAP = 1, FIN_CGM_RFRF = 2 };

// recurrences of cgm_invert_template_threaded.hpp:118-141 (pap) and :183-252 (rfrf)
__device__ inline void cgm_finalize(CgmState* s, int fin, double total, double pap_scale)
{
  if (fin == FIN_CGM_PAP) {
    const double pap = pap_scale * total;
    s->pap = pap;
    s->betap = s->betaa;
    s->betaa = -s->rr / pap;
    for (int k = 0; k < s->nshift; k++) {
      if (!s->run[k]) continue;
      const double ratio = s->betap / (s->betaa * s->alpha * (1 - s->zas[k] / s->zps[k]) +
                                       s->betap * (1 - s->shift[k] * s->betaa));
      if (isnan(ratio)) { s->nan = 1; s->done = 1; }
      s->zfs[k] = s->zas[k] * ratio;
      s->betas_prev[k] = s->betas[k];
      s->betas[k] = s->betaa * ratio;
    }
  } else {
    const double rfrf = total;
    s->rfrf = rfrf;
    s->alpha = rfrf / s->rr;
    for (int k = 0; k < s->nshift; k++) {
      if (!s->run[k]) continue;
      s->alphas[k] = s->alpha * s->zfs[k] * s->betas[k] / (s->zas[k] * s->betaa);
      s->zps[k] = s->zas[k];
      s->zas[k] = s->zfs[k];
    }
    s->rr = rfrf;
    s->iter += 1;
    // the C2 just run left its sol_k terms when lag was clear (the next C2 applies
    // them); a shift that stops, or a source that ends, owes them to the settle kernel
    const bool owed = s->lag == 0;
    for (int k = 0; k < s->nshift; k++) {
      if (!s->run[k]) continue;
      s->owe[k] = 0;
      s->res[k] = s->rr * s->zfs[k] * s->zfs[k] / s->b2;
      if (s->res[k] < s->target[k]) {
        s->run[k] = 0;
        s->nrun -= 1;
        if (owed) s->owe[k] = 1;
      }
    }
    if (isnan(rfrf)) { s->nan = 1; s->done = 1; }
    if (s->nrun <= 0 || s->iter >= s->niter) s->done = 1;
    if (s->done && owed)
      for (int k = 0; k < s->nshift; k++)
        if (s->run[k]) s->owe[k] = 1;
    if (s->mixed && !s->done) {
      int kr = -1;
      for (int k = 0; k < s->nshift; k++)
        if (s->run[k] && (kr < 0 || s->shift[k] < s->shift[kr])) kr = k;
      s->kref = kr;
      if (kr >= 0 && s->res[kr] < s->delta2 * s->rel_ref) s->reliable = 1;
    }
    // before a reliable update C3 applies the running shifts' owed terms
    if (s->reliable && owed)
      for (int k = 0; k < s->nshift; k++)
        if (s->run[k]) s->owe[k] = 2;
    // C3 leaves the ps_k update to the next C2 unless the source stops here or a
    // reliable update (which replaces r) comes first
    s->pend = !s->done && !s->reliable;
    s->lag = owed && s->pend;
  }
}

__global__ void __launch_bounds__(1024) k_cgm_finalize(const double* partial, int64_t nchunks, CgmState* st, int fin,
                                                       double pap_scale)
{
  __shared__ double vl[1024];
  CgmState* s = st + blockIdx.x;
  if (s->done) return;  // uniform over the workgroup
  if (s->reliable) {
    // first finalisation after the flag: the flagged iteration is complete
    // (its p/ps update has run); later launches idle until the host's update
    if (threadIdx.x == 0) s->reliable = 2;
    return;
  }
  const double tot = block_sum_partials(partial + (int64_t)blockIdx.x * nchunks, nchunks, vl);
  if (threadIdx.x == 0) cgm_finalize(s, fin, tot, pap_scale);
}

// ------------------------------------------------------------------ persistent CGM
// A chunk of multi-shift iterations in one cooperative launch: the five phases of an
// iteration (C1, the pap finalisation, C2, the rfrf finalisation, C3) run the same
// block functions as the separate kernels over the same logical blocks, each
// workgroup walking a contiguous share of its XCD's blocks, with a grid barrier
// between phases.  The finalisations sum the chunk partials with 256 threads in the
// 1024-lane order, so every result has the bits of the separate launches.
//
// The barrier waits a bounded time: a workgroup that spins past spin_limit raises
// `abort`, every workgroup leaves at its next barrier, and the host reports the
// launch as failed; no wait outlasts the limit whatever happens to the grid.
struct GridBar {
  unsigned count, gen;
  int abort, pad;
};

struct CgmPArgs {
  Geo g;
  LinkView U;
  SpinorView p, t, r, sol, ps, none;
  int64_t sstride, nchunks, blocks_hop, blocks_upd;
  CgmState* st;
  double* partial;
  Shape sh_hop, sh_upd;
  int nr, niters, upd_threads;
  unsigned spin_limit;
  GridBar* bar;
};

// the logical blocks [lo, hi) of this workgroup's XCD and the step between the
// blocks of one workgroup (workgroup b runs on XCD b % 8, the hardware's order)
__device__ __forceinline__ void persist_range(int64_t nblk, int64_t* first, int64_t* hi, int64_t* step)
{
  const uint32_t b = blockIdx.x, nb = gridDim.x;
  const uint32_t xcd = b & 7u, bi = b >> 3;
  const int64_t nbx = (int64_t)(nb >> 3) + (xcd < (nb & 7u) ? 1 : 0);
  const int64_t q8 = nblk >> 3, r8 = nblk & 7;
  const int64_t lo = (int64_t)xcd * q8 + ((int64_t)xcd < r8 ? (int64_t)xcd : r8);
  *hi = lo + q8 + ((int64_t)xcd < r8 ? 1 : 0);
  *first = lo + bi;
  *step = nbx;
}

// grid barrier; returns false (uniformly over the workgroup) once the grid aborted
__device__ __forceinline__ bool grid_sync(GridBar* bar, unsigned* gen, unsigned spin_limit, int* s_ok)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this workgroup's writes, device-wide
    const unsigned g0 = *gen;
    const unsigned arrived = atomicAdd(&bar->count, 1u) + 1u;
    if (arrived == gridDim.x) {
      atomicExch(&bar->count, 0u);
      __threadfence();
      atomicAdd(&bar->gen, 1u);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g0) {
        if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (++spins > spin_limit) {
          atomicExch(&bar->abort, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    *gen = g0 + 1;
    __threadfence();  // the other workgroups' writes, here
    *s_ok = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  }
  __syncthreads();
  return *s_ok != 0;
}

// one source's finalisation as k_cgm_finalize does it, with 256 threads
__device__ __forceinline__ void cgm_finalize_block(CgmState* s, const double* pr, int64_t nchunks, int fin,
                                                   double pap_scale, double* vl)
{
  if (s->done) return;  // uniform over the workgroup
  if (s->reliable) {
    if (threadIdx.x == 0) s->reliable = 2;
    return;
  }
  const double tot = block_sum_partials_256(pr, nchunks, vl);
  if (threadIdx.x == 0) cgm_finalize(s, fin, tot, pap_scale);
}

template <int SP, int LK, int RB>
__global__ void __launch_bounds__(256) k_cgm_persistent(CgmPArgs A)
{
  constexpr int kStage = RB * StStage<SP>::BYTES;
  __shared__ __attribute__((aligned(16))) char lds[kStage > 8192 ? kStage : 8192];
  __shared__ int s_ok;
  unsigned gen = 0;  // barrier generation (thread 0)
  const int* skip = (const int*)&A.st->done;
  const int sstr = (int)(sizeof(CgmState) / sizeof(int));
  int64_t f, hi, step;
  for (int it = 0; it < A.niters; it++) {
    // C1: t = Doe p, |t|^2 partials
    persist_range(A.blocks_hop, &f, &hi, &step);
    for (int64_t lb = f; lb < hi; lb += step) {
      __syncthreads();  // the last block's LDS reads are done
      st_hop_block<SP, LK, 1, RB>(lb, lds, A.g, 1, A.t, A.p, A.U, 1.0, 0.0, nullptr, A.none, 1, A.sh_hop, A.nr,
                                  A.nchunks, A.partial, skip, sstr, 2);
    }
    if (!grid_sync(A.bar, &gen, A.spin_limit, &s_ok)) return;
    for (int s = blockIdx.x; s < A.nr; s += gridDim.x)
      cgm_finalize_block(A.st + s, A.partial + (int64_t)s * A.nchunks, A.nchunks, FIN_CGM_PAP, 0.25, (double*)lds);
    if (!grid_sync(A.bar, &gen, A.spin_limit, &s_ok)) return;
    // C2: hop + shift and r updates, |r|^2 partials
    for (int64_t lb = f; lb < hi; lb += step) {
      __syncthreads();
      cgm_hop2_block<SP, LK, RB>(lb, lds, A.g, A.t, A.U, A.r, A.sol, A.ps, A.sstride, A.st, A.sh_hop, A.nr,
                                 A.nchunks, A.partial);
    }
    if (!grid_sync(A.bar, &gen, A.spin_limit, &s_ok)) return;
    for (int s = blockIdx.x; s < A.nr; s += gridDim.x)
      cgm_finalize_block(A.st + s, A.partial + (int64_t)s * A.nchunks, A.nchunks, FIN_CGM_RFRF, 0.25, (double*)lds);
    if (!grid_sync(A.bar, &gen, A.spin_limit, &s_ok)) return;
    // C3: p update (and ps before a reliable update)
    int64_t uf, uhi, ustep;
    persist_range(A.blocks_upd, &uf, &uhi, &ustep);
    if ((int)threadIdx.x < A.upd_threads)
      for (int64_t lb = uf; lb < uhi; lb += ustep)
        cgm_update_block<SP>(lb, A.g, A.r, A.p, A.ps, A.sol, A.sstride, A.st, A.sh_upd);
    if (!grid_sync(A.bar, &gen, A.spin_limit, &s_ok)) return;
  }
}

// sol_k -= betas_k ps_k for the shifts whose term the last C2 left and that no
// later C2 applies (owe_k == 1: the shift stopped or the source ended); the host
// launches it after a poll that finds such a shift and then clears the flags
template <int SP>
__global__ void __launch_bounds__(256) k_cgm_settle(Geo g, SpinorView sol, SpinorView ps, int64_t sstride,
                                                    const CgmState* st, Shape SH)
{
  using R = RealOfS<SP>;
  const Map m = map_thread(SH, g.Vh);
  const CgmState* S = st + m.r;
  if (!m.active) return;
  for (int k = 0; k < S->nshift; k++) {
    if (S->owe[k] != 1) continue;
    SpinorView sk = sol, pk = ps;
    sk.data = (char*)sol.data + k * sstride * (SP == SP_F64 ? 16 : 8);
    pk.data = (char*)ps.data + k * sstride * (SP == SP_F64 ? 16 : 8);
    cx<R> x[3], q[3];
    SpAcc<SP, R, 3>::load(sk, g.stride, m.e, x, m.r);
    SpAcc<SP, R, 3>::load(pk, g.stride, m.e, q, m.r);
    const R bk = (R)S->betas[k];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      x[c].re = fma(-bk, q[c].re, x[c].re);
      x[c].im = fma(-bk, q[c].im, x[c].im);
    }
    SpAcc<SP, R, 3>::store(sk, g.stride, m.e, x, m.r);
  }
}

int launch_cgm_settle(nq_ctx* ctx, const SpinorView& sol, const SpinorView& ps, int64_t sstride, void* states, int nr)
{
  const Geo& g = ctx->geo;
  const Launch L = launch_shape(g, nr);
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "cgm_settle", 0.0, &ev);
  if (sol.prec == SP_F64)
    hipLaunchKernelGGL(k_cgm_settle<SP_F64>, dim3((unsigned)L.blocks), dim3(L.threads), 0, ctx->stream, g, sol, ps,
                       sstride, (const CgmState*)states, L.s);
  else
    hipLaunchKernelGGL(k_cgm_settle<SP_F32>, dim3((unsigned)L.blocks), dim3(L.threads), 0, ctx->stream, g, sol, ps,
                       sstride, (const CgmState*)states, L.s);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

// ------------------------------------------------------------------ launchers
static double csite(int sp) { return sp == SP_F64 ? 48.0 : 24.0; }
static double lbytes(int lk) { return lk == LK_F64_18 ? 144.0 : 72.0; }

// Right-hand sides per wavefront of the staggered hops: one for a single right-hand
// side; for FP32 batches of five or more, 6 (12 sources x 12 shifts at 32^3x64: the
// multi-shift hop 1214 against 1245 us with RB = 4; the 16-system refinement CG's
// fused tail 511 against 534-557 us with RB = 8 or 4); else 4 (FP64: RB = 6 takes
// 208 VGPRs and 72 KB of LDS, two wavefronts per SIMD, and measured 2 % slower).
// NQ_ST_RB = 1, 2, 4, 6 or 8 (8: FP32 only) overrides; every choice gives the same
// bits.  The launch shape for RB: W = 4 chunks per workgroup, one RB group per
// wavefront.
static int st_rb(int nr, int sp)
{
  if (nr <= 1) return 1;
  const char* e = getenv("NQ_ST_RB");
  const int v = e ? atoi(e) : (sp == SP_F32 && nr >= 5 ? 6 : 4);
  if (v == 8) return sp == SP_F32 ? 8 : 4;
  return v == 1 || v == 2 || v == 6 ? v : 4;
}
static Shape st_shape(int nr, int RB)
{
  Shape sh;
  sh.W = 4;
  sh.NRB = 1;
  sh.NRG = sh.NR = (nr + RB - 1) / RB;
  return sh;
}

template <int SP, int LK, int MODE>
static void st_hop_go(nq_ctx* ctx, int RB, int op, const SpinorView& out, const SpinorView& in, const LinkView& U,
                      double c1, double c2, const double* coef, const SpinorView& aux, int write, int nr,
                      const int* skip_flags, int skip_stride, int skip_n)
{
  const Geo& g = ctx->geo;
  const Shape sh = st_shape(nr, RB);
  const int64_t nchunks = (g.Vh + 63) / 64;
  const dim3 grid((unsigned)((nchunks + 3) / 4 * sh.NRG)), blk(256);
  switch (RB) {
  case 1: hipLaunchKernelGGL((k_st_hop<SP, LK, MODE, 1>), grid, blk, 0, ctx->stream, g, op, out, in, U, c1, c2, coef, aux,
                             write, sh, nr, nchunks, ctx->partial, skip_flags, skip_stride, skip_n); break;
  case 2: hipLaunchKernelGGL((k_st_hop<SP, LK, MODE, 2>), grid, blk, 0, ctx->stream, g, op, out, in, U, c1, c2, coef, aux,
                             write, sh, nr, nchunks, ctx->partial, skip_flags, skip_stride, skip_n); break;
  case 6: hipLaunchKernelGGL((k_st_hop<SP, LK, MODE, 6>), grid, blk, 0, ctx->stream, g, op, out, in, U, c1, c2, coef, aux,
                             write, sh, nr, nchunks, ctx->partial, skip_flags, skip_stride, skip_n); break;
  case 8:
    if constexpr (SP == SP_F32) {
      hipLaunchKernelGGL((k_st_hop<SP, LK, MODE, 8>), grid, blk, 0, ctx->stream, g, op, out, in, U, c1, c2, coef, aux,
                         write, sh, nr, nchunks, ctx->partial, skip_flags, skip_stride, skip_n);
      break;
    }
    [[fallthrough]];
  default: hipLaunchKernelGGL((k_st_hop<SP, LK, MODE, 4>), grid, blk, 0, ctx->stream, g, op, out, in, U, c1, c2, coef,
                              aux, write, sh, nr, nchunks, ctx->partial, skip_flags, skip_stride, skip_n); break;
  }
}

int launch_st_hop(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U, double c1,
                  double c2, const double* coef_dev, const SpinorView& aux, int mode, int write, int nr, int fin,
                  CgState* st, const int* skip_flags, int skip_stride, int skip_n)
{
  const Geo& g = ctx->geo;
  if (in.ncomp != 3) return fail(NQ_EINVAL, "staggered hop needs colour fields");
  if (nr < 1 || nr > 16) return fail(NQ_EINVAL, "right-hand sides per launch must be 1..16");
  const int64_t nchunks = (g.Vh + 63) / 64;
  if (nchunks * nr > ctx->max_partials) return fail(NQ_EINVAL, "st hop: reduction scratch too small");
  const int sp = in.prec, lk = U.prec;
  if ((mode == 0 || write) && out.prec != sp) return fail(NQ_EINVAL, "st hop: output precision mismatch");
  if (aux.data && aux.prec != sp) return fail(NQ_EINVAL, "st hop: aux precision mismatch");
  hipEvent_t ev = nullptr;
  // links once for all right-hand sides, the fields per right-hand side
  const double by = 8 * lbytes(lk) + csite(sp) * (1 + (aux.data ? 1 : 0) + ((mode == 0 || write) ? 1 : 0)) * nr;
  prof_begin(ctx, mode == 0 ? "st_hop" : "st_hop_red", by * (double)g.Vh, &ev);
  const int RB = st_rb(nr, sp);
#define NQ_ST(SPV, LKV)                                                                                              \
  switch (mode) {                                                                                                    \
  case 0: st_hop_go<SPV, LKV, 0>(ctx, RB, op, out, in, U, c1, c2, coef_dev, aux, write, nr, skip_flags, skip_stride, \
                                 skip_n); break;                                                                     \
  case 1: st_hop_go<SPV, LKV, 1>(ctx, RB, op, out, in, U, c1, c2, coef_dev, aux, write, nr, skip_flags, skip_stride, \
                                 skip_n); break;                                                                     \
  case 6: st_hop_go<SPV, LKV, 6>(ctx, RB, op, out, in, U, c1, c2, coef_dev, aux, write, nr, skip_flags, skip_stride, \
                                 skip_n); break;                                                                     \
  default: st_hop_go<SPV, LKV, 4>(ctx, RB, op, out, in, U, c1, c2, coef_dev, aux, write, nr, skip_flags,             \
                                  skip_stride, skip_n); break;                                                       \
  }
  if (sp == SP_F64 && lk == LK_F64_18) { NQ_ST(SP_F64, LK_F64_18) }
  else if (sp == SP_F32 && lk == LK_F32_18) { NQ_ST(SP_F32, LK_F32_18) }
  else return fail(NQ_EINVAL, "st hop: unsupported precision combination");
#undef NQ_ST
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  if (mode != 0 && fin != FIN_NONE) NQ_TRY(launch_finalize(ctx, nr, st, fin, skip_n > 1 ? 2 : 0));
  return NQ_OK;
}

int launch_st_cgtail(nq_ctx* ctx, const SpinorView& t, const LinkView& U, const double* coef_dev, const SpinorView& p,
                     const SpinorView& x, const SpinorView& r, CgState* st, int fin, int nr)
{
  const Geo& g = ctx->geo;
  if (t.ncomp != 3 || p.ncomp != 3) return fail(NQ_EINVAL, "staggered CG tail needs colour fields");
  if (nr < 1 || nr > 16) return fail(NQ_EINVAL, "right-hand sides per launch must be 1..16");
  const int sp = t.prec, lk = U.prec;
  if (p.prec != sp || x.prec != sp || r.prec != sp) return fail(NQ_EINVAL, "st CG tail: precision mismatch");
  const int64_t nchunks = (g.Vh + 63) / 64;
  if (nchunks * nr > ctx->max_partials) return fail(NQ_EINVAL, "st CG tail: reduction scratch too small");
  hipEvent_t ev = nullptr;
  // links once; per system t gathered, p read, x and r read and written
  prof_begin(ctx, "st_cgtail", (8 * lbytes(lk) + 6 * csite(sp) * nr) * (double)g.Vh, &ev);
  const int RB = st_rb(nr, sp);
  const Shape sh = st_shape(nr, RB);
  const dim3 grid((unsigned)((nchunks + 3) / 4 * sh.NRG)), blk(256);
#define NQ_TL(SPV, LKV, RBV)                                                                                        \
  hipLaunchKernelGGL((k_st_cgtail<SPV, LKV, RBV>), grid, blk, 0, ctx->stream, g, t, U, coef_dev, p, x, r, st, sh, nr, \
                     nchunks, ctx->partial)
  if (sp == SP_F64 && lk == LK_F64_18) {
    if (RB == 1) NQ_TL(SP_F64, LK_F64_18, 1);
    else if (RB == 2) NQ_TL(SP_F64, LK_F64_18, 2);
    else if (RB == 6) NQ_TL(SP_F64, LK_F64_18, 6);
    else NQ_TL(SP_F64, LK_F64_18, 4);
  } else if (sp == SP_F32 && lk == LK_F32_18) {
    if (RB == 1) NQ_TL(SP_F32, LK_F32_18, 1);
    else if (RB == 2) NQ_TL(SP_F32, LK_F32_18, 2);
    else if (RB == 6) NQ_TL(SP_F32, LK_F32_18, 6);
    else if (RB == 8) NQ_TL(SP_F32, LK_F32_18, 8);
    else NQ_TL(SP_F32, LK_F32_18, 4);
  } else return fail(NQ_EINVAL, "st CG tail: unsupported precision combination");
#undef NQ_TL
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return launch_finalize(ctx, nr, st, fin, 2);
}

// sol64_k += sol_k ; sol_k = 0 for all shifts k of source r
template <int SP>
__global__ void __launch_bounds__(256) k_cgm_flush(Geo g, SpinorView x64, SpinorView sol, int64_t sstride, int nshift,
                                                   const CgmState* st, Shape SH, int only_reliable)
{
  using R = RealOfS<SP>;
  const Map m = map_thread(SH, g.Vh);
  const CgmState* S = st + m.r;
  if (!m.active || (only_reliable && (!S->reliable || S->done))) return;
  for (int k = 0; k < nshift; k++) {
    SpinorView xk = x64, sk = sol;
    xk.data = (char*)x64.data + k * sstride * 16;
    sk.data = (char*)sol.data + k * sstride * (SP == SP_F64 ? 16 : 8);
    cx<double> a[3];
    cx<R> b[3];
    SpAcc<SP_F64, double, 3>::load(xk, g.stride, m.e, a, m.r);
    SpAcc<SP, R, 3>::load(sk, g.stride, m.e, b, m.r);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[c] = a[c] + mk<double>((double)b[c].re, (double)b[c].im);
      b[c] = mk<R>(0, 0);
    }
    SpAcc<SP_F64, double, 3>::store(xk, g.stride, m.e, a, m.r);
    SpAcc<SP, R, 3>::store(sk, g.stride, m.e, b, m.r);
  }
}

int launch_cgm_flush(nq_ctx* ctx, const SpinorView& x64, const SpinorView& sol, int64_t sstride, int nshift,
                     void* states, int nr, int only_reliable)
{
  const Geo& g = ctx->geo;
  const Launch L = launch_shape(g, nr);
  if (x64.prec != SP_F64) return fail(NQ_EINVAL, "cgm flush: FP64 accumulators expected");
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "cgm_flush", (sol.prec == SP_F64 ? 64.0 : 56.0) * nshift * nr * (double)g.Vh, &ev);
  if (sol.prec == SP_F64)
    hipLaunchKernelGGL(k_cgm_flush<SP_F64>, grid, blk, 0, ctx->stream, g, x64, sol, sstride, nshift,
                       (const CgmState*)states, L.s, only_reliable);
  else
    hipLaunchKernelGGL(k_cgm_flush<SP_F32>, grid, blk, 0, ctx->stream, g, x64, sol, sstride, nshift,
                       (const CgmState*)states, L.s, only_reliable);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

// ------------------------------------------------------------------ generic CGM vectors
// For operators whose matvec is launched separately (the lx twisted-mass Q^2):
//   v2: sol_k -= betas_k ps_k (running k) ; r += betaa s ; |r|^2
//   v3: p = r + alpha p ; ps_k = zfs_k r + alphas_k ps_k (running k)
// FP64 fields of NC complex components; one source (state st[0]).
template <int NC>
__global__ void __launch_bounds__(256) k_cgm_vec2(Geo g, SpinorView s, SpinorView r, SpinorView sol, SpinorView ps,
                                                  int64_t sstride, const CgmState* st, Shape SH, int64_t nchunks,
                                                  double* partial)
{
  const Map m = map_thread(SH, g.Vh);
  double red = 0;
  if (m.active && !st->done) {
    for (int k = 0; k < st->nshift; k++) {
      if (!st->run[k]) continue;
      SpinorView sk = sol, pk = ps;
      sk.data = (char*)sol.data + k * sstride * 16;
      pk.data = (char*)ps.data + k * sstride * 16;
      cx<double> x[NC], q[NC];
      SpAcc<SP_F64, double, NC, kNT>::load(sk, g.stride, m.e, x);
      SpAcc<SP_F64, double, NC, kNT>::load(pk, g.stride, m.e, q);
      const double b = st->betas[k];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        x[c].re = fma(-b, q[c].re, x[c].re);
        x[c].im = fma(-b, q[c].im, x[c].im);
      }
      SpAcc<SP_F64, double, NC, kNT>::store(sk, g.stride, m.e, x);
    }
    cx<double> rv[NC], sv[NC];
    SpAcc<SP_F64, double, NC>::load(r, g.stride, m.e, rv);
    SpAcc<SP_F64, double, NC, kNT>::load(s, g.stride, m.e, sv);
    const double ba = st->betaa;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      rv[c].re = fma(ba, sv[c].re, rv[c].re);
      rv[c].im = fma(ba, sv[c].im, rv[c].im);
      red += rv[c].re * rv[c].re + rv[c].im * rv[c].im;
    }
    SpAcc<SP_F64, double, NC>::store(r, g.stride, m.e, rv);
  }
  store_partial(red, partial, nchunks, 0, m.chunk);
}

template <int NC>
__global__ void __launch_bounds__(256) k_cgm_vec3(Geo g, SpinorView r, SpinorView p, SpinorView ps, int64_t sstride,
                                                  const CgmState* st, Shape SH)
{
  const Map m = map_thread(SH, g.Vh);
  if (!m.active || st->done) return;
  cx<double> rv[NC], pv[NC];
  SpAcc<SP_F64, double, NC>::load(r, g.stride, m.e, rv);
  SpAcc<SP_F64, double, NC>::load(p, g.stride, m.e, pv);
  const double a = st->alpha;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    pv[c].re = fma(a, pv[c].re, rv[c].re);
    pv[c].im = fma(a, pv[c].im, rv[c].im);
  }
  SpAcc<SP_F64, double, NC>::store(p, g.stride, m.e, pv);
  for (int k = 0; k < st->nshift; k++) {
    if (!st->run[k]) continue;
    SpinorView pk = ps;
    pk.data = (char*)ps.data + k * sstride * 16;
    cx<double> q[NC];
    SpAcc<SP_F64, double, NC, kNT>::load(pk, g.stride, m.e, q);
    const double z = st->zfs[k], al = st->alphas[k];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      q[c].re = fma(al, q[c].re, z * rv[c].re);
      q[c].im = fma(al, q[c].im, z * rv[c].im);
    }
    SpAcc<SP_F64, double, NC, kNT>::store(pk, g.stride, m.e, q);
  }
}

int launch_cgm_finalize(nq_ctx* ctx, void* states, int nr, int fin, double pap_scale)
{
  const int64_t nchunks = (ctx->geo.Vh + 63) / 64;
  hipLaunchKernelGGL(k_cgm_finalize, dim3(nr), dim3(1024), 0, ctx->stream, ctx->partial, nchunks, (CgmState*)states,
                     fin == 1 ? (int)FIN_CGM_PAP : (int)FIN_CGM_RFRF, pap_scale);
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

int launch_cgm_vec2(nq_ctx* ctx, const SpinorView& s, const SpinorView& r, const SpinorView& sol, const SpinorView& ps,
                    int64_t sstride, void* states)
{
  const Geo& g = ctx->geo;
  const Launch L = launch_shape(g, 1);
  if (r.ncomp != 24) return fail(NQ_EINVAL, "cgm vectors: lx spincolor fields expected");
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "cgm_vec2", 0.0, &ev);
  hipLaunchKernelGGL(k_cgm_vec2<24>, dim3((unsigned)L.blocks), dim3(L.threads), 0, ctx->stream, g, s, r, sol, ps, sstride,
                     (const CgmState*)states, L.s, L.nchunks, ctx->partial);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return launch_cgm_finalize(ctx, states, 1, 2, 1.0);
}

int launch_cgm_vec3(nq_ctx* ctx, const SpinorView& r, const SpinorView& p, const SpinorView& ps, int64_t sstride,
                    void* states)
{
  const Geo& g = ctx->geo;
  const Launch L = launch_shape(g, 1);
  if (r.ncomp != 24) return fail(NQ_EINVAL, "cgm vectors: lx spincolor fields expected");
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "cgm_vec3", 0.0, &ev);
  hipLaunchKernelGGL(k_cgm_vec3<24>, dim3((unsigned)L.blocks), dim3(L.threads), 0, ctx->stream, g, r, p, ps, sstride,
                     (const CgmState*)states, L.s);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int cgm_state_bytes() { return (int)sizeof(CgmState); }

template <int SP, int LK, int RB>
static int persistent_go(nq_ctx* ctx, CgmPArgs& A)
{
  const void* fn = (const void*)k_cgm_persistent<SP, LK, RB>;
  static const int occ = [fn] {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, fn, 256, 0) != hipSuccess) o = 0;
    return o;
  }();
  if (occ < 1 || ctx->ncu < 1) return fail(NQ_EDEVICE, "persistent CGM: no resident workgroup");
  const int grid = occ * ctx->ncu;
  void* args[] = {&A};
  NQ_HIP(hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(256), args, 0, ctx->stream));
  return NQ_OK;
}

int launch_cgm_persistent(nq_ctx* ctx, const LinkView& U, const SpinorView& p, const SpinorView& t,
                          const SpinorView& r, const SpinorView& sol, const SpinorView& ps, int64_t sstride,
                          void* states, int nr, int niters, void* bar, unsigned spin_limit)
{
  const Geo& g = ctx->geo;
  const int sp = p.prec, lk = U.prec;
  if (niters < 1) return NQ_OK;
  if (nr < 1 || nr > 16) return fail(NQ_EINVAL, "persistent CGM: 1..16 sources");
  CgmPArgs A;
  A.g = g;
  A.U = U;
  A.p = p;
  A.t = t;
  A.r = r;
  A.sol = sol;
  A.ps = ps;
  A.none.ncomp = 3;
  A.sstride = sstride;
  A.nchunks = (g.Vh + 63) / 64;
  const int RB = st_rb(nr, sp);
  A.sh_hop = st_shape(nr, RB);
  A.blocks_hop = (A.nchunks + 3) / 4 * A.sh_hop.NRG;
  const Launch L = launch_shape(g, nr);
  A.sh_upd = L.s;
  A.blocks_upd = L.blocks;
  A.upd_threads = L.threads;
  A.st = (CgmState*)states;
  A.partial = ctx->partial;
  A.nr = nr;
  A.niters = niters;
  A.spin_limit = spin_limit;
  A.bar = (GridBar*)bar;
  if (A.nchunks * nr > ctx->max_partials) return fail(NQ_EINVAL, "persistent CGM: reduction scratch too small");
  NQ_HIP(hipMemsetAsync(bar, 0, sizeof(GridBar), ctx->stream));
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "cgm_persistent", 0.0, &ev);
  int rc = NQ_OK;
  if (sp == SP_F64 && lk == LK_F64_18) {
    if (RB == 1) rc = persistent_go<SP_F64, LK_F64_18, 1>(ctx, A);
    else if (RB == 2) rc = persistent_go<SP_F64, LK_F64_18, 2>(ctx, A);
    else if (RB == 6) rc = persistent_go<SP_F64, LK_F64_18, 6>(ctx, A);
    else rc = persistent_go<SP_F64, LK_F64_18, 4>(ctx, A);
  } else if (sp == SP_F32 && lk == LK_F32_18) {
    if (RB == 1) rc = persistent_go<SP_F32, LK_F32_18, 1>(ctx, A);
    else if (RB == 2) rc = persistent_go<SP_F32, LK_F32_18, 2>(ctx, A);
    else if (RB == 6) rc = persistent_go<SP_F32, LK_F32_18, 6>(ctx, A);
    else if (RB == 8) rc = persistent_go<SP_F32, LK_F32_18, 8>(ctx, A);
    else rc = persistent_go<SP_F32, LK_F32_18, 4>(ctx, A);
  } else {
    rc = fail(NQ_EINVAL, "persistent CGM: unsupported precision combination");
  }
  prof_end(ctx, ev);
  return rc;
}

int cgm_grid_bar_bytes() { return (int)sizeof(GridBar); }

int launch_cgm_iteration(nq_ctx* ctx, const LinkView& U, const SpinorView& p, const SpinorView& t,
                         const SpinorView& r, const SpinorView& sol, const SpinorView& ps, int64_t sstride,
                         void* states, int nr, int nbusy, int nrun)
{
  const Geo& g = ctx->geo;
  const Launch L = launch_shape(g, nr);
  CgmState* st = (CgmState*)states;
  const int sp = p.prec, lk = U.prec;
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  // C1: t = Doe p ; |t|^2
  SpinorView none;
  none.ncomp = 3;
  hipEvent_t ev = nullptr;
  // ledger bytes (algorithmic, per site): the links once for all sources; per busy
  // source p gathered and t written (C1); t gathered, r read and written, and per
  // running shift ps read and written and, every other iteration, sol read and
  // written (C2, which carries the ps update and pairs the sol updates: 3 fields
  // per running shift on average); r read, p read and written (C3; its ps update
  // runs only before a reliable update)
  const double cs = csite(sp);
  prof_begin(ctx, "cgm_hop1", (8 * lbytes(lk) + 2 * cs * nbusy) * (double)g.Vh, &ev);
  const int RB = st_rb(nr, sp);
  const int* skip = (const int*)&st->done;
  const int sstr = (int)(sizeof(CgmState) / sizeof(int));
  if (sp == SP_F64)
    st_hop_go<SP_F64, LK_F64_18, 1>(ctx, RB, 1, t, p, U, 1.0, 0.0, nullptr, none, 1, nr, skip, sstr, 2);
  else
    st_hop_go<SP_F32, LK_F32_18, 1>(ctx, RB, 1, t, p, U, 1.0, 0.0, nullptr, none, 1, nr, skip, sstr, 2);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  hipLaunchKernelGGL(k_cgm_finalize, dim3(nr), dim3(1024), 0, ctx->stream, ctx->partial, L.nchunks, st,
                     (int)FIN_CGM_PAP, 0.25);
  NQ_HIP(hipGetLastError());
  // C2
  prof_begin(ctx, "cgm_hop2_update", (8 * lbytes(lk) + 3 * cs * nbusy + 3 * cs * nrun) * (double)g.Vh, &ev);
  {
    const Shape sh = st_shape(nr, RB);
    const dim3 grid2((unsigned)((L.nchunks + 3) / 4 * sh.NRG)), blk2(256);
#define NQ_C2(SPV, LKV)                                                                                           \
  switch (RB) {                                                                                                   \
  case 1: hipLaunchKernelGGL((k_cgm_hop2<SPV, LKV, 1>), grid2, blk2, 0, ctx->stream, g, t, U, r, sol, ps, sstride, st, \
                             sh, nr, L.nchunks, ctx->partial); break;                                             \
  case 2: hipLaunchKernelGGL((k_cgm_hop2<SPV, LKV, 2>), grid2, blk2, 0, ctx->stream, g, t, U, r, sol, ps, sstride, st, \
                             sh, nr, L.nchunks, ctx->partial); break;                                             \
  case 6: hipLaunchKernelGGL((k_cgm_hop2<SPV, LKV, 6>), grid2, blk2, 0, ctx->stream, g, t, U, r, sol, ps, sstride, st, \
                             sh, nr, L.nchunks, ctx->partial); break;                                             \
  default: hipLaunchKernelGGL((k_cgm_hop2<SPV, LKV, 4>), grid2, blk2, 0, ctx->stream, g, t, U, r, sol, ps, sstride,  \
                              st, sh, nr, L.nchunks, ctx->partial); break;                                        \
  }
    if (sp == SP_F64) { NQ_C2(SP_F64, LK_F64_18) }
    else if (RB == 8)
      hipLaunchKernelGGL((k_cgm_hop2<SP_F32, LK_F32_18, 8>), grid2, blk2, 0, ctx->stream, g, t, U, r, sol, ps, sstride, st,
                         sh, nr, L.nchunks, ctx->partial);
    else { NQ_C2(SP_F32, LK_F32_18) }
#undef NQ_C2
  }
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  hipLaunchKernelGGL(k_cgm_finalize, dim3(nr), dim3(1024), 0, ctx->stream, ctx->partial, L.nchunks, st,
                     (int)FIN_CGM_RFRF, 0.25);
  NQ_HIP(hipGetLastError());
  // C3
  prof_begin(ctx, "cgm_update", 3 * cs * nbusy * (double)g.Vh, &ev);
  if (sp == SP_F64)
    hipLaunchKernelGGL(k_cgm_update<SP_F64>, grid, blk, 0, ctx->stream, g, r, p, ps, sol, sstride, st, L.s);
  else
    hipLaunchKernelGGL(k_cgm_update<SP_F32>, grid, blk, 0, ctx->stream, g, r, p, ps, sol, sstride, st, L.s);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

// Rooted-staggered force accumulation of one pole (summ_the_rootst_eoimpr_quark_force,
// hmc/fermions/rootst_eoimpr_quark_force.cpp:54-82) on the links of parity par:
//   F[x][mu] += wsign * eta_mu(x) u1b[x][mu] a(x+mu) b(x)^+
// (even: a = v = Doe chi, b = chi, wsign = +w; odd: a = chi, b = v, wsign = -w),
// eta_mu(x) = prod_{nu<mu} (-1)^{x_nu} (getStagphaseOfLx, geometry_lx.cpp:703-724).
// u1b and F are in nissa's eo order ([Vh][4] complex, [Vh][4][3][3] complex); every
// (site, mu) entry belongs to one thread.
__global__ void __launch_bounds__(256) k_st_force(Geo g, int par, SpinorView a, SpinorView b, const double* u1b,
                                                  double wsign, double* F)
{
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const Site s = site_of(g, par, e);
  const int64_t ieo = nissa_lx(g, s.t, s.x, s.y, s.z) >> 1;
  cx<double> bv[3];
  SpAcc<SP_F64, double, 3>::load(b, g.stride, e, bv);
  const int c[4] = {s.t, s.x, s.y, s.z};
  int eta = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (mu > 0) eta *= 1 - 2 * (c[mu - 1] & 1);
    const uint32_t n = neigh_eo(g, s, mu, true);
    cx<double> av[3];
    SpAcc<SP_F64, double, 3>::load(a, g.stride, n, av);
    const double* u = u1b + (ieo * 4 + mu) * 2;
    const cx<double> ph = mk<double>(u[0] * wsign * eta, u[1] * wsign * eta);
    double* f = F + (ieo * 4 + mu) * 18;
#pragma unroll
    for (int c1 = 0; c1 < 3; c1++) {
      const cx<double> pa = ph * av[c1];
#pragma unroll
      for (int c2 = 0; c2 < 3; c2++) {
        const cx<double> t = pa * conjg(bv[c2]);
        f[(c1 * 3 + c2) * 2] += t.re;
        f[(c1 * 3 + c2) * 2 + 1] += t.im;
      }
    }
  }
}

int launch_st_force(nq_ctx* ctx, int par, const SpinorView& a, const SpinorView& b, const double* u1b, double wsign,
                    double* F)
{
  const Geo& g = ctx->geo;
  if (a.ncomp != 3 || b.ncomp != 3 || a.prec != SP_F64 || b.prec != SP_F64)
    return fail(NQ_EINVAL, "staggered force needs FP64 colour fields");
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "st_force", (double)g.Vh * (48.0 * 5 + 4 * (16.0 + 2 * 144.0)), &ev);
  hipLaunchKernelGGL(k_st_force, dim3((unsigned)((g.Vh + 255) / 256)), dim3(256), 0, ctx->stream, g, par, a, b, u1b,
                     wsign, F);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

}  // namespace nq
