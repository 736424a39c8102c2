// Fused BLAS-1 kernels and deterministic reductions.
//
// Reference: the CG/CGM vector updates (cg_invert_template_threaded.hpp:82-124,
// cgm_invert_template_threaded.hpp:148-221) and the scalar products with their
// stride-halving tree (src/base/field.hpp:316-367,863-905).  Here each update is
// one pass that also produces the norm it feeds (per-thread FP64 partials,
// wavefront butterflies, fixed-order chunk sums finalised by the last block),
// and every scalar of the Krylov recurrences stays in device memory.
// Operands may be stored in FP64, FP32 or FP16(+norm); arithmetic is FP64.
// All kernels take the (64-site chunk x right-hand side) mapping; state[r]
// holds the scalars of right-hand side r.

#include "nq_device.hpp"

namespace nq {

template <int NC>
__device__ __forceinline__ void load_any(const SpinorView& v, int64_t stride, uint32_t e, cx<double> s[NC], int r)
{
  if (v.prec == SP_F64) SpAcc<SP_F64, double, NC>::load(v, stride, e, s, r);
  else if (v.prec == SP_F32) SpAcc<SP_F32, double, NC>::load(v, stride, e, s, r);
  else SpAcc<SP_F16, double, NC>::load(v, stride, e, s, r);
}
template <int NC>
__device__ __forceinline__ void store_any(const SpinorView& v, int64_t stride, uint32_t e, const cx<double> s[NC], int r)
{
  if (v.prec == SP_F64) SpAcc<SP_F64, double, NC>::store(v, stride, e, s, r);
  else if (v.prec == SP_F32) SpAcc<SP_F32, double, NC>::store(v, stride, e, s, r);
  else SpAcc<SP_F16, double, NC>::store(v, stride, e, s, r);
}

struct BlasArgs {
  Geo g;
  Shape s;
  int64_t nchunks;
  double* partial;
  CgState* st;
  int fin;
  int pred;  // 0 always, 1 skip rhs when done, 2 skip unless reliable
};

__device__ __forceinline__ bool skip_rhs(const BlasArgs& a, int r)
{
  if (!a.st || !a.pred) return false;
  const CgState* s = a.st + r;
  if (s->done) return true;
  return (a.pred == 2 && !s->reliable) || (a.pred == 3 && s->reliable);
}

// y = a x + b y ; optional |y|^2 -> fin
template <int NC, bool RED>
__global__ void __launch_bounds__(256) k_axpby(BlasArgs A, SpinorView y, double ar, double ai, SpinorView x, double br,
                                                double bi)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const bool skip = skip_rhs(A, m.r);
  double nrm = 0;
  if (m.active && !skip) {
    cx<double> yv[NC], xv[NC];
    const cx<double> a = mk(ar, ai), b = mk(br, bi);
    if (x.data) load_any<NC>(x, A.g.stride, m.e, xv, m.r);
    else
#pragma unroll
      for (int k = 0; k < NC; k++) xv[k] = mk(0.0, 0.0);
    if (br != 0.0 || bi != 0.0) load_any<NC>(y, A.g.stride, m.e, yv, m.r);
    else
#pragma unroll
      for (int k = 0; k < NC; k++) yv[k] = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      yv[k] = a * xv[k] + b * yv[k];
      nrm += yv[k].re * yv[k].re + yv[k].im * yv[k].im;
    }
    store_any<NC>(y, A.g.stride, m.e, yv, m.r);
  }
  if constexpr (RED) store_partial(nrm, A.partial, A.nchunks, m.r, m.chunk);
}

// out = a x + b y + c z (real coefficients; y/z may be absent; out may be absent);
// optional |out|^2 -> fin
template <int NC, bool RED>
__global__ void __launch_bounds__(256) k_lincomb(BlasArgs A, SpinorView out, double a, SpinorView x, double b,
                                                  SpinorView y, double c, SpinorView z)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const bool skip = skip_rhs(A, m.r);
  double nrm = 0;
  if (m.active && !skip) {
    cx<double> o[NC], t[NC];
    load_any<NC>(x, A.g.stride, m.e, o, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) o[k] = scal(o[k], a);
    if (y.data) {
      load_any<NC>(y, A.g.stride, m.e, t, m.r);
#pragma unroll
      for (int k = 0; k < NC; k++) { o[k].re = fma(b, t[k].re, o[k].re); o[k].im = fma(b, t[k].im, o[k].im); }
    }
    if (z.data) {
      load_any<NC>(z, A.g.stride, m.e, t, m.r);
#pragma unroll
      for (int k = 0; k < NC; k++) { o[k].re = fma(c, t[k].re, o[k].re); o[k].im = fma(c, t[k].im, o[k].im); }
    }
#pragma unroll
    for (int k = 0; k < NC; k++) nrm += o[k].re * o[k].re + o[k].im * o[k].im;
    if (out.data) store_any<NC>(out, A.g.stride, m.e, o, m.r);
  }
  if constexpr (RED) store_partial(nrm, A.partial, A.nchunks, m.r, m.chunk);
}

// Re <a, b> -> fin
template <int NC>
__global__ void __launch_bounds__(256) k_dot(BlasArgs A, SpinorView a, SpinorView b)
{
  const Map m = map_thread(A.s, A.g.Vh);
  double acc = 0;
  if (m.active && !skip_rhs(A, m.r)) {
    cx<double> av[NC], bv[NC];
    load_any<NC>(a, A.g.stride, m.e, av, m.r);
    load_any<NC>(b, A.g.stride, m.e, bv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) acc += av[k].re * bv[k].re + av[k].im * bv[k].im;
  }
  store_partial(acc, A.partial, A.nchunks, m.r, m.chunk);
}

// p = r + beta p, beta from state[r]; skipped once that right-hand side is done or
// while it waits for a reliable update.  With x given, first x += alpha p (the CG
// solution update moved out of the matvec tail, so that p is read once per
// iteration)
template <int NC>
__global__ void __launch_bounds__(256) k_pupdate(BlasArgs A, SpinorView p, SpinorView r, SpinorView x)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const CgState* s = A.st + m.r;
  if (s->done || !m.active || (A.pred == 2 ? !s->reliable : s->reliable)) return;
  const double beta = s->beta;
  cx<double> pv[NC], rv[NC];
  load_any<NC>(p, A.g.stride, m.e, pv, m.r);
  load_any<NC>(r, A.g.stride, m.e, rv, m.r);
  if (x.data) {
    const double alpha = s->alpha;
    cx<double> xv[NC];
    load_any<NC>(x, A.g.stride, m.e, xv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      xv[k].re = fma(alpha, pv[k].re, xv[k].re);
      xv[k].im = fma(alpha, pv[k].im, xv[k].im);
    }
    store_any<NC>(x, A.g.stride, m.e, xv, m.r);
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    pv[k].re = fma(beta, pv[k].re, rv[k].re);
    pv[k].im = fma(beta, pv[k].im, rv[k].im);
  }
  store_any<NC>(p, A.g.stride, m.e, pv, m.r);
}

// p_out = r + beta p_in for one right-hand side (p_in, p_out distinct buffers: the
// one-source mixed CG alternates between two).  With x, every other iteration (xstep)
// x += alpha_prev p_prev + alpha p_in, p_prev being the previous iteration's p still in
// p_out: x is read and written half as often, and the x terms of an iteration that
// does not apply them wait in p_out until the next one.
template <int NC>
__global__ void __launch_bounds__(256) k_pupdate_pair(BlasArgs A, SpinorView pin, SpinorView pout, SpinorView r,
                                                      SpinorView x)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const CgState* s = A.st + m.r;
  if (s->done || !m.active || s->reliable) return;
  const double beta = s->beta;
  cx<double> pv[NC], rv[NC];
  load_any<NC>(pin, A.g.stride, m.e, pv, m.r);
  load_any<NC>(r, A.g.stride, m.e, rv, m.r);
  if (x.data && s->xstep) {
    const double a0 = s->alpha_prev, a1 = s->alpha;
    cx<double> xv[NC], qv[NC];
    load_any<NC>(x, A.g.stride, m.e, xv, m.r);
    load_any<NC>(pout, A.g.stride, m.e, qv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      xv[k].re = fma(a1, pv[k].re, fma(a0, qv[k].re, xv[k].re));
      xv[k].im = fma(a1, pv[k].im, fma(a0, qv[k].im, xv[k].im));
    }
    store_any<NC>(x, A.g.stride, m.e, xv, m.r);
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    pv[k].re = fma(beta, pv[k].re, rv[k].re);
    pv[k].im = fma(beta, pv[k].im, rv[k].im);
  }
  store_any<NC>(pout, A.g.stride, m.e, pv, m.r);
}

// x += alpha p ; r -= alpha s ; |r|^2 (plain CG update, cg_invert_template_threaded.hpp:82-102)
template <int NC>
__global__ void __launch_bounds__(256) k_cg_update(BlasArgs A, SpinorView x, SpinorView p, SpinorView r, SpinorView s)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const CgState* S = A.st + m.r;
  double nrm = 0;
  if (m.active && !S->done && !S->reliable) {
    const double alpha = S->alpha;
    cx<double> xv[NC], pv[NC], rv[NC], sv[NC];
    load_any<NC>(x, A.g.stride, m.e, xv, m.r);
    load_any<NC>(p, A.g.stride, m.e, pv, m.r);
    load_any<NC>(r, A.g.stride, m.e, rv, m.r);
    load_any<NC>(s, A.g.stride, m.e, sv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      xv[k].re = fma(alpha, pv[k].re, xv[k].re);
      xv[k].im = fma(alpha, pv[k].im, xv[k].im);
      rv[k].re = fma(-alpha, sv[k].re, rv[k].re);
      rv[k].im = fma(-alpha, sv[k].im, rv[k].im);
      nrm += rv[k].re * rv[k].re + rv[k].im * rv[k].im;
    }
    store_any<NC>(x, A.g.stride, m.e, xv, m.r);
    store_any<NC>(r, A.g.stride, m.e, rv, m.r);
  }
  store_partial(nrm, A.partial, A.nchunks, m.r, m.chunk);
}

// x64 += xs ; xs = 0 (reliable update), only for right-hand sides flagged reliable
template <int NC>
__global__ void __launch_bounds__(256) k_reliable_accum(BlasArgs A, SpinorView x64, SpinorView xs, SpinorView p,
                                                         SpinorView p_prev)
{
  const Map m = map_thread(A.s, A.g.Vh);
  const CgState* s = A.st + m.r;
  if (s->done || !s->reliable || !m.active) return;
  cx<double> a[NC], b[NC];
  load_any<NC>(x64, A.g.stride, m.e, a, m.r);
  load_any<NC>(xs, A.g.stride, m.e, b, m.r);
  if (p_prev.data && s->xlag) {
    // paired x updates: the previous iteration's alpha p, left to the skipped update
    const double alpha = s->alpha_prev;
    cx<double> pv[NC];
    load_any<NC>(p_prev, A.g.stride, m.e, pv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      b[k].re = fma(alpha, pv[k].re, b[k].re);
      b[k].im = fma(alpha, pv[k].im, b[k].im);
    }
  }
  if (p.data) {
    // the last iteration's alpha p, not yet folded into xs
    const double alpha = s->alpha;
    cx<double> pv[NC];
    load_any<NC>(p, A.g.stride, m.e, pv, m.r);
#pragma unroll
    for (int k = 0; k < NC; k++) {
      b[k].re = fma(alpha, pv[k].re, b[k].re);
      b[k].im = fma(alpha, pv[k].im, b[k].im);
    }
  }
#pragma unroll
  for (int k = 0; k < NC; k++) { a[k] = a[k] + b[k]; b[k] = mk(0.0, 0.0); }
  store_any<NC>(x64, A.g.stride, m.e, a, m.r);
  store_any<NC>(xs, A.g.stride, m.e, b, m.r);
}

// out = in (precision conversion); in absent -> zero
template <int NC>
__global__ void __launch_bounds__(256) k_convert(BlasArgs A, SpinorView out, SpinorView in)
{
  const Map m = map_thread(A.s, A.g.Vh);
  if (!m.active || skip_rhs(A, m.r)) return;
  cx<double> v[NC];
  if (in.data) load_any<NC>(in, A.g.stride, m.e, v, m.r);
  else
#pragma unroll
    for (int k = 0; k < NC; k++) v[k] = mk(0.0, 0.0);
  store_any<NC>(out, A.g.stride, m.e, v, m.r);
}

// out_h = a_h x + b_h y for the upper (spins 0,1) and lower (spins 2,3) halves of a
// spincolor: the site-local (1/2k + i mu g5)^(+-1) pieces of the tm operator
__global__ void __launch_bounds__(256) k_spin_diag(BlasArgs A, SpinorView out, SpinorView x, SpinorView y, double aur,
                                                    double aui, double alr, double ali, double bur, double bui,
                                                    double blr, double bli)
{
  const Map m = map_thread(A.s, A.g.Vh);
  if (!m.active || skip_rhs(A, m.r)) return;
  cx<double> o[12], t[12];
  load_any<12>(x, A.g.stride, m.e, t, m.r);
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = (k < 6 ? mk(aur, aui) : mk(alr, ali)) * t[k];
  if (y.data) {
    load_any<12>(y, A.g.stride, m.e, t, m.r);
#pragma unroll
    for (int k = 0; k < 12; k++) fma_to(o[k], k < 6 ? mk(bur, bui) : mk(blr, bli), t[k]);
  }
  store_any<12>(out, A.g.stride, m.e, o, m.r);
}

// One block: for each right-hand side still running, sum its chunk partials with
// 256 lanes (lane v takes chunks v, v+256, ... in four interleaved chains),
// combine the lanes by a fixed pairwise tree and finalise its state.
__global__ void __launch_bounds__(1024) k_finalize(const double* partial, int64_t nchunks, CgState* st, int fin,
                                                   int pred_reliable)
{
  __shared__ double vl[1024];
  CgState* s = st + blockIdx.x;
  // uniform over the workgroup
  if (s->done || (pred_reliable == 1 && !s->reliable) || (pred_reliable == 2 && s->reliable)) return;
  const double tot = block_sum_partials(partial + (int64_t)blockIdx.x * nchunks, nchunks, vl);
  if (threadIdx.x == 0) finalize(s, fin, tot);
}

int launch_finalize(nq_ctx* ctx, int nr, CgState* st, int fin, int pred_reliable)
{
  if (!st || fin == FIN_NONE) return fail(NQ_EINVAL, "finalize without state");
  const int64_t nchunks = (ctx->geo.Vh + 63) / 64;
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "reduce_finalize", 8.0 * nchunks * nr, &ev);
  hipLaunchKernelGGL(k_finalize, dim3(nr), dim3(1024), 0, ctx->stream, ctx->partial, nchunks, st, fin, pred_reliable);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

// ------------------------------------------------------------------ launchers
Launch launch_shape(const Geo& g, int nr)
{
  Launch L;
  L.s.NR = nr;
  L.s.NRB = nr < 4 ? nr : 4;
  L.s.NRG = (nr + L.s.NRB - 1) / L.s.NRB;
  L.s.W = L.s.NRB == 1 ? 4 : (L.s.NRB == 2 ? 2 : 1);
  L.nchunks = (g.Vh + 63) / 64;
  L.blocks = (L.nchunks + L.s.W - 1) / L.s.W * L.s.NRG;
  L.threads = 64 * L.s.W * L.s.NRB;
  return L;
}

static double site_bytes(const SpinorView& v)
{
  if (!v.data) return 0;
  switch (v.prec) {
  case SP_F64: return 16.0 * v.ncomp;
  case SP_F32: return 8.0 * v.ncomp;
  default: return 4.0 * v.ncomp + 4;
  }
}

static int make_args(nq_ctx* ctx, int nr, CgState* st, int fin, int pred, BlasArgs* A, Launch* L)
{
  if (nr < 1 || nr > 16) return fail(NQ_EINVAL, "right-hand sides per launch must be 1..16");
  *L = launch_shape(ctx->geo, nr);
  if (L->nchunks * nr > ctx->max_partials) return fail(NQ_EINVAL, "reduction scratch too small");
  if (((fin != FIN_NONE && fin != FIN_PARTIAL) || pred) && !st) return fail(NQ_EINVAL, "reduction without solver state");
  A->g = ctx->geo;
  A->s = L->s;
  A->nchunks = L->nchunks;
  A->partial = ctx->partial;
  A->st = st;
  A->fin = fin;
  A->pred = pred;
  return NQ_OK;
}

int launch_axpby(nq_ctx* ctx, const SpinorView& y, const double a[2], const SpinorView& x, const double b[2], int fin,
                 CgState* st, int pred, int nr)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, fin, pred, &A, &L));
  const bool red = fin != FIN_NONE;
  hipEvent_t ev = nullptr;
  prof_begin(ctx, red ? "blas_axpby_norm" : "blas_axpby",
             nr * (site_bytes(x) + site_bytes(y) * ((b[0] != 0 || b[1] != 0) ? 2 : 1)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (y.ncomp == 24) {
    if (red) hipLaunchKernelGGL((k_axpby<24, true>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
    else hipLaunchKernelGGL((k_axpby<24, false>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
  } else if (y.ncomp == 12) {
    if (red) hipLaunchKernelGGL((k_axpby<12, true>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
    else hipLaunchKernelGGL((k_axpby<12, false>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
  } else {
    if (red) hipLaunchKernelGGL((k_axpby<3, true>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
    else hipLaunchKernelGGL((k_axpby<3, false>), grid, blk, 0, ctx->stream, A, y, a[0], a[1], x, b[0], b[1]);
  }
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  if (red && fin != FIN_PARTIAL) NQ_TRY(launch_finalize(ctx, nr, st, fin, pred == 2 ? 1 : (pred == 3 ? 2 : 0)));
  return NQ_OK;
}

int launch_lincomb(nq_ctx* ctx, const SpinorView& out, double a, const SpinorView& x, double b, const SpinorView& y,
                   double c, const SpinorView& z, int fin, CgState* st, int nr, int pred)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, fin, pred, &A, &L));
  const bool red = fin != FIN_NONE;
  hipEvent_t ev = nullptr;
  prof_begin(ctx, red ? "blas_lincomb_norm" : "blas_lincomb",
             nr * (site_bytes(out) + site_bytes(x) + site_bytes(y) + site_bytes(z)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (x.ncomp == 24) {
    if (red) hipLaunchKernelGGL((k_lincomb<24, true>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
    else hipLaunchKernelGGL((k_lincomb<24, false>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
  } else if (x.ncomp == 12) {
    if (red) hipLaunchKernelGGL((k_lincomb<12, true>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
    else hipLaunchKernelGGL((k_lincomb<12, false>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
  } else {
    if (red) hipLaunchKernelGGL((k_lincomb<3, true>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
    else hipLaunchKernelGGL((k_lincomb<3, false>), grid, blk, 0, ctx->stream, A, out, a, x, b, y, c, z);
  }
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  if (red && fin != FIN_PARTIAL) NQ_TRY(launch_finalize(ctx, nr, st, fin, pred == 2 ? 1 : (pred == 3 ? 2 : 0)));
  return NQ_OK;
}

int launch_dot(nq_ctx* ctx, const SpinorView& a, const SpinorView& b, int fin, CgState* st, int nr, int pred)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, fin, pred, &A, &L));
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "blas_dot", nr * (site_bytes(a) + site_bytes(b)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (a.ncomp == 24) hipLaunchKernelGGL(k_dot<24>, grid, blk, 0, ctx->stream, A, a, b);
  else if (a.ncomp == 12) hipLaunchKernelGGL(k_dot<12>, grid, blk, 0, ctx->stream, A, a, b);
  else hipLaunchKernelGGL(k_dot<3>, grid, blk, 0, ctx->stream, A, a, b);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  NQ_TRY(launch_finalize(ctx, nr, st, fin, pred == 2 ? 1 : (pred == 3 ? 2 : 0)));
  return NQ_OK;
}

int launch_cg_pupdate(nq_ctx* ctx, const SpinorView& p, const SpinorView& r, CgState* st, int nr, int pred,
                      const SpinorView& x)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, FIN_NONE, pred == 2 ? 2 : 1, &A, &L));
  if (x.data && x.ncomp != p.ncomp) return fail(NQ_EINVAL, "pupdate: component mismatch");
  hipEvent_t ev = nullptr;
  const double xb = x.data ? 2.0 * site_bytes(x) : 0.0;
  prof_begin(ctx, x.data ? "blas_xpupdate" : "blas_pupdate",
             nr * (2 * site_bytes(p) + site_bytes(r) + xb) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (p.ncomp == 24) hipLaunchKernelGGL(k_pupdate<24>, grid, blk, 0, ctx->stream, A, p, r, x);
  else if (p.ncomp == 12) hipLaunchKernelGGL(k_pupdate<12>, grid, blk, 0, ctx->stream, A, p, r, x);
  else hipLaunchKernelGGL(k_pupdate<3>, grid, blk, 0, ctx->stream, A, p, r, x);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_cg_update(nq_ctx* ctx, const SpinorView& x, const SpinorView& p, const SpinorView& r, const SpinorView& s,
                     CgState* st, int fin, int nr)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, fin, 1, &A, &L));
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "blas_cg_update",
             nr * (2 * site_bytes(x) + site_bytes(p) + 2 * site_bytes(r) + site_bytes(s)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (x.ncomp == 24) hipLaunchKernelGGL(k_cg_update<24>, grid, blk, 0, ctx->stream, A, x, p, r, s);
  else if (x.ncomp == 12) hipLaunchKernelGGL(k_cg_update<12>, grid, blk, 0, ctx->stream, A, x, p, r, s);
  else hipLaunchKernelGGL(k_cg_update<3>, grid, blk, 0, ctx->stream, A, x, p, r, s);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return launch_finalize(ctx, nr, st, fin, 2);
}

__global__ void k_clear_reliable(CgState* st, int nr)
{
  if ((int)threadIdx.x < nr) st[threadIdx.x].reliable = 0;
}

int launch_clear_reliable(nq_ctx* ctx, CgState* st, int nr)
{
  if (nr < 1 || nr > 64) return fail(NQ_EINVAL, "clear_reliable: bad count");
  hipLaunchKernelGGL(k_clear_reliable, dim3(1), dim3(64), 0, ctx->stream, st, nr);
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

int launch_cg_pupdate_pair(nq_ctx* ctx, const SpinorView& p_in, const SpinorView& p_out, const SpinorView& r,
                           CgState* st, const SpinorView& x)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, 1, st, FIN_NONE, 1, &A, &L));
  if (p_in.data == p_out.data) return fail(NQ_EINVAL, "pupdate pair: p_in and p_out must differ");
  if (p_in.ncomp != p_out.ncomp || (x.data && x.ncomp != p_in.ncomp)) return fail(NQ_EINVAL, "pupdate pair: component mismatch");
  hipEvent_t ev = nullptr;
  // x and the previous p every other iteration: half their bytes on average
  const double xb = x.data ? 0.5 * (2.0 * site_bytes(x) + site_bytes(p_out)) : 0.0;
  prof_begin(ctx, x.data ? "blas_xpupdate" : "blas_pupdate",
             (site_bytes(p_in) + site_bytes(p_out) + site_bytes(r) + xb) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (p_in.ncomp == 12) hipLaunchKernelGGL(k_pupdate_pair<12>, grid, blk, 0, ctx->stream, A, p_in, p_out, r, x);
  else return fail(NQ_EINVAL, "pupdate pair: spincolor fields expected");
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_reliable_accum(nq_ctx* ctx, const SpinorView& x64, const SpinorView& xs, CgState* st, int nr,
                          const SpinorView& p, const SpinorView& p_prev)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, FIN_NONE, 2, &A, &L));
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "blas_reliable_accum",
             nr * (2 * (site_bytes(x64) + site_bytes(xs)) + (p.data ? site_bytes(p) : 0)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (x64.ncomp == 12) hipLaunchKernelGGL(k_reliable_accum<12>, grid, blk, 0, ctx->stream, A, x64, xs, p, p_prev);
  else hipLaunchKernelGGL(k_reliable_accum<3>, grid, blk, 0, ctx->stream, A, x64, xs, p, p_prev);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_convert(nq_ctx* ctx, const SpinorView& out, const SpinorView& in, int nr, CgState* st, int pred)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, st, FIN_NONE, pred, &A, &L));
  if (in.data && in.ncomp != out.ncomp) return fail(NQ_EINVAL, "convert: component mismatch");
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "blas_convert", nr * (site_bytes(out) + site_bytes(in)) * (double)ctx->geo.Vh, &ev);
  const dim3 grid((unsigned)L.blocks), blk(L.threads);
  if (out.ncomp == 24) hipLaunchKernelGGL(k_convert<24>, grid, blk, 0, ctx->stream, A, out, in);
  else if (out.ncomp == 12) hipLaunchKernelGGL(k_convert<12>, grid, blk, 0, ctx->stream, A, out, in);
  else hipLaunchKernelGGL(k_convert<3>, grid, blk, 0, ctx->stream, A, out, in);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_spin_diag(nq_ctx* ctx, const SpinorView& out, const double a[2][2], const SpinorView& x,
                     const double b[2][2], const SpinorView& y, int nr)
{
  BlasArgs A;
  Launch L;
  NQ_TRY(make_args(ctx, nr, nullptr, FIN_NONE, 0, &A, &L));
  if (x.ncomp != 12 || out.ncomp != 12) return fail(NQ_EINVAL, "spin_diag needs spincolor fields");
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "blas_spin_diag", nr * (site_bytes(out) + site_bytes(x) + site_bytes(y)) * (double)ctx->geo.Vh, &ev);
  hipLaunchKernelGGL(k_spin_diag, dim3((unsigned)L.blocks), dim3(L.threads), 0, ctx->stream, A, out, x, y, a[0][0],
                     a[0][1], a[1][0], a[1][1], b[0][0], b[0][1], b[1][0], b[1][1]);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_zero(nq_ctx* ctx, const SpinorView& v, int nr)
{
  SpinorView none;
  none.ncomp = v.ncomp;
  return launch_convert(ctx, v, none, nr, nullptr, 0);
}

// the scratch state slot used by the synchronous reductions is the last one
static CgState* scratch_state(nq_ctx* ctx) { return ctx->state + (ctx->nstate - 1); }

int reduce_norm2(nq_ctx* ctx, const SpinorView& v, double* out)
{
  SpinorView none;
  none.ncomp = v.ncomp;
  CgState* st = scratch_state(ctx);
  NQ_TRY(launch_lincomb(ctx, none, 1.0, v, 0, none, 0, none, FIN_STORE0, st, 1, 0));
  NQ_HIP(hipMemcpyAsync(&ctx->state_host[ctx->nstate - 1], st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->state_host[ctx->nstate - 1].extra[0];
  return NQ_OK;
}

int reduce_dot(nq_ctx* ctx, const SpinorView& a, const SpinorView& b, double* out)
{
  CgState* st = scratch_state(ctx);
  NQ_TRY(launch_dot(ctx, a, b, FIN_STORE0, st, 1, 0));
  NQ_HIP(hipMemcpyAsync(&ctx->state_host[ctx->nstate - 1], st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->state_host[ctx->nstate - 1].extra[0];
  return NQ_OK;
}

}  // namespace nq
