// Twisted-mass operators and solvers on the device.
//
// Reference path (src/dirac_operators/tmD_eoprec/dirac_operator_tmD_eoprec_portable.cpp,
// src/inverters/twisted_mass/cg_invert_tmD_eoprec.cpp,
// src/inverters/templates/cg_invert_template_threaded.hpp):
//   K(mu)      = g5 (D_oo(mu) - 1/4 H_oe D_ee(mu)^-1 H_eo)           (2 hop launches)
//   K K^+      = K(+mu) K(-mu)                                        (4 hop launches)
//   CG on K K^+ with nissa's squared relative residual stop, then the eo
//   reconstruction (eqs. 8-11 of doc/eo_inverter.lyx) and the FP64 check.
// Every site-local factor (D_ee^-1, D_oo, g5, the 1/4 and 1/2) is folded into
// the hop epilogues; the CG scalars live on the device; the host only polls
// the convergence flag once per chunk of iterations.
//
// <p, K K^+ p> is computed as |K(-mu) p|^2 in the epilogue of the second hop
// (K(+mu)^+ = K(-mu)), so the matvec needs no extra pass for the dot product;
// the x/r updates and |r|^2 are fused into the last hop of the matvec.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <functional>

#include "nq_internal.hpp"
#include "tm_batch.hpp"

namespace nq {

using cplx = std::complex<double>;

static void setc(double c[2][2], cplx up, cplx lo)
{
  c[0][0] = up.real(); c[0][1] = up.imag();
  c[1][0] = lo.real(); c[1][1] = lo.imag();
}

// D_ee(mu)^-1 on the hop: upper (a - i mu)/n, lower (a + i mu)/n  (..._portable.hpp:205-221)
static HopEpi epi_inv_dee(double kappa, double mu)
{
  HopEpi e;
  const double a = 1 / (2 * kappa), n = a * a + mu * mu;
  setc(e.c1, cplx(a / n, -mu / n), cplx(a / n, mu / n));
  return e;
}
// g5 (D_oo(mu) in - hop/4)  (..._portable.cpp:21-32 with tmDee of :179-194)
static HopEpi epi_kout(double kappa, double mu, const SpinorView& in)
{
  HopEpi e;
  const double a = 1 / (2 * kappa);
  setc(e.c1, cplx(-0.25, 0), cplx(0.25, 0));
  setc(e.c2, cplx(a, mu), cplx(-a, mu));
  e.aux = in;
  return e;
}
// Q = g5 D on one parity: g5((1/2k + i mu g5) in - hop/2)  (dirac_operator_tmQ_portable.cpp:164-174)
static HopEpi epi_q(double kappa, double mu, const SpinorView& in)
{
  HopEpi e;
  const double a = 1 / (2 * kappa);
  setc(e.c1, cplx(-0.5, 0), cplx(0.5, 0));
  setc(e.c2, cplx(a, mu), cplx(-a, mu));
  e.aux = in;
  return e;
}

struct TmSetup {
  nq_ctx* ctx;
  LinkView U;
  double anis;
  double kappa;
};

static int tm_setup(nq_ctx* ctx, const nq_gauge* g, int lk, const double* anis, double kappa, TmSetup* s)
{
  if (!g || g->ctx != ctx) return fail(NQ_EINVAL, "gauge handle belongs to another context");
  if (!(kappa != 0) || std::isnan(kappa)) return fail(NQ_EINVAL, "bad kappa");
  s->ctx = ctx;
  s->anis = anis ? *anis : 1.0;
  s->kappa = kappa;
  return link_view(g, lk, &s->U);
}

// FP64 links of the work around a sloppy solve (reliable updates, eo preparation and
// reconstruction): 12-real when the gauge field allows exact compression (two rows
// stored, the third rebuilt in FP64), which cuts the FP64 hop's link bytes by a third;
// the plain FP64 solver and the final FP64 check keep the 18-real links.
static int lk64_for(int prec, const nq_gauge* g)
{
  return (prec != NQ_PREC_F64 && g->compressible) ? LK_F64_12 : LK_F64_18;
}

// K(mu): out_odd = K in_odd, using tmp_even
static int apply_kern(const TmSetup& S, double mu, const SpinorView& out, const SpinorView& in, const SpinorView& tmp,
                      CgState* st_skip = nullptr, int pred = 0)
{
  HopEpi a = epi_inv_dee(S.kappa, mu);
  a.st = st_skip;
  a.pred_reliable = pred;
  NQ_TRY(launch_tm_hop(S.ctx, NQ_EVEN, tmp, in, S.U, S.anis, a));
  HopEpi b = epi_kout(S.kappa, mu, in);
  b.st = st_skip;
  b.pred_reliable = pred;
  return launch_tm_hop(S.ctx, NQ_ODD, out, tmp, S.U, S.anis, b);
}

// r = b - K(+mu)K(-mu) x with |r|^2 -> fin on st (mode 3: residual tail)
static int residual_kernsq(const TmSetup& S, double mu, const SpinorView& r, const SpinorView& rs, const SpinorView& x,
                           const SpinorView& b, const SpinorView& t1, const SpinorView& tmp, CgState* st, int fin,
                           int pred)
{
  NQ_TRY(apply_kern(S, -mu, t1, x, tmp, st, pred));
  HopEpi a = epi_inv_dee(S.kappa, mu);
  a.st = st;
  a.pred_reliable = pred;
  NQ_TRY(launch_tm_hop(S.ctx, NQ_EVEN, tmp, t1, S.U, S.anis, a));
  HopEpi e = epi_kout(S.kappa, mu, t1);
  for (int h = 0; h < 2; h++)
    for (int k = 0; k < 2; k++) { e.c1[h][k] = -e.c1[h][k]; e.c2[h][k] = -e.c2[h][k]; }
  setc(e.c3, 1.0, 1.0);
  e.aux2 = b;
  e.mode = 3;
  e.r = r;
  e.rs = rs;
  e.st = st;
  e.fin = fin;
  e.pred_reliable = pred;
  SpinorView none;
  return launch_tm_hop(S.ctx, NQ_ODD, none, tmp, S.U, S.anis, e);
}

static int lk_of_prec(int prec, int* sp, int* lk)
{
  switch (prec) {
  case NQ_PREC_F64: *sp = SP_F64; *lk = LK_F64_18; return NQ_OK;
  case NQ_PREC_F32: *sp = SP_F32; *lk = LK_F32_18; return NQ_OK;
  case NQ_PREC_F16: *sp = SP_F16; *lk = LK_F32_12; return NQ_OK;
  case NQ_PREC_F16H: *sp = SP_F16; *lk = LK_F16_12; return NQ_OK;
  default: return fail(NQ_EINVAL, "unknown precision mode %d", prec);
  }
}

static CgState* scratch_state(nq_ctx* ctx) { return ctx->state + (ctx->nstate - 1); }

// Drop of |r| between two FP64 reliable updates of a mixed solve: the fewest updates
// for which no drop goes below the smallest one the inner storage keeps reliable
// (FP16 spinors: 3e-3, FP32: 3e-4), spread evenly over the requested reduction
// sqrt(residue) of |r|/|b|.  Measured at 32^3x64 to residue 1e-20: FP16 storage with
// drops of 1e-2 (5 updates) and 3.2e-3 (4) both take 64 iterations, 1e-3 takes 69 and
// 3e-4 99; FP32 takes 64 iterations with drops down to 3e-4 (3 updates).
static double reliable_delta(int sp, double residue)
{
  const double dmin = sp == SP_F16 ? 3e-3 : 3e-4;
  const double red = std::sqrt(std::max(residue, 1e-300));
  if (!(red < dmin)) return dmin;
  const int n = (int)std::ceil(std::log(red) / std::log(dmin) - 1e-9);
  return std::pow(red, 1.0 / n);
}

// Upload a host-initialised state to device slot i (async on the ctx stream).
static int put_state(nq_ctx* ctx, int i, const CgState& s)
{
  ctx->state_host[i] = s;
  NQ_HIP(hipMemcpyAsync(ctx->state + i, ctx->state_host + i, sizeof(CgState), hipMemcpyHostToDevice, ctx->stream));
  return NQ_OK;
}
static int get_states(nq_ctx* ctx, int i0, int n)
{
  NQ_HIP(hipMemcpyAsync(ctx->state_host + i0, ctx->state + i0, sizeof(CgState) * n, hipMemcpyDeviceToHost, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  return NQ_OK;
}

// iterations to enqueue before the next poll, from the observed convergence rate
static int next_chunk(double rr0, double rr, double target, int done_iters, int last_chunk)
{
  if (done_iters <= 0 || !(rr > 0) || !(rr0 > rr)) return last_chunk;
  const double rate = std::log(rr / rr0) / done_iters;  // log decrease per iteration (<0)
  const double need = std::log(target / rr) / rate;
  int n = (int)std::ceil(need) + 1;
  if (n < 2) n = 2;
  if (n > 64) n = 64;
  return n;
}

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st)
  {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  double stop()
  {
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms * 1e-3;
  }
  ~Timer()
  {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// CG on K K^+ (FP64 throughout), cg_invert semantics (cg_invert_template_threaded.hpp:19-150)
// kx (optional, FP64 odd): receives K(-mu) x of the final residual, the eo solve's
// eq. 10, so that the caller need not apply K(-mu) again
static int cg_kernsq_f64(const TmSetup& S, double mu, int niter, double residue, const SpinorView& x,
                         const SpinorView* guess, const SpinorView& b, nq_solve_stats* out,
                         const SpinorView* kx = nullptr)
{
  nq_ctx* ctx = S.ctx;
  SpinorView r, p, t1, tmp;
  NQ_TRY(alloc_spinor(ctx, "cg.r", SP_F64, 12, &r));
  NQ_TRY(alloc_spinor(ctx, "cg.p", SP_F64, 12, &p));
  if (kx) t1 = *kx;
  else NQ_TRY(alloc_spinor(ctx, "cg.t1", SP_F64, 12, &t1));
  NQ_TRY(alloc_spinor(ctx, "cg.tmp", SP_F64, 12, &tmp));
  Timer T(ctx->stream);
  double b2 = 0;
  NQ_TRY(reduce_norm2(ctx, b, &b2));
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgState s0 = {};
  s0.b2 = b2;
  s0.target = residue * b2;
  s0.niter = niter;
  NQ_TRY(put_state(ctx, 0, s0));
  CgState* st = ctx->state;
  SpinorView none;
  none.ncomp = 12;
  if (guess) {
    NQ_TRY(launch_convert(ctx, x, *guess));
    NQ_TRY(residual_kernsq(S, mu, r, none, x, b, t1, tmp, st, FIN_RR_INIT, 0));
  } else {
    NQ_TRY(launch_zero(ctx, x));
    NQ_TRY(launch_lincomb(ctx, r, 1.0, b, 0, none, 0, none, FIN_RR_INIT, st));
  }
  NQ_TRY(launch_convert(ctx, p, r));
  NQ_TRY(get_states(ctx, 0, 1));
  const double rr0 = ctx->state_host[0].rr;
  int chunk = 8, launched = 0;
  while (!ctx->state_host[0].done) {
    for (int i = 0; i < chunk; i++) {
      HopEpi a = epi_inv_dee(S.kappa, -mu);
      a.st = st;
      NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmp, p, S.U, S.anis, a));
      HopEpi b1 = epi_kout(S.kappa, -mu, p);
      b1.mode = 1;
      b1.st = st;
      b1.fin = FIN_PAP;
      NQ_TRY(launch_tm_hop(ctx, NQ_ODD, t1, tmp, S.U, S.anis, b1));
      HopEpi a2 = epi_inv_dee(S.kappa, mu);
      a2.st = st;
      NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmp, t1, S.U, S.anis, a2));
      HopEpi b2 = epi_kout(S.kappa, mu, t1);
      b2.mode = 2;
      b2.st = st;
      b2.fin = FIN_RR;
      b2.x = x;
      b2.p = p;
      b2.r = r;
      NQ_TRY(launch_tm_hop(ctx, NQ_ODD, none, tmp, S.U, S.anis, b2));
      NQ_TRY(launch_cg_pupdate(ctx, p, r, st));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, 1));
    const CgState& h = ctx->state_host[0];
    chunk = next_chunk(rr0, h.rr, h.target, h.iter, chunk);
    if (launched > niter + 64) break;
  }
  const CgState h = ctx->state_host[0];
  if (h.nan) return fail(NQ_ENAN, "nanned");
  // last calculation of the residual (cg_invert_template_threaded.hpp:131-135)
  CgState sc = {};
  NQ_TRY(put_state(ctx, ctx->nstate - 1, sc));
  NQ_TRY(residual_kernsq(S, mu, r, none, x, b, t1, tmp, scratch_state(ctx), FIN_STORE0, 0));
  NQ_TRY(get_states(ctx, ctx->nstate - 1, 1));
  const double true_rr = ctx->state_host[ctx->nstate - 1].extra[0];
  out->iters = h.iter;
  out->reliable_updates = 0;
  out->rel_res_sq = true_rr / b2;
  out->converged = h.iter < niter;
  out->seconds = T.stop();
  if (!out->converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations)", h.iter);
  return NQ_OK;
}

// Mixed precision CG with reliable updates: sloppy Krylov iteration (FP32 or FP16
// spinors, FP32/FP16 links), FP64 solution accumulation and true-residual
// recomputation whenever |r| drops by delta (or the sloppy residual reaches the
// target); convergence is declared only on the FP64 residual.
// kx as in cg_kernsq_f64: every FP64 residual computes K(-mu) x64 there, so after
// convergence it holds that of the final x64
static int cg_kernsq_mixed(const TmSetup& S64, const TmSetup& Ss, int sp, double mu, int niter, double residue,
                           double delta, const SpinorView& x64, const SpinorView* guess, const SpinorView& b,
                           nq_solve_stats* out, const SpinorView* kx = nullptr)
{
  nq_ctx* ctx = S64.ctx;
  SpinorView r64, t64, tmp64, xs, rs, p, pb, t1s, tmps;
  NQ_TRY(alloc_spinor(ctx, "cg.r", SP_F64, 12, &r64));
  if (kx) t64 = *kx;
  else NQ_TRY(alloc_spinor(ctx, "cg.t1", SP_F64, 12, &t64));
  NQ_TRY(alloc_spinor(ctx, "cg.tmp", SP_F64, 12, &tmp64));
  // sloppy solution increment and residual stay FP32 (FP16 storage of either,
  // with per-site norms, measured 64 -> 70 iterations at 32^3x64 and a slower solve)
  NQ_TRY(alloc_spinor(ctx, "mcg.xs", SP_F32, 12, &xs));
  NQ_TRY(alloc_spinor(ctx, "mcg.rs", SP_F32, 12, &rs));
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "mcg.p16" : "mcg.p32", sp, 12, &p));
  // the search direction alternates between two buffers, so that the p update can
  // apply two iterations' x terms at once (iteration j reads p[j % 2], writes p[(j+1) % 2])
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "mcg.p16b" : "mcg.p32b", sp, 12, &pb));
  const SpinorView pv[2] = {p, pb};
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "mcg.t16" : "mcg.t32", sp, 12, &t1s));
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "mcg.e16" : "mcg.e32", sp, 12, &tmps));
  Timer T(ctx->stream);
  double b2 = 0;
  NQ_TRY(reduce_norm2(ctx, b, &b2));
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgState s0 = {};
  s0.b2 = b2;
  s0.target = residue * b2;
  s0.niter = niter;
  s0.mixed = 1;
  s0.delta = delta;
  NQ_TRY(put_state(ctx, 0, s0));
  CgState* st = ctx->state;
  SpinorView none;
  none.ncomp = 12;
  if (guess) {
    NQ_TRY(launch_convert(ctx, x64, *guess));
    NQ_TRY(residual_kernsq(S64, mu, r64, none, x64, b, t64, tmp64, st, FIN_RR_INIT, 0));
  } else {
    NQ_TRY(launch_zero(ctx, x64));
    NQ_TRY(launch_lincomb(ctx, r64, 1.0, b, 0, none, 0, none, FIN_RR_INIT, st));
  }
  NQ_TRY(launch_convert(ctx, rs, r64));
  NQ_TRY(launch_convert(ctx, p, rs));
  NQ_TRY(launch_zero(ctx, xs));
  NQ_TRY(get_states(ctx, 0, 1));
  const double rr_init = ctx->state_host[0].rr;
  // The sloppy iterations idle (device flag) once a reliable update is due; the host
  // polls after chunks sized from the observed convergence rate to end where the
  // next update is expected, and then runs the FP64 update unconditionally.
  int chunk = 8, launched = 0;
  while (!ctx->state_host[0].done) {
    // iteration j of the solve uses p[j % 2]; the chunk starts at the device's count
    const int par = ctx->state_host[0].iter & 1;
    for (int i = 0; i < chunk; i++) {
      const SpinorView& p = pv[(par + i) & 1];
      const SpinorView& pn = pv[(par + i + 1) & 1];
      HopEpi a = epi_inv_dee(Ss.kappa, -mu);
      a.st = st;
      a.pred_reliable = 2;
      NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmps, p, Ss.U, Ss.anis, a));
      // <p, K K^+ p> = |K(-mu) p|^2 reduced here, finalised (alpha) by an extra
      // workgroup of the next hop, which does not need alpha
      HopEpi b1 = epi_kout(Ss.kappa, -mu, p);
      b1.mode = 1;
      b1.st = st;
      b1.fin = FIN_PARTIAL;
      b1.pred_reliable = 2;
      NQ_TRY(launch_tm_hop(ctx, NQ_ODD, t1s, tmps, Ss.U, Ss.anis, b1));
      HopEpi a2 = epi_inv_dee(Ss.kappa, mu);
      a2.st = st;
      a2.pred_reliable = 2;
      a2.fold_fin = FIN_PAP;
      a2.fold_pred = 2;
      NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmps, t1s, Ss.U, Ss.anis, a2));
      HopEpi b2 = epi_kout(Ss.kappa, mu, t1s);
      b2.mode = 2;
      b2.st = st;
      b2.fin = FIN_RR_SLOPPY;
      b2.pred_reliable = 2;
      b2.p = p;
      b2.r = rs;
      NQ_TRY(launch_tm_hop(ctx, NQ_ODD, none, tmps, Ss.U, Ss.anis, b2));
      // p_next = rs + beta p and, every other iteration, xs += the two iterations'
      // alpha p terms
      NQ_TRY(launch_cg_pupdate_pair(ctx, p, pn, rs, st, xs));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, 1));
    if (ctx->state_host[0].reliable && !ctx->state_host[0].done) {
      // reliable update after iteration j = iter - 1, whose p update was skipped:
      // x64 += xs + (the previous iteration's alpha p when still owed) + alpha p, xs = 0;
      // r64 = b - A x64 in FP64; rs = r64; p[(j+1) % 2] = rs + beta p[j % 2]
      const int j = ctx->state_host[0].iter - 1;
      NQ_TRY(launch_reliable_accum(ctx, x64, xs, st, 1, pv[j & 1], pv[(j + 1) & 1]));
      NQ_TRY(residual_kernsq(S64, mu, r64, rs, x64, b, t64, tmp64, st, FIN_RR_TRUE, 1));
      NQ_TRY(launch_cg_pupdate_pair(ctx, pv[j & 1], pv[(j + 1) & 1], rs, st));
      NQ_TRY(get_states(ctx, 0, 1));
    }
    const CgState& h = ctx->state_host[0];
    if (h.nan || h.done) break;
    // iterations until |r| is expected to fall below delta * r0 (or the target)
    const double goal = std::max(h.delta * h.delta * h.r0 * h.r0, h.target);
    const double slope = h.iter > 0 && h.rr < rr_init ? std::log(h.rr / rr_init) / h.iter : -0.3;
    int need = (int)std::ceil(std::log(goal / h.rr) / slope);
    chunk = std::max(1, std::min(32, need));
    if (launched > niter + 64) break;
  }
  const CgState h = ctx->state_host[0];
  if (h.nan) return fail(NQ_ENAN, "nanned");
  out->iters = h.iter;
  out->reliable_updates = h.nreliable;
  out->rel_res_sq = h.true_rr / b2;
  out->converged = h.true_rr < h.target;
  out->seconds = T.stop();
  if (!out->converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations, rel res %g)", h.iter, h.true_rr / b2);
  return NQ_OK;
}

static int cg_kernsq_any(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu, int niter,
                         double residue, int prec, const SpinorView& x, const SpinorView* guess, const SpinorView& b,
                         nq_solve_stats* st, const SpinorView* kx = nullptr)
{
  int sp, lk;
  NQ_TRY(lk_of_prec(prec, &sp, &lk));
  if (niter <= 0) return fail(NQ_EINVAL, "niter must be positive");
  if (!(residue > 0)) return fail(NQ_EINVAL, "residue must be positive");
  TmSetup S64;
  NQ_TRY(tm_setup(ctx, g, lk64_for(prec, g), anis, kappa, &S64));
  nq_solve_stats dummy;
  if (!st) st = &dummy;
  std::memset(st, 0, sizeof *st);
  if (prec == NQ_PREC_F64) return cg_kernsq_f64(S64, mu, niter, residue, x, guess, b, st, kx);
  TmSetup Ss;
  NQ_TRY(tm_setup(ctx, g, lk, anis, kappa, &Ss));
  return cg_kernsq_mixed(S64, Ss, sp, mu, niter, residue, reliable_delta(sp, residue), x, guess, b, st, kx);
}

static int apply_Q_views(const TmSetup& S, double mu, const nq_field* out, const nq_field* in)
{
  for (int p = 0; p < 2; p++)
    NQ_TRY(launch_tm_hop(S.ctx, p, view_of(out, p), view_of(in, !p), S.U, S.anis, epi_q(S.kappa, mu, view_of(in, p))));
  return NQ_OK;
}

// ---------------------------------------------------------------- batched solves
// Coefficient tables of the four K K^+ epilogues (and the residual one) for nr
// right-hand sides with per-RHS mu; one table is double[16][12] (c1, c2, c3).
enum { TB_INVDEE_M = 0, TB_KOUT_M = 1, TB_INVDEE_P = 2, TB_KOUT_P = 3, TB_RES = 4, TB_N = 5 };

static void put_coef(double* t, const HopEpi& e)
{
  for (int h = 0; h < 2; h++)
    for (int k = 0; k < 2; k++) {
      t[2 * h + k] = e.c1[h][k];
      t[4 + 2 * h + k] = e.c2[h][k];
      t[8 + 2 * h + k] = e.c3[h][k];
    }
}

static int upload_kernsq_coefs(nq_ctx* ctx, double kappa, int nr, const double* mus, double** dev)
{
  NQ_TRY(ctx_temp(ctx, "bcg.coef", sizeof(double) * TB_N * 16 * 12, (void**)dev));
  std::vector<double> h(TB_N * 16 * 12, 0.0);
  SpinorView none;
  for (int r = 0; r < nr; r++) {
    const double mu = mus[r];
    put_coef(&h[(TB_INVDEE_M * 16 + r) * 12], epi_inv_dee(kappa, -mu));
    put_coef(&h[(TB_KOUT_M * 16 + r) * 12], epi_kout(kappa, -mu, none));
    put_coef(&h[(TB_INVDEE_P * 16 + r) * 12], epi_inv_dee(kappa, mu));
    put_coef(&h[(TB_KOUT_P * 16 + r) * 12], epi_kout(kappa, mu, none));
    HopEpi e = epi_kout(kappa, mu, none);
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 2; k++) { e.c1[a][k] = -e.c1[a][k]; e.c2[a][k] = -e.c2[a][k]; }
    setc(e.c3, 1.0, 1.0);
    put_coef(&h[(TB_RES * 16 + r) * 12], e);
  }
  NQ_HIP(hipMemcpyAsync(*dev, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  return NQ_OK;
}

static HopEpi bepi(const double* coef, int table, int nr, CgState* st, int pred)
{
  HopEpi e;
  e.coef = coef + table * 16 * 12;
  e.nrhs = nr;
  e.batched = 1;
  e.st = st;
  e.pred_reliable = pred;
  return e;
}

// r = b - K(+mu_r)K(-mu_r) x for nr right-hand sides (mode-3 tail; rs optional)
static int residual_kernsq_batched(const TmSetup& S, const double* coef, int nr, const SpinorView& r,
                                   const SpinorView& rs, const SpinorView& x, const SpinorView& b, const SpinorView& t1,
                                   const SpinorView& tmp, CgState* st, int fin, int pred)
{
  nq_ctx* ctx = S.ctx;
  NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmp, x, S.U, S.anis, bepi(coef, TB_INVDEE_M, nr, st, pred)));
  HopEpi k1 = bepi(coef, TB_KOUT_M, nr, st, pred);
  k1.aux = x;
  NQ_TRY(launch_tm_hop(ctx, NQ_ODD, t1, tmp, S.U, S.anis, k1));
  NQ_TRY(launch_tm_hop(ctx, NQ_EVEN, tmp, t1, S.U, S.anis, bepi(coef, TB_INVDEE_P, nr, st, pred)));
  HopEpi e = bepi(coef, TB_RES, nr, st, pred);
  e.aux = t1;
  e.aux2 = b;
  e.mode = 3;
  e.r = r;
  e.rs = rs;
  e.fin = fin;
  SpinorView none;
  none.ncomp = 12;
  return launch_tm_hop(ctx, NQ_ODD, none, tmp, S.U, S.anis, e);
}

// ---------------------------------------------------------------- twisted clover operator
// (the solvers follow in the twisted clover section below)
static SpinorView clov_view(const nq_clover* cl, int par)
{
  SpinorView v;
  v.data = cl->cl[par].p;
  v.prec = SP_F64;
  v.ncomp = 36;
  return v;
}

// inverse twisted clover blocks of parity par for (kappa, mu), computed on first use
static int clov_inv_view(nq_clover* cl, int par, double kappa, double mu, SpinorView* v)
{
  nq_ctx* c = cl->ctx;
  if (!cl->inv[par].p) NQ_TRY(dev_alloc(cl->inv[par], (size_t)72 * 16 * c->geo.stride));
  v->data = cl->inv[par].p;
  v->prec = SP_F64;
  v->ncomp = 72;
  if (!cl->inv_valid[par] || cl->inv_kappa[par] != kappa || cl->inv_mu[par] != mu) {
    NQ_TRY(launch_clover_invert(c, clov_view(cl, par), *v, kappa, mu));
    cl->inv_valid[par] = 1;
    cl->inv_kappa[par] = kappa;
    cl->inv_mu[par] = mu;
    cl->invs_valid[0][par] = cl->invs_valid[1][par] = 0;
  }
  return NQ_OK;
}

// packed copies in the sloppy spinor precision sp (FP32 or FP16) of the clover blocks
// and of the inverse blocks (for (kappa, mu))
static int clov_view_s(nq_clover* cl, int par, int sp, SpinorView* v)
{
  nq_ctx* c = cl->ctx;
  const int i = sp == SP_F16 ? 1 : 0;
  if (!cl->cls[i][par].p) NQ_TRY(dev_alloc(cl->cls[i][par], (size_t)36 * (i ? 4 : 8) * c->geo.stride));
  v->data = cl->cls[i][par].p;
  v->prec = i ? SP_F16 : SP_F32;
  v->ncomp = 36;
  if (!cl->cls_valid[i][par]) {
    NQ_TRY(launch_blocks_sloppy(c, 36, clov_view(cl, par), *v));
    cl->cls_valid[i][par] = 1;
  }
  return NQ_OK;
}
static int clov_inv_view_s(nq_clover* cl, int par, double kappa, double mu, int sp, SpinorView* v)
{
  nq_ctx* c = cl->ctx;
  const int i = sp == SP_F16 ? 1 : 0;
  SpinorView v64;
  NQ_TRY(clov_inv_view(cl, par, kappa, mu, &v64));
  if (!cl->invs[i][par].p) NQ_TRY(dev_alloc(cl->invs[i][par], (size_t)72 * (i ? 4 : 8) * c->geo.stride));
  v->data = cl->invs[i][par].p;
  v->prec = i ? SP_F16 : SP_F32;
  v->ncomp = 72;
  if (!cl->invs_valid[i][par]) {
    NQ_TRY(launch_blocks_sloppy(c, 72, v64, *v));
    cl->invs_valid[i][par] = 1;
  }
  return NQ_OK;
}

struct ClovSetup {
  TmSetup S;
  nq_clover* cl;
  SpinorView inv_e;  // inverse blocks of the even sites for (kappa, mu)
  SpinorView cl_o;   // clover blocks of the odd sites
  double mu;
  SpinorView te2, uo;  // scratch: the even intermediate of K; odd (the FP64 check)
};

static int clov_setup(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                      ClovSetup* s)
{
  if (!cl || cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &s->S));
  s->cl = cl;
  s->mu = mu;
  NQ_TRY(clov_inv_view(cl, NQ_EVEN, kappa, mu, &s->inv_e));
  s->cl_o = clov_view(cl, NQ_ODD);
  NQ_TRY(alloc_spinor(c, "clov.te2", SP_F64, 12, &s->te2));
  NQ_TRY(alloc_spinor(c, "clov.uo", SP_F64, 12, &s->uo));
  return NQ_OK;
}

// the same operator in a sloppy precision: FP32 (FP32 18-real links, FP32 spinors and
// blocks) or FP16 (FP16 spinors with per-site norms, FP32 or, for NQ_PREC_F16H, FP16
// 12-real links, FP16 blocks)
static int clov_setup_sloppy(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                             double mu, int sp, int lk, ClovSetup* s)
{
  NQ_TRY(tm_setup(c, g, lk, anis, kappa, &s->S));
  s->cl = cl;
  s->mu = mu;
  NQ_TRY(clov_inv_view_s(cl, NQ_EVEN, kappa, mu, sp, &s->inv_e));
  NQ_TRY(clov_view_s(cl, NQ_ODD, sp, &s->cl_o));
  const bool h = sp == SP_F16;
  NQ_TRY(alloc_spinor(c, h ? "clov.te2_16" : "clov.te2_32", sp, 12, &s->te2));
  return NQ_OK;
}

static const double kOne[2][2] = {{1, 0}, {1, 0}};

// K(dag): out_odd from in_odd; `last` carries the epilogue mode (0 write, 1 norm,
// 2 CG tail) of the second hop, whose c1/c2/aux are set here
// Two hop launches: the first applies the inverse blocks of the even sites to the
// hop in its epilogue (te2 = D_ee(dag)^-1 H in), the second the D_oo(dag) blocks to
// `in` read at its own site (out = g5 (D_oo in - 1/4 H te2)), so neither the hop nor
// the D_oo product goes through memory.
// The first hop idles when the last one does (same state and predicate); fold_fin:
// the finalisation that the first hop carries for the reduction launched before
static int clov_kern(const ClovSetup& s, int dag, const SpinorView& out, const SpinorView& in, HopEpi last,
                     int fold_fin = FIN_NONE)
{
  nq_ctx* c = s.S.ctx;
  HopEpi h;
  h.st = last.st;
  h.pred_reliable = last.pred_reliable;
  h.fold_fin = fold_fin;
  h.fold_pred = last.pred_reliable;
  h.blk = s.inv_e;
  h.blk_mode = 1;
  h.blk_dag = dag;
  h.blk_kappa = s.S.kappa;
  h.blk_mu = s.mu;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, s.te2, in, s.S.U, s.S.anis, h));
  setc(last.c1, -0.25, 0.25);
  setc(last.c2, 1.0, -1.0);
  last.aux = in;
  last.blk = s.cl_o;
  last.blk_mode = 2;
  last.blk_dag = dag;
  last.blk_kappa = s.S.kappa;
  last.blk_mu = s.mu;
  return launch_tm_hop(c, NQ_ODD, out, s.te2, s.S.U, s.S.anis, last);
}

// ---------------------------------------------------------------- TmBatch
int64_t TmBatch::bytes_per_slot(const Geo& g, int prec, bool clover)
{
  const int64_t s = g.stride;
  const int64_t sloppy = prec == NQ_PREC_F32 ? 12 * 8 * s : 12 * 4 * s + 4 * s;
  int64_t b = 5 * 12 * 16 * s;  // b, x64, r64, t64, tmp64
  if (prec != NQ_PREC_F64) {
    b += 2 * 12 * 8 * s;  // xs, rs
    b += 3 * sloppy;      // p, t1, tmp
  } else {
    b += 12 * 16 * s;  // p
  }
  if (clover) b += 2 * 12 * 16 * s + (prec != NQ_PREC_F64 ? 2 * sloppy : 0);  // te2, uo (FP64 and sloppy)
  return b;
}

int TmBatch::setup(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, int prec, int group,
                   nq_clover* cl, double clover_mu)
{
  if (group < 0 || group >= kMaxGroups) return fail(NQ_EINVAL, "batch group %d out of range", group);
  NQ_TRY(lk_of_prec(prec, &sp_, &lk_));
  ctx_ = ctx;
  g_ = g;
  kappa_ = kappa;
  has_anis_ = anis != nullptr;
  anis_ = anis ? *anis : 1.0;
  prec_ = prec;
  group_ = group;
  mixed_ = prec != NQ_PREC_F64;
  S_ = ctx->state + 16 * group;
  TmSetup chk;
  NQ_TRY(tm_setup(ctx, g, cl ? LK_F64_18 : lk64_for(prec, g), anis, kappa, &chk));  // validates kappa / builds links
  if (mixed_) NQ_TRY(tm_setup(ctx, g, lk_, anis, kappa, &chk));
  const std::string t = "tb" + std::to_string(group) + ".";
  const int n = kMaxSlots;
  NQ_TRY(ctx_temp(ctx, t + "coef", sizeof(double) * TB_N * 16 * 12, (void**)&coef_));
  hcoef_.assign(TB_N * 16 * 12, 0.0);
  NQ_TRY(alloc_spinor_nr(ctx, t + "b", SP_F64, 12, n, &b_));
  NQ_TRY(alloc_spinor_nr(ctx, t + "x", SP_F64, 12, n, &x64_));
  NQ_TRY(alloc_spinor_nr(ctx, t + "r", SP_F64, 12, n, &r64_));
  NQ_TRY(alloc_spinor_nr(ctx, t + "t1", SP_F64, 12, n, &t64_));
  NQ_TRY(alloc_spinor_nr(ctx, t + "tmp", SP_F64, 12, n, &tmp64_));
  NQ_TRY(alloc_spinor(ctx, t + "et", SP_F64, 12, &et_));
  if (mixed_) {
    NQ_TRY(alloc_spinor_nr(ctx, t + "xs", SP_F32, 12, n, &xs_));
    NQ_TRY(alloc_spinor_nr(ctx, t + "rs", SP_F32, 12, n, &rs_));
    NQ_TRY(alloc_spinor_nr(ctx, t + "p" + std::to_string(sp_), sp_, 12, n, &p_));
    NQ_TRY(alloc_spinor_nr(ctx, t + "t" + std::to_string(sp_), sp_, 12, n, &t1s_));
    NQ_TRY(alloc_spinor_nr(ctx, t + "e" + std::to_string(sp_), sp_, 12, n, &tmps_));
  } else {
    NQ_TRY(alloc_spinor_nr(ctx, t + "p", SP_F64, 12, n, &p_));
    xs_ = x64_;
    rs_ = r64_;
    t1s_ = t64_;
    tmps_ = tmp64_;
  }
  cl_ = cl;
  clmu_ = clover_mu;
  if (cl) {
    // the group's inverse blocks of the even sites for (kappa, mu) (the clover handle
    // caches one pair only, and groups of other masses may share the handle)
    if (cl->ctx != ctx) return fail(NQ_EINVAL, "clover handle belongs to another context");
    void* p = nullptr;
    NQ_TRY(ctx_temp(ctx, t + "clinv", (size_t)72 * 16 * ctx->geo.stride, &p));
    clinv64_.data = p;
    clinv64_.prec = SP_F64;
    clinv64_.ncomp = 72;
    NQ_TRY(launch_clover_invert(ctx, clov_view(cl, NQ_EVEN), clinv64_, kappa, clover_mu));
    clo64_ = clov_view(cl, NQ_ODD);
    NQ_TRY(alloc_spinor_nr(ctx, t + "cte2", SP_F64, 12, n, &te2_64_));
    NQ_TRY(alloc_spinor_nr(ctx, t + "cuo", SP_F64, 12, n, &uo64_));
    NQ_TRY(alloc_spinor(ctx, t + "cte1", SP_F64, 12, &cte2_));
    if (mixed_) {
      // blocks in the sloppy spinor precision (packed FP32 or FP16)
      const bool h16 = sp_ == SP_F16;
      NQ_TRY(ctx_temp(ctx, t + (h16 ? "clinv16" : "clinv32"), (size_t)72 * (h16 ? 4 : 8) * ctx->geo.stride, &p));
      clinv_s_.data = p;
      clinv_s_.prec = h16 ? SP_F16 : SP_F32;
      clinv_s_.ncomp = 72;
      NQ_TRY(launch_blocks_sloppy(ctx, 72, clinv64_, clinv_s_));
      NQ_TRY(clov_view_s(cl, NQ_ODD, sp_, &clo_s_));
      NQ_TRY(alloc_spinor_nr(ctx, t + "cte2_" + std::to_string(sp_), sp_, 12, n, &te2s_));
      NQ_TRY(alloc_spinor_nr(ctx, t + "cuo_" + std::to_string(sp_), sp_, 12, n, &uos_));
    } else {
      clinv_s_ = clinv64_;
      clo_s_ = clo64_;
      te2s_ = te2_64_;
      uos_ = uo64_;
    }
  }
  // every slot starts idle
  for (int i = 0; i < n; i++) {
    CgState s = {};
    s.done = 1;
    ctx->state_host[16 * group + i] = s;
    active_[i] = false;
  }
  NQ_HIP(hipMemcpyAsync(S_, ctx->state_host + 16 * group, sizeof(CgState) * n, hipMemcpyHostToDevice, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  chunk_ = 8;
  return NQ_OK;
}

int TmBatch::nactive() const
{
  int n = 0;
  for (int i = 0; i < kMaxSlots; i++) n += active_[i];
  return n;
}

int TmBatch::hi() const
{
  int h = 0;
  for (int i = 0; i < kMaxSlots; i++)
    if (active_[i]) h = i + 1;
  return h;
}

int TmBatch::upload_row(int slot, double mu)
{
  SpinorView none;
  double* h = hcoef_.data();
  if (cl_) {
    // twisted clover: the mass sits in the blocks; the hop of D_ee^-1 H is plain
    // (c1 = 1), the K out hop forms g5 (uo - 1/4 hop), the residual b - g5 (uo - 1/4 hop)
    HopEpi plain, kout, res;
    setc(kout.c1, -0.25, 0.25);
    setc(kout.c2, 1.0, -1.0);
    setc(res.c1, 0.25, -0.25);
    setc(res.c2, -1.0, 1.0);
    setc(res.c3, 1.0, 1.0);
    put_coef(&h[(TB_INVDEE_M * 16 + slot) * 12], plain);
    put_coef(&h[(TB_KOUT_M * 16 + slot) * 12], kout);
    put_coef(&h[(TB_INVDEE_P * 16 + slot) * 12], plain);
    put_coef(&h[(TB_KOUT_P * 16 + slot) * 12], kout);
    put_coef(&h[(TB_RES * 16 + slot) * 12], res);
  } else {
    put_coef(&h[(TB_INVDEE_M * 16 + slot) * 12], epi_inv_dee(kappa_, -mu));
    put_coef(&h[(TB_KOUT_M * 16 + slot) * 12], epi_kout(kappa_, -mu, none));
    put_coef(&h[(TB_INVDEE_P * 16 + slot) * 12], epi_inv_dee(kappa_, mu));
    put_coef(&h[(TB_KOUT_P * 16 + slot) * 12], epi_kout(kappa_, mu, none));
    HopEpi e = epi_kout(kappa_, mu, none);
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 2; k++) { e.c1[a][k] = -e.c1[a][k]; e.c2[a][k] = -e.c2[a][k]; }
    setc(e.c3, 1.0, 1.0);
    put_coef(&h[(TB_RES * 16 + slot) * 12], e);
  }
  for (int t = 0; t < TB_N; t++)
    NQ_HIP(hipMemcpyAsync(coef_ + (t * 16 + slot) * 12, &h[(t * 16 + slot) * 12], sizeof(double) * 12,
                          hipMemcpyHostToDevice, ctx_->stream));
  return NQ_OK;
}

static SpinorView slot_view(const SpinorView& v, int slot)
{
  SpinorView o = v;
  const int eb = v.prec == SP_F64 ? 16 : (v.prec == SP_F32 ? 8 : 4);
  o.data = (char*)v.data + (int64_t)slot * v.rstride * eb;
  if (v.norm) o.norm = v.norm + (int64_t)slot * (v.rstride / v.ncomp);
  return o;
}

int TmBatch::start(int slot, double mu, double residue, int niter, const SpinorView& src_e, const SpinorView& src_o,
                   const SpinorView* guess)
{
  if (slot < 0 || slot >= kMaxSlots) return fail(NQ_EINVAL, "slot %d out of range", slot);
  if (active_[slot]) return fail(NQ_EINVAL, "slot %d is busy", slot);
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  if (cl_ && mu != clmu_) return fail(NQ_EINVAL, "twisted clover batch: mu %g differs from the group's %g", mu, clmu_);
  if (cl_ && guess) return fail(NQ_EINVAL, "twisted clover batch: no guess");
  nq_ctx* c = ctx_;
  TmSetup S;
  NQ_TRY(tm_setup(c, g_, lk64(), has_anis_ ? &anis_ : nullptr, kappa_, &S));
  NQ_TRY(upload_row(slot, mu));
  mu_[slot] = mu;
  const double a = 1 / (2 * kappa_), n = a * a + mu * mu;
  SpinorView none;
  none.ncomp = 12;
  // eq. 8a/8b: b = g5 (eta_o + 1/2 H_oe D_ee^-1 eta_e)
  if (cl_) {
    NQ_TRY(launch_clover_apply(c, 1, et_, src_e, clinv64_, none, kappa_, mu, 0, kOne, nullptr));
  } else {
    double ca[2][2], cb[2][2] = {{0, 0}, {0, 0}};
    setc(ca, cplx(a / n, -mu / n), cplx(a / n, mu / n));
    NQ_TRY(launch_spin_diag(c, et_, ca, src_e, cb, none));
  }
  HopEpi e8;
  setc(e8.c1, 0.5, -0.5);
  setc(e8.c2, 1.0, -1.0);
  e8.aux = src_o;
  const SpinorView bs = slot_view(b_, slot);
  NQ_TRY(launch_tm_hop(c, NQ_ODD, bs, et_, S.U, S.anis, e8));
  // solver state of the slot; FIN_INIT_B2 completes it from |b|^2 on the device
  CgState s0 = {};
  s0.target = residue;
  s0.niter = niter;
  s0.mixed = mixed_;
  s0.delta = reliable_delta(sp_, residue);
  c->state_host[16 * group_ + slot] = s0;
  NQ_HIP(hipMemcpyAsync(S_ + slot, c->state_host + 16 * group_ + slot, sizeof(CgState), hipMemcpyHostToDevice,
                        c->stream));
  NQ_TRY(launch_lincomb(c, slot_view(r64_, slot), 1.0, bs, 0, none, 0, none, FIN_INIT_B2, S_ + slot, 1));
  if (guess) {
    // x = guess, r = b - K K^+ x (|r|^2 and the stop test of the start, FIN_RR_INIT)
    NQ_TRY(launch_convert(c, slot_view(x64_, slot), *guess));
    NQ_TRY(residual_kernsq(S, mu, slot_view(r64_, slot), none, slot_view(x64_, slot), bs, slot_view(t64_, slot), et_,
                           S_ + slot, FIN_RR_INIT, 0));
  } else {
    NQ_TRY(launch_zero(c, slot_view(x64_, slot)));
  }
  if (mixed_) {
    NQ_TRY(launch_convert(c, slot_view(rs_, slot), slot_view(r64_, slot)));
    NQ_TRY(launch_zero(c, slot_view(xs_, slot)));
  }
  NQ_TRY(launch_convert(c, slot_view(p_, slot), slot_view(rs_, slot)));
  active_[slot] = true;
  niter_[slot] = niter;
  launched_[slot] = 0;
  return NQ_OK;
}

int TmBatch::lk64() const { return cl_ ? LK_F64_18 : lk64_for(prec_, g_); }

// K(dag) in = g5 (D_oo(dag) in - 1/4 H D_ee(dag)^-1 H in) on nr slots: a plain hop, the
// batched inverse blocks, the batched D_oo blocks of `in`, and the K out hop whose
// epilogue (last: its coefficient rows, mode and tail) combines them
int TmBatch::clov_kern_batched(const LinkView& U, double anis, int dag, const SpinorView& out, const SpinorView& in,
                               const SpinorView& tmp, const SpinorView& te2, const SpinorView& uo,
                               const SpinorView& inv, const SpinorView& clo, HopEpi last, int nr, int pred)
{
  nq_ctx* c = ctx_;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, tmp, in, U, anis, bepi(coef_, dag ? TB_INVDEE_M : TB_INVDEE_P, nr, last.st, pred)));
  NQ_TRY(launch_clover_apply_batched(c, 1, dag, te2, tmp, inv, kappa_, clmu_, nr, last.st, pred));
  NQ_TRY(launch_clover_apply_batched(c, 0, dag, uo, in, clo, kappa_, clmu_, nr, last.st, pred));
  last.aux = uo;
  return launch_tm_hop(c, NQ_ODD, out, te2, U, anis, last);
}

int TmBatch::enqueue(int nr, int iters)
{
  nq_ctx* c = ctx_;
  TmSetup Ss;
  NQ_TRY(tm_setup(c, g_, mixed_ ? lk_ : LK_F64_18, has_anis_ ? &anis_ : nullptr, kappa_, &Ss));
  const int pred = mixed_ ? 2 : 0;
  SpinorView none;
  none.ncomp = 12;
  if (cl_) {
    // twisted clover: K(1) p (|.|^2 = <p, K K^+ p>), then K(0) with the CG tail
    for (int it = 0; it < iters; it++) {
      HopEpi b1 = bepi(coef_, TB_KOUT_M, nr, S_, pred);
      b1.mode = 1;
      b1.fin = FIN_PAP;
      NQ_TRY(clov_kern_batched(Ss.U, Ss.anis, 1, t1s_, p_, tmps_, te2s_, uos_, clinv_s_, clo_s_, b1, nr, pred));
      HopEpi b2e = bepi(coef_, TB_KOUT_P, nr, S_, pred);
      b2e.mode = 2;
      b2e.fin = mixed_ ? FIN_RR_SLOPPY : FIN_RR;
      b2e.p = p_;
      b2e.r = rs_;
      if (!mixed_) b2e.x = xs_;
      NQ_TRY(clov_kern_batched(Ss.U, Ss.anis, 0, none, t1s_, tmps_, te2s_, uos_, clinv_s_, clo_s_, b2e, nr, pred));
      if (mixed_) NQ_TRY(launch_cg_pupdate(c, p_, rs_, S_, nr, 1, xs_));
      else NQ_TRY(launch_cg_pupdate(c, p_, rs_, S_, nr));
    }
    return NQ_OK;
  }
  for (int it = 0; it < iters; it++) {
    NQ_TRY(launch_tm_hop(c, NQ_EVEN, tmps_, p_, Ss.U, Ss.anis, bepi(coef_, TB_INVDEE_M, nr, S_, pred)));
    HopEpi b1 = bepi(coef_, TB_KOUT_M, nr, S_, pred);
    b1.aux = p_;
    b1.mode = 1;
    b1.fin = FIN_PAP;
    NQ_TRY(launch_tm_hop(c, NQ_ODD, t1s_, tmps_, Ss.U, Ss.anis, b1));
    NQ_TRY(launch_tm_hop(c, NQ_EVEN, tmps_, t1s_, Ss.U, Ss.anis, bepi(coef_, TB_INVDEE_P, nr, S_, pred)));
    HopEpi b2e = bepi(coef_, TB_KOUT_P, nr, S_, pred);
    b2e.aux = t1s_;
    b2e.mode = 2;
    b2e.fin = mixed_ ? FIN_RR_SLOPPY : FIN_RR;
    b2e.p = p_;
    b2e.r = rs_;
    if (mixed_) {
      // sloppy: xs += alpha p moves into the p update (one read of p); a slot whose
      // update is pending skips it and the reliable update folds alpha p in
      NQ_TRY(launch_tm_hop(c, NQ_ODD, none, tmps_, Ss.U, Ss.anis, b2e));
      NQ_TRY(launch_cg_pupdate(c, p_, rs_, S_, nr, 1, xs_));
    } else {
      // FP64: the slot that converges stops before its p update, so x stays in the tail
      b2e.x = xs_;
      NQ_TRY(launch_tm_hop(c, NQ_ODD, none, tmps_, Ss.U, Ss.anis, b2e));
      NQ_TRY(launch_cg_pupdate(c, p_, rs_, S_, nr));
    }
  }
  return NQ_OK;
}

// Replays of captured iteration chunks.  Every launch of an iteration has fixed
// arguments (batch buffers, coefficient rows and states of this batch), so a chunk
// of 2^k iterations for a given number of slots is captured once into a HIP graph
// and replayed; a chunk of any length is a sum of such graphs.  The first chunk of a
// slot count runs eagerly (it also settles the autotuned launch shapes); the launch
// ledger, when enabled, needs eager launches too.  NQ_GRAPHS=0 disables graphs.
int TmBatch::run_graphs(int nr, int iters)
{
  nq_ctx* c = ctx_;
  for (int k = 6; k >= 0; k--) {
    const int n = 1 << k;
    while (iters >= n) {
      const int key = nr * 128 + n;
      auto it = graphs_.find(key);
      if (it == graphs_.end()) {
        hipGraph_t gr = nullptr;
        NQ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        c->capturing = true;
        const int rc = enqueue(nr, n);
        c->capturing = false;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc != NQ_OK) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc;
        }
        if (e != hipSuccess) return fail(NQ_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) return fail(NQ_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
        it = graphs_.emplace(key, ex).first;
      }
      NQ_HIP(hipGraphLaunch(it->second, c->stream));
      iters -= n;
    }
  }
  return NQ_OK;
}

TmBatch::~TmBatch()
{
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
}

int TmBatch::run(int iters)
{
  const int nr = hi();
  if (nr == 0 || iters <= 0) return NQ_OK;
  const char* env = getenv("NQ_GRAPHS");
  const bool graphs = !(env && env[0] == '0') && !ctx_->profiling && eager_done_.count(nr);
  if (graphs) NQ_TRY(run_graphs(nr, iters));
  else NQ_TRY(enqueue(nr, iters));
  eager_done_.insert(nr);
  for (int i = 0; i < nr; i++)
    if (active_[i]) launched_[i] += iters;
  return NQ_OK;
}

int TmBatch::poll(std::vector<int>* finished)
{
  const int nr = hi();
  if (nr == 0) return NQ_OK;
  nq_ctx* c = ctx_;
  const int base = 16 * group_;
  NQ_TRY(get_states(c, base, nr));
  bool any_rel = false;
  for (int i = 0; i < nr; i++) any_rel |= active_[i] && c->state_host[base + i].reliable && !c->state_host[base + i].done;
  if (any_rel) {
    // reliable update of the flagged slots: x64 += xs + alpha p, xs = 0; r64 = b - A x64
    // in FP64, rs = r64; p = rs + beta p; then clear the flags
    TmSetup S64;
    NQ_TRY(tm_setup(c, g_, lk64(), has_anis_ ? &anis_ : nullptr, kappa_, &S64));
    NQ_TRY(launch_reliable_accum(c, x64_, xs_, S_, nr, p_));
    if (cl_) {
      // r64 = b - K(0) K(1) x64 (FP64 blocks and links), rs = r64
      NQ_TRY(clov_kern_batched(S64.U, S64.anis, 1, t64_, x64_, tmp64_, te2_64_, uo64_, clinv64_, clo64_,
                               bepi(coef_, TB_KOUT_M, nr, S_, 1), nr, 1));
      HopEpi e = bepi(coef_, TB_RES, nr, S_, 1);
      e.aux2 = b_;
      e.mode = 3;
      e.r = r64_;
      e.rs = rs_;
      e.fin = FIN_RR_TRUE_HOLD;
      SpinorView none;
      none.ncomp = 12;
      NQ_TRY(clov_kern_batched(S64.U, S64.anis, 0, none, t64_, tmp64_, te2_64_, uo64_, clinv64_, clo64_, e, nr, 1));
    } else {
      NQ_TRY(residual_kernsq_batched(S64, coef_, nr, r64_, rs_, x64_, b_, t64_, tmp64_, S_, FIN_RR_TRUE_HOLD, 1));
    }
    NQ_TRY(launch_cg_pupdate(c, p_, rs_, S_, nr, 2));
    NQ_TRY(launch_clear_reliable(c, S_, nr));
    NQ_TRY(get_states(c, base, nr));
  }
  int next = 64;
  for (int i = 0; i < nr; i++) {
    if (!active_[i]) continue;
    const CgState& h = c->state_host[base + i];
    // a slot ends when its state says so; the launch count bounds the wait (every
    // launched iteration either advances the state or idles a finished slot)
    if (h.done || launched_[i] > 2 * niter_[i] + 64) {
      finished->push_back(i);
      continue;
    }
    const double goal = mixed_ ? std::max(h.delta * h.delta * h.r0 * h.r0, h.target) : h.target;
    next = std::min(next, next_chunk(h.b2, h.rr, goal, h.iter, chunk_));
  }
  chunk_ = std::max(2, next);
  return NQ_OK;
}

int TmBatch::finish(int slot, const SpinorView& src_e, const SpinorView& src_o, const SpinorView& sol_e,
                    const SpinorView& sol_o, nq_solve_stats* st, double* check, const SpinorView* guess_out)
{
  if (slot < 0 || slot >= kMaxSlots || !active_[slot]) return fail(NQ_EINVAL, "slot %d is not active", slot);
  nq_ctx* c = ctx_;
  active_[slot] = false;
  const CgState h = c->state_host[16 * group_ + slot];
  if (!h.done) {
    // stopped by the launch bound: park the slot so that later launches skip it
    CgState& hs = c->state_host[16 * group_ + slot];
    hs.done = 1;
    NQ_HIP(hipMemcpyAsync(S_ + slot, &hs, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
  }
  nq_solve_stats out = {};
  out.iters = h.iter;
  out.reliable_updates = h.nreliable;
  if (h.nan) {
    if (st) *st = out;
    return fail(NQ_ENAN, h.b2 > 0 ? "nanned" : "invalid norm");
  }
  const double mu = mu_[slot];
  TmSetup S;
  NQ_TRY(tm_setup(c, g_, lk64(), has_anis_ ? &anis_ : nullptr, kappa_, &S));
  const double a = 1 / (2 * kappa_), n = a * a + mu * mu;
  const cplx zi_u(a / n, -mu / n), zi_l(a / n, mu / n);
  if (guess_out) NQ_TRY(launch_convert(c, *guess_out, slot_view(x64_, slot)));
  SpinorView none;
  none.ncomp = 12;
  if (cl_) {
    // eq. 10: sol_o = K(1) y ; eq. 11: sol_e = D_ee^-1 (eta_e + 1/2 H_eo sol_o), with the
    // group's inverse blocks (cg_invert_tmclovD_eoprec.cpp:200-225)
    ClovSetup cs;
    cs.S = S;
    cs.cl = cl_;
    cs.inv_e = clinv64_;
    cs.cl_o = clo64_;
    cs.mu = clmu_;
    cs.te2 = cte2_;
    NQ_TRY(clov_kern(cs, 1, sol_o, slot_view(x64_, slot), HopEpi()));
    HopEpi e11;
    setc(e11.c1, 0.5, 0.5);
    setc(e11.c2, 1.0, 1.0);
    e11.aux = src_e;
    NQ_TRY(launch_tm_hop(c, NQ_EVEN, cte2_, sol_o, S.U, S.anis, e11));
    NQ_TRY(launch_clover_apply(c, 1, sol_e, cte2_, clinv64_, none, kappa_, mu, 0, kOne, nullptr));
  } else {
    // eq. 10: sol_o = K(-mu) y ; eq. 11: sol_e = D_ee^-1 (eta_e + 1/2 H_eo sol_o)
    NQ_TRY(apply_kern(S, -mu, sol_o, slot_view(x64_, slot), sol_e));
    HopEpi e11;
    setc(e11.c1, 0.5 * zi_u, 0.5 * zi_l);
    setc(e11.c2, zi_u, zi_l);
    e11.aux = src_e;
    NQ_TRY(launch_tm_hop(c, NQ_EVEN, sol_e, sol_o, S.U, S.anis, e11));
  }
  // FP64 check with the 18-real links: |D sol - src|^2 / |src|^2
  TmSetup Sc;
  NQ_TRY(tm_setup(c, g_, LK_F64_18, has_anis_ ? &anis_ : nullptr, kappa_, &Sc));
  CgState* sc = c->state + 48 + 2 * group_;
  c->state_host[48 + 2 * group_] = CgState{};
  c->state_host[49 + 2 * group_] = CgState{};
  NQ_HIP(hipMemcpyAsync(sc, c->state_host + 48 + 2 * group_, 2 * sizeof(CgState), hipMemcpyHostToDevice, c->stream));
  const SpinorView sol_v[2] = {sol_e, sol_o}, src_v[2] = {src_e, src_o};
  for (int p = 0; p < 2; p++) {
    HopEpi ec;
    setc(ec.c1, -0.5, -0.5);
    setc(ec.c3, -1.0, -1.0);
    if (cl_) {
      // (z + Cl) sol_p - 1/2 hop sol_!p - src_p
      NQ_TRY(launch_clover_apply(c, 0, cte2_, sol_v[p], clov_view(cl_, p), none, kappa_, mu, 0, kOne, nullptr));
      setc(ec.c2, 1.0, 1.0);
      ec.aux = cte2_;
    } else {
      setc(ec.c2, cplx(a, mu), cplx(a, -mu));
      ec.aux = sol_v[p];
    }
    ec.aux2 = src_v[p];
    ec.mode = 1;
    ec.write = 0;
    ec.st = sc;
    ec.fin = p == 0 ? FIN_STORE0 : FIN_STORE1;
    NQ_TRY(launch_tm_hop(c, p, none, sol_v[!p], Sc.U, Sc.anis, ec));
  }
  CgState* sn = c->state + 48 + 2 * group_ + 1;
  NQ_TRY(launch_lincomb(c, none, 1.0, src_e, 0, none, 0, none, FIN_STORE0, sn));
  NQ_TRY(launch_lincomb(c, none, 1.0, src_o, 0, none, 0, none, FIN_STORE1, sn + 0));
  NQ_TRY(get_states(c, 48 + 2 * group_, 2));
  // state 48+2g: extra[0] = |.|^2 even, extra[1] = |.|^2 odd; state 49+2g: |src_e|^2, |src_o|^2
  const CgState& r0 = c->state_host[48 + 2 * group_];
  const CgState& r1 = c->state_host[49 + 2 * group_];
  const double chk = (r0.extra[0] + r0.extra[1]) / (r1.extra[0] + r1.extra[1]);
  if (check) *check = chk;
  out.rel_res_sq = chk;
  out.converged = mixed_ ? h.true_rr < h.target : h.iter < h.niter;
  if (st) *st = out;
  if (!out.converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations)", h.iter);
  return NQ_OK;
}

// CG on Q(-mu) Q(+mu) over lx (inv_tmQ2_cg), FP64, cg_invert semantics
// Q(mu) on lx fields: writes out = Q(mu) in
typedef std::function<int(double mu, const nq_field* out, const nq_field* in)> ApplyQ;

// cg_invert on Q(-mu)Q(+mu) over lx fields (inv_tmQ2_cg / inv_tmclovQ2_cg): the matvec
// is two Q applies and <p, Q(-mu)Q(+mu) p> = |Q(+mu) p|^2 (Q(mu)^+ = Q(-mu)).  FP64.
static int q2_cg_f64(nq_ctx* ctx, const ApplyQ& Q, double mu, int niter, double residue, const nq_field* sol,
                     const nq_field* guess, const nq_field* src, nq_solve_stats* out)
{
  Timer T(ctx->stream);
  // lx temporaries as fields (even half then odd half) so parity views apply
  SpinorView rv, pv, tv, sv;
  NQ_TRY(alloc_spinor_nr(ctx, "q2.r", SP_F64, 12, 2, &rv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.p", SP_F64, 12, 2, &pv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.t", SP_F64, 12, 2, &tv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.s", SP_F64, 12, 2, &sv));
  auto as_field = [&](const SpinorView& v) {
    nq_field f;
    f.ctx = ctx;
    f.kind = NQ_SPINCOLOR;
    f.coverage = NQ_LX;
    f.buf.p = v.data;
    return f;
  };
  nq_field fr = as_field(rv), fp = as_field(pv), ft = as_field(tv), fs = as_field(sv);
  struct Release {
    nq_field* f[4];
    ~Release()
    {
      for (auto* x : f) x->buf.p = nullptr;
    }
  } rel{{&fr, &fp, &ft, &fs}};
  const SpinorView x = view_lx_flat(sol), b = view_lx_flat(src), r = view_lx_flat(&fr), p = view_lx_flat(&fp),
                   s = view_lx_flat(&fs);
  SpinorView none;
  none.ncomp = 24;
  CgState* st = ctx->state;
  NQ_TRY(put_state(ctx, 0, CgState{}));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, 0, none, 0, none, FIN_STORE0, st));
  NQ_TRY(get_states(ctx, 0, 1));
  const double b2 = ctx->state_host[0].extra[0];
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgState s0 = {};
  s0.b2 = b2;
  s0.target = residue * b2;
  s0.niter = niter;
  NQ_TRY(put_state(ctx, 0, s0));
  auto q2 = [&](const nq_field* out_f, const nq_field* in_f) -> int {
    NQ_TRY(Q(+mu, &ft, in_f));
    return Q(-mu, out_f, &ft);
  };
  if (guess) {
    if (guess != sol) NQ_TRY(launch_convert(ctx, x, view_lx_flat(guess)));
    NQ_TRY(q2(&fs, sol));
    NQ_TRY(launch_lincomb(ctx, r, 1.0, b, -1.0, s, 0, none, FIN_RR_INIT, st));
  } else {
    NQ_TRY(launch_zero(ctx, x));
    NQ_TRY(launch_lincomb(ctx, r, 1.0, b, 0, none, 0, none, FIN_RR_INIT, st));
  }
  NQ_TRY(launch_convert(ctx, p, r));
  NQ_TRY(get_states(ctx, 0, 1));
  const double rr0 = ctx->state_host[0].rr;
  int chunk = 8, launched = 0;
  while (!ctx->state_host[0].done) {
    for (int i = 0; i < chunk; i++) {
      // s = Q(-mu) Q(+mu) p ; <p, s> = |Q(+mu) p|^2
      NQ_TRY(Q(+mu, &ft, &fp));
      NQ_TRY(launch_lincomb(ctx, none, 1.0, view_lx_flat(&ft), 0, none, 0, none, FIN_PAP, st, 1, 1));
      NQ_TRY(Q(-mu, &fs, &ft));
      NQ_TRY(launch_cg_update(ctx, x, p, r, s, st, FIN_RR));
      NQ_TRY(launch_cg_pupdate(ctx, p, r, st));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, 1));
    const CgState& h = ctx->state_host[0];
    chunk = next_chunk(rr0, h.rr, h.target, h.iter, chunk);
    if (launched > niter + 64) break;
  }
  const CgState h = ctx->state_host[0];
  if (h.nan) return fail(NQ_ENAN, "nanned");
  // last calculation of the residual
  NQ_TRY(q2(&fs, sol));
  NQ_TRY(put_state(ctx, 1, CgState{}));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, -1.0, s, 0, none, FIN_STORE0, st + 1));
  NQ_TRY(get_states(ctx, 1, 1));
  out->iters = h.iter;
  out->rel_res_sq = ctx->state_host[1].extra[0] / b2;
  out->converged = h.iter < niter;
  out->seconds = T.stop();
  if (!out->converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations)", h.iter);
  return NQ_OK;
}

// Multi-mass CG on Q(-m_k)Q(+m_k) = Q(0)^2 + m_k^2 over lx (inv_tmQ2_cgm,
// cgm_invert_tmQ2.cpp:46-53 with cgm_invert's recurrences): the matvec is Q(0)^2 (four
// hop launches), <p, Q(0)^2 p> = |Q(0) p|^2, the shifts are m_k^2.  FP64.
// The Q apply comes as a callback (twisted mass or twisted clover).
static int q2_cgm_f64(nq_ctx* ctx, const ApplyQ& Q, int nmass, const double* masses, int niter,
                      const double* residues, nq_field** sols, const nq_field* src, nq_solve_stats* st,
                      double* true_res)
{
  const int64_t fsz = 24 * ctx->geo.stride;  // elements of one lx field
  SpinorView rv, pv, tv, sv, ps, xs;
  NQ_TRY(alloc_spinor_nr(ctx, "q2.r", SP_F64, 12, 2, &rv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.p", SP_F64, 12, 2, &pv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.t", SP_F64, 12, 2, &tv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2.s", SP_F64, 12, 2, &sv));
  NQ_TRY(alloc_spinor_nr(ctx, "q2m.ps", SP_F64, 24, nmass, &ps));
  NQ_TRY(alloc_spinor_nr(ctx, "q2m.x", SP_F64, 24, nmass, &xs));
  auto as_field = [&](void* p) {
    nq_field f;
    f.ctx = ctx;
    f.kind = NQ_SPINCOLOR;
    f.coverage = NQ_LX;
    f.buf.p = p;
    return f;
  };
  nq_field fr = as_field(rv.data), fp = as_field(pv.data), ft = as_field(tv.data), fs = as_field(sv.data);
  struct Release {
    nq_field* f[4];
    ~Release()
    {
      for (auto* x : f) x->buf.p = nullptr;
    }
  } rel{{&fr, &fp, &ft, &fs}};
  const SpinorView b = view_lx_flat(src), r = view_lx_flat(&fr), p = view_lx_flat(&fp), t = view_lx_flat(&ft),
                   s = view_lx_flat(&fs);
  SpinorView none;
  none.ncomp = 24;
  Timer T(ctx->stream);
  NQ_TRY(put_state(ctx, 0, CgState{}));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, 0, none, 0, none, FIN_STORE0, ctx->state));
  NQ_TRY(get_states(ctx, 0, 1));
  const double b2 = ctx->state_host[0].extra[0];
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgmState* dst = nullptr;
  CgmState* hst = nullptr;
  NQ_TRY(ctx_temp(ctx, "q2m.state", sizeof(CgmState), (void**)&dst));
  NQ_HIP(hipHostMalloc((void**)&hst, sizeof(CgmState), hipHostMallocDefault));
  struct HostFree {
    CgmState* p;
    ~HostFree() { (void)hipHostFree(p); }
  } hf{hst};
  CgmState h0 = {};
  h0.b2 = b2;
  h0.rr = b2;
  h0.betaa = 1;
  h0.nshift = nmass;
  h0.nrun = nmass;
  h0.niter = niter;
  for (int k = 0; k < nmass; k++) {
    h0.shift[k] = masses[k] * masses[k];
    h0.target[k] = residues[k];
    h0.zps[k] = h0.zas[k] = 1;
    h0.run[k] = 1;
  }
  *hst = h0;
  NQ_HIP(hipMemcpyAsync(dst, hst, sizeof(CgmState), hipMemcpyHostToDevice, ctx->stream));
  for (int k = 0; k < nmass; k++) {
    SpinorView xk = xs, pk = ps;
    xk.data = (char*)xs.data + k * fsz * 16;
    pk.data = (char*)ps.data + k * fsz * 16;
    NQ_TRY(launch_zero(ctx, xk));
    NQ_TRY(launch_convert(ctx, pk, b));
  }
  NQ_TRY(launch_convert(ctx, p, b));
  NQ_TRY(launch_convert(ctx, r, b));
  int chunk = 8, launched = 0;
  for (;;) {
    for (int it = 0; it < chunk; it++) {
      NQ_TRY(Q(0.0, &ft, &fp));  // t = Q(0) p
      NQ_TRY(launch_lincomb(ctx, none, 1.0, t, 0, none, 0, none, FIN_PARTIAL, nullptr));
      NQ_TRY(launch_cgm_finalize(ctx, dst, 1, 1, 1.0));  // pap = |t|^2
      NQ_TRY(Q(0.0, &fs, &ft));  // s = Q(0)^2 p
      NQ_TRY(launch_cgm_vec2(ctx, s, r, xs, ps, fsz, dst));
      NQ_TRY(launch_cgm_vec3(ctx, r, p, ps, fsz, dst));
    }
    launched += chunk;
    NQ_HIP(hipMemcpyAsync(hst, dst, sizeof(CgmState), hipMemcpyDeviceToHost, ctx->stream));
    NQ_HIP(hipStreamSynchronize(ctx->stream));
    if (hst->nan) return fail(NQ_ENAN, "nanned");
    if (hst->done || launched > niter + 64) break;
    chunk = std::min(64, chunk * 2);
  }
  const int iters = hst->iter;
  // copy out and the FP64 check |(Q(-m)Q(+m)) x_k - b|^2 / |b|^2
  double worst = 0;
  for (int k = 0; k < nmass; k++) {
    SpinorView xk = xs;
    xk.data = (char*)xs.data + k * fsz * 16;
    NQ_TRY(launch_convert(ctx, view_lx_flat(sols[k]), xk));
    NQ_TRY(Q(+masses[k], &ft, sols[k]));
    NQ_TRY(Q(-masses[k], &fs, &ft));
    NQ_TRY(put_state(ctx, 0, CgState{}));
    NQ_TRY(launch_lincomb(ctx, none, 1.0, s, -1.0, b, 0, none, FIN_STORE0, ctx->state));
    NQ_TRY(get_states(ctx, 0, 1));
    const double tr = ctx->state_host[0].extra[0] / b2;
    if (true_res) true_res[k] = tr;
    worst = std::max(worst, tr);
  }
  if (st) {
    *st = nq_solve_stats{};
    st->iters = iters;
    st->rel_res_sq = worst;
    st->converged = iters < niter;
    st->seconds = T.stop();
  }
  if (iters >= niter) return fail(NQ_ENOCONV, "exit without converging");
  return NQ_OK;
}

}  // namespace nq

using namespace nq;

namespace {
struct Guard {
  std::lock_guard<std::recursive_mutex> lk;
  explicit Guard(nq_ctx* c) : lk(c->mu) { (void)hipSetDevice(c->device); }
};
int check_field(const nq_ctx* c, const nq_field* f, int kind, int cov, const char* what)
{
  if (!f) return fail(NQ_EINVAL, "%s: null field", what);
  if (f->ctx != c) return fail(NQ_EINVAL, "%s: field belongs to another context", what);
  if (f->kind != kind || f->coverage != cov) return fail(NQ_EINVAL, "%s: wrong field kind/coverage", what);
  return NQ_OK;
}

int check_multi(const nq_ctx* c, int nrhs, int niter, double residue, nq_field** sols, nq_field** guesses,
                const nq_field* const* srcs, const char* what)
{
  if (nrhs < 1) return fail(NQ_EINVAL, "need at least one source");
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  for (int i = 0; i < nrhs; i++) {
    NQ_TRY(check_field(c, srcs[i], NQ_SPINCOLOR, NQ_LX, what));
    NQ_TRY(check_field(c, sols[i], NQ_SPINCOLOR, NQ_LX, what));
    if (guesses && guesses[i]) NQ_TRY(check_field(c, guesses[i], NQ_SPINCOLOR, NQ_ODD, what));
    for (int j = 0; j < nrhs; j++)
      if (sols[i] == srcs[j]) return fail(NQ_EINVAL, "solution aliases a source");
  }
  return NQ_OK;
}

// Streams nrhs sources through one batch: sources enter free slots as earlier ones
// finish (eq. 8 and the CG start), the batch iterates, finished slots run eqs. 10-11
// and the FP64 check.
int run_batch_solves(nq_ctx* c, TmBatch& B, int nrhs, const double* mus, int niter, double residue,
                     nq_field** sols, nq_field** guesses, const nq_field* const* srcs, nq_solve_stats* st)
{
  std::vector<nq_solve_stats> local(nrhs);
  std::vector<int> owner(TmBatch::kMaxSlots, -1);
  int rc = NQ_OK, next = 0, ndone = 0;
  auto fill = [&]() -> int {
    for (int s = 0; s < TmBatch::kMaxSlots && next < nrhs; s++)
      if (!B.active(s)) {
        SpinorView gv;
        const bool hg = guesses && guesses[next];
        if (hg) gv = view_of(guesses[next], 0);
        NQ_TRY(B.start(s, mus[next], residue, niter, view_of(srcs[next], 0), view_of(srcs[next], 1),
                       hg ? &gv : nullptr));
        owner[s] = next++;
      }
    return NQ_OK;
  };
  NQ_TRY(fill());
  while (ndone < nrhs) {
    NQ_TRY(B.run(B.suggested_chunk()));
    std::vector<int> fin;
    NQ_TRY(B.poll(&fin));
    for (int s : fin) {
      const int i = owner[s];
      SpinorView gv;
      const bool hg = guesses && guesses[i];
      if (hg) gv = view_of(guesses[i], 0);
      const int r = B.finish(s, view_of(srcs[i], 0), view_of(srcs[i], 1), view_of(sols[i], 0), view_of(sols[i], 1),
                             &local[i], nullptr, hg ? &gv : nullptr);
      if (r == NQ_ENOCONV) rc = r;
      else if (r != NQ_OK) return r;
      owner[s] = -1;
      ndone++;
    }
    NQ_TRY(fill());
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  if (st)
    for (int i = 0; i < nrhs; i++) st[i] = local[i];
  return rc == NQ_ENOCONV ? fail(NQ_ENOCONV, "exit without converging") : NQ_OK;
}
}  // namespace

// ---------------------------------------------------------------- allocation-free K K^+
// Scratch of nq_tm_kern_square_noalloc: the even intermediate and the odd K(-mu) in of
// one FP64 K K^+ apply (two spincolor fields of `stride` sites, 384 B per site), in a
// device array that each GPU holds from module load to process end.  The host never
// passes the array's name to a kernel (in host code it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per device with
// that device current, and is kept per device.  Every call on one GPU shares its copy,
// so calls that overlap on one GPU race; nothing in it has to start at zero (both
// fields are written in full before they are read).
static constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// device address of this device's scratch (the caller has made `dev` current)
int noalloc_scratch(int dev, char** out)
{
  if (dev < 0 || dev >= kMaxDevices) return fail(NQ_EINVAL, "device %d out of range for the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NQ_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  *out = (char*)g_noalloc_ptr[dev];
  return NQ_OK;
}
}  // namespace

extern "C" {

int nq_tm_hop(nq_ctx* c, const nq_gauge* g, const double* anis, nq_field* out, const nq_field* in)
{
  if (!c || !out || !in) return fail(NQ_EINVAL, "null argument");
  if (in->coverage == NQ_LX || out->coverage == NQ_LX || in->coverage == out->coverage)
    return fail(NQ_EINVAL, "calling with messed up parities");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, in->coverage, "nq_tm_hop"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, out->coverage, "nq_tm_hop"));
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, 1.0, &S));
  HopEpi e;
  NQ_TRY(launch_tm_hop(c, out->coverage, view_of(out, 0), view_of(in, 0), S.U, S.anis, e));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tm_Dee(nq_ctx* c, double kappa, double mu, int inverse, nq_field* out, const nq_field* in)
{
  if (!c || !out || !in) return fail(NQ_EINVAL, "null argument");
  if (in->coverage == NQ_LX || out->coverage != in->coverage) return fail(NQ_EINVAL, "nq_tm_Dee: parity mismatch");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, in->coverage, "nq_tm_Dee"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, in->coverage, "nq_tm_Dee"));
  Guard G(c);
  const double a = 1 / (2 * kappa), n = a * a + mu * mu;
  double ca[2][2], cb[2][2] = {{0, 0}, {0, 0}};
  if (inverse) setc(ca, cplx(a / n, -mu / n), cplx(a / n, mu / n));
  else setc(ca, cplx(a, mu), cplx(a, -mu));
  SpinorView none;
  NQ_TRY(launch_spin_diag(c, view_of(out, 0), ca, view_of(in, 0), cb, none));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tm_kern(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, nq_field* out,
               const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern"));
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  SpinorView tmp;
  NQ_TRY(alloc_spinor(c, "op.tmp", SP_F64, 12, &tmp));
  NQ_TRY(apply_kern(S, mu, view_of(out, 0), view_of(in, 0), tmp));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tm_kern_square(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, nq_field* out,
                      const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern_square"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern_square"));
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  SpinorView tmp, t1;
  NQ_TRY(alloc_spinor(c, "op.tmp", SP_F64, 12, &tmp));
  NQ_TRY(alloc_spinor(c, "op.t1", SP_F64, 12, &t1));
  NQ_TRY(apply_kern(S, -mu, t1, view_of(in, 0), tmp));
  NQ_TRY(apply_kern(S, +mu, view_of(out, 0), t1, tmp));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

// nq_tm_kern_square with its two temporaries in the library's static device scratch:
// no device or host-pinned allocation, and the launch ledger stays off for it.  The
// gauge handle's FP64 18-real links are built when the handle is created.
int nq_tm_kern_square_noalloc(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu,
                              nq_field* out, const nq_field* in)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern_square_noalloc"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_ODD, "nq_tm_kern_square_noalloc"));
  const size_t field = (size_t)12 * 16 * (size_t)c->geo.stride;
  if (2 * field > kNoAllocScratchBytes)
    return fail(NQ_EINVAL, "nq_tm_kern_square_noalloc: %lld sites per parity need %zu bytes of scratch, the limit is %zu "
                           "(%lld sites per parity)",
                (long long)c->geo.Vh, 2 * field, kNoAllocScratchBytes,
                (long long)(kNoAllocScratchBytes / (2 * 12 * 16) / 64 * 64));
  if (g->ctx != c) return fail(NQ_EINVAL, "nq_tm_kern_square_noalloc: gauge belongs to another context");
  if (!g->l64.p) return fail(NQ_EINVAL, "nq_tm_kern_square_noalloc: gauge handle without FP64 links");
  Guard G(c);
  char* base = nullptr;
  NQ_TRY(noalloc_scratch(c->device, &base));
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  SpinorView tmp, t1;
  tmp.data = base;
  t1.data = base + field;
  const bool prof = c->profiling;
  c->profiling = false;
  int rc = apply_kern(S, -mu, t1, view_of(in, 0), tmp);
  if (rc == NQ_OK) rc = apply_kern(S, +mu, view_of(out, 0), t1, tmp);
  c->profiling = prof;
  NQ_TRY(rc);
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}


int nq_tm_apply_Q(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, nq_field* out,
                  const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_LX, "nq_tm_apply_Q"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_LX, "nq_tm_apply_Q"));
  if (in == out) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  NQ_TRY(apply_Q_views(S, mu, out, in));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tm_apply_Q2(nq_ctx* c, const nq_gauge* g, double kappa, double mu, nq_field* out, const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_LX, "nq_tm_apply_Q2"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_LX, "nq_tm_apply_Q2"));
  if (in == out) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, nullptr, kappa, &S));
  nq_field tmp;
  tmp.ctx = c;
  tmp.kind = NQ_SPINCOLOR;
  tmp.coverage = NQ_LX;
  SpinorView tv;
  NQ_TRY(alloc_spinor_nr(c, "op.q2tmp", SP_F64, 12, 2, &tv));
  tmp.buf.p = tv.data;
  NQ_TRY(apply_Q_views(S, +mu, &tmp, in));
  NQ_TRY(apply_Q_views(S, -mu, out, &tmp));
  tmp.buf.p = nullptr;
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tm_cg_kernsq(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, int niter,
                    double residue, int prec, nq_field* sol, const nq_field* guess, const nq_field* src,
                    nq_solve_stats* st)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_ODD, "nq_tm_cg_kernsq"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_ODD, "nq_tm_cg_kernsq"));
  if (guess) NQ_TRY(check_field(c, guess, NQ_SPINCOLOR, NQ_ODD, "nq_tm_cg_kernsq"));
  Guard G(c);
  SpinorView gv;
  if (guess) gv = view_of(guess, 0);
  return cg_kernsq_any(c, g, kappa, anis, mu, niter, residue, prec, view_of(sol, 0), guess ? &gv : nullptr,
                       view_of(src, 0), st);
}

// inv_tmD_cg_eoprec_native + check (cg_invert_tmD_eoprec.cpp:102-150,192-212).
// guess_Koo (odd, optional): the start of the eq. 9 CG, overwritten with its solution
// ("if a guess was passed, return new one", :137)
int nq_tm_solve_eoprec(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, int niter,
                       double residue, int prec, nq_field* sol, nq_field* guess_Koo, const nq_field* src,
                       nq_solve_stats* st, double* check)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tm_solve_eoprec"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_LX, "nq_tm_solve_eoprec"));
  if (guess_Koo) NQ_TRY(check_field(c, guess_Koo, NQ_SPINCOLOR, NQ_ODD, "nq_tm_solve_eoprec (guess_Koo)"));
  if (sol == src) return fail(NQ_EINVAL, "sol==src");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, lk64_for(prec, g), anis, kappa, &S));
  const double a = 1 / (2 * kappa), n = a * a + mu * mu;
  const cplx zi_u(a / n, -mu / n), zi_l(a / n, mu / n);
  SpinorView se = view_of(src, 0), so = view_of(src, 1), xe = view_of(sol, 0), xo = view_of(sol, 1);
  SpinorView et, phi, y;
  NQ_TRY(alloc_spinor(c, "eo.et", SP_F64, 12, &et));
  NQ_TRY(alloc_spinor(c, "eo.phi", SP_F64, 12, &phi));
  NQ_TRY(alloc_spinor(c, "eo.y", SP_F64, 12, &y));
  SpinorView none;
  none.ncomp = 12;
  // eq. 8a: et = D_ee^-1 eta_e
  double ca[2][2], cb[2][2] = {{0, 0}, {0, 0}};
  setc(ca, zi_u, zi_l);
  NQ_TRY(launch_spin_diag(c, et, ca, se, cb, none));
  // eq. 8b: phi = g5 (eta_o + 1/2 H_oe et)
  HopEpi e8;
  setc(e8.c1, 0.5, -0.5);
  setc(e8.c2, 1.0, -1.0);
  e8.aux = so;
  NQ_TRY(launch_tm_hop(c, NQ_ODD, phi, et, S.U, S.anis, e8));
  // eq. 9
  nq_solve_stats local;
  if (!st) st = &local;
  // the CG's FP64 residuals compute K(-mu) y into sol_o: after convergence that is eq. 10
  SpinorView gv;
  if (guess_Koo) gv = view_of(guess_Koo, 0);
  int rc = cg_kernsq_any(c, g, kappa, anis, mu, niter, residue, prec, y, guess_Koo ? &gv : nullptr, phi, st, &xo);
  if (rc != NQ_OK && rc != NQ_ENOCONV) return rc;
  if (guess_Koo) NQ_TRY(launch_convert(c, gv, y));
  // eq. 10: sol_o = K(-mu) y  (sol_e used as scratch), unless the CG left it there
  if (rc != NQ_OK) NQ_TRY(apply_kern(S, -mu, xo, y, xe));
  // eq. 11: sol_e = D_ee^-1 (eta_e + 1/2 H_eo sol_o)
  HopEpi e11;
  setc(e11.c1, 0.5 * zi_u, 0.5 * zi_l);
  setc(e11.c2, zi_u, zi_l);
  e11.aux = se;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, xe, xo, S.U, S.anis, e11));
  if (check) {
    // |D sol - src|^2 / |src|^2, D = g5 Q, per parity: -1/2 hop + (a + i mu g5) sol - src
    // (18-real FP64 links: nissa's operator as stored)
    TmSetup Sc;
    NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &Sc));
    double n2[2] = {0, 0};
    for (int p = 0; p < 2; p++) {
      CgState sc = {};
      NQ_TRY(put_state(c, c->nstate - 1, sc));
      HopEpi ec;
      setc(ec.c1, -0.5, -0.5);
      setc(ec.c2, cplx(a, mu), cplx(a, -mu));
      setc(ec.c3, -1.0, -1.0);
      ec.aux = view_of(sol, p);
      ec.aux2 = view_of(src, p);
      ec.mode = 1;
      ec.write = 0;
      ec.st = scratch_state(c);
      ec.fin = FIN_STORE0;
      NQ_TRY(launch_tm_hop(c, p, none, view_of(sol, !p), Sc.U, Sc.anis, ec));
      NQ_TRY(get_states(c, c->nstate - 1, 1));
      n2[p] = c->state_host[c->nstate - 1].extra[0];
    }
    double s2 = 0;
    NQ_TRY(reduce_norm2(c, view_lx_flat(src), &s2));
    *check = (n2[0] + n2[1]) / s2;
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  return rc == NQ_ENOCONV ? fail(NQ_ENOCONV, "exit without converging") : NQ_OK;
}

int nq_tmQ2_cg(nq_ctx* c, const nq_gauge* g, double kappa, double mu, int niter, double residue, nq_field* sol,
               const nq_field* guess, const nq_field* src, nq_solve_stats* st)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmQ2_cg"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_LX, "nq_tmQ2_cg"));
  if (guess) NQ_TRY(check_field(c, guess, NQ_SPINCOLOR, NQ_LX, "nq_tmQ2_cg"));
  if (sol == src) return fail(NQ_EINVAL, "sol==src");
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, nullptr, kappa, &S));
  nq_solve_stats local;
  if (!st) st = &local;
  std::memset(st, 0, sizeof *st);
  const ApplyQ Q = [&](double m, const nq_field* o, const nq_field* i) { return apply_Q_views(S, m, o, i); };
  return q2_cg_f64(c, Q, mu, niter, residue, sol, guess, src, st);
}

// inv_tmD_cg_eoprec for several sources at once: eqs. 8 and 10-11 per source, the
// K K^+ CG batched over up to 16 sources per group
int nq_tm_solve_eoprec_multi(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, int nrhs,
                             const double* mus, int niter, double residue, int prec, nq_field** sols,
                             nq_field** guesses, const nq_field* const* srcs, nq_solve_stats* st)
{
  if (!c || !mus || !sols || !srcs) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_multi(c, nrhs, niter, residue, sols, guesses, srcs, "nq_tm_solve_eoprec_multi"));
  Guard G(c);
  TmBatch B;
  NQ_TRY(B.setup(c, g, kappa, anis, prec, 0));
  return run_batch_solves(c, B, nrhs, mus, niter, residue, sols, guesses, srcs, st);
}

int nq_tmQ2_cgm(nq_ctx* c, const nq_gauge* g, double kappa, int nmass, const double* masses, int niter,
                const double* residues, nq_field** sols, const nq_field* src, nq_solve_stats* st, double* true_res)
{
  if (!c || !masses || !residues || !sols) return fail(NQ_EINVAL, "null argument");
  if (nmass < 1 || nmass > kMaxShift) return fail(NQ_EINVAL, "need 1..%d masses", kMaxShift);
  if (niter <= 0) return fail(NQ_EINVAL, "bad niter");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmQ2_cgm"));
  for (int k = 0; k < nmass; k++) {
    NQ_TRY(check_field(c, sols[k], NQ_SPINCOLOR, NQ_LX, "nq_tmQ2_cgm"));
    if (sols[k] == src) return fail(NQ_EINVAL, "sol==src");
  }
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, nullptr, kappa, &S));
  const ApplyQ Q = [&](double m, const nq_field* o, const nq_field* i) { return apply_Q_views(S, m, o, i); };
  return q2_cgm_f64(c, Q, nmass, masses, niter, residues, sols, src, st, true_res);
}

// inv_tmDQ_cgm (cgm_invert_tmQ2.cpp:55-75): the same on g5 src
int nq_tmDQ_cgm(nq_ctx* c, const nq_gauge* g, double kappa, int nmass, const double* masses, int niter,
                const double* residues, nq_field** sols, const nq_field* src, nq_solve_stats* st, double* true_res)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmDQ_cgm"));
  Guard G(c);
  SpinorView gv;
  NQ_TRY(alloc_spinor_nr(c, "q2m.g5src", SP_F64, 12, 2, &gv));
  nq_field g5src;
  g5src.ctx = c;
  g5src.kind = NQ_SPINCOLOR;
  g5src.coverage = NQ_LX;
  g5src.buf.p = gv.data;
  double ca[2][2], cb[2][2] = {{0, 0}, {0, 0}};
  setc(ca, 1.0, -1.0);
  SpinorView none;
  int rc = NQ_OK;
  for (int p = 0; p < 2 && rc == NQ_OK; p++) rc = launch_spin_diag(c, view_of(&g5src, p), ca, view_of(src, p), cb, none);
  if (rc == NQ_OK) rc = nq_tmQ2_cgm(c, g, kappa, nmass, masses, niter, residues, sols, &g5src, st, true_res);
  g5src.buf.p = nullptr;
  return rc;
}

int nq_tm_kern_square_host(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, double* out,
                           const double* in)
{
  if (!c || !out || !in) return fail(NQ_EINVAL, "null argument");
  Guard G(c);
  nq_field *fi = nullptr, *fo = nullptr;
  NQ_TRY(nq_field_create(c, NQ_SPINCOLOR, NQ_ODD, &fi));
  int rc = nq_field_create(c, NQ_SPINCOLOR, NQ_ODD, &fo);
  if (rc == NQ_OK) rc = nq_field_upload(fi, in);
  if (rc == NQ_OK) rc = nq_tm_kern_square(c, g, kappa, anis, mu, fo, fi);
  if (rc == NQ_OK) rc = nq_field_download(fo, out);
  nq_field_destroy(fi);
  nq_field_destroy(fo);
  return rc;
}

int nq_tm_solve_eoprec_host(nq_ctx* c, const nq_gauge* g, double kappa, const double* anis, double mu, int niter,
                            double residue, int prec, double* sol, double* guess_Koo, const double* src,
                            nq_solve_stats* st, double* check)
{
  if (!c || !sol || !src) return fail(NQ_EINVAL, "null argument");
  Guard G(c);
  nq_field *fs = nullptr, *fx = nullptr, *fg = nullptr;
  NQ_TRY(nq_field_create(c, NQ_SPINCOLOR, NQ_LX, &fs));
  int rc = nq_field_create(c, NQ_SPINCOLOR, NQ_LX, &fx);
  if (rc == NQ_OK && guess_Koo) rc = nq_field_create(c, NQ_SPINCOLOR, NQ_ODD, &fg);
  if (rc == NQ_OK) rc = nq_field_upload(fs, src);
  if (rc == NQ_OK && fg) rc = nq_field_upload(fg, guess_Koo);
  int rs = rc == NQ_OK ? nq_tm_solve_eoprec(c, g, kappa, anis, mu, niter, residue, prec, fx, fg, fs, st, check) : rc;
  if (rs == NQ_OK || rs == NQ_ENOCONV) {
    rc = nq_field_download(fx, sol);
    if (rc == NQ_OK && fg) rc = nq_field_download(fg, guess_Koo);
  }
  nq_field_destroy(fs);
  nq_field_destroy(fx);
  nq_field_destroy(fg);
  return rc == NQ_OK ? rs : rc;
}

int nq_bench_tm_hop(nq_ctx* c, const nq_gauge* g, int prec, int napply, nq_field* out, const nq_field* in,
                    double* ms, double* bytes)
{
  if (!c || napply <= 0) return fail(NQ_EINVAL, "bad argument");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_bench_tm_hop"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_EVEN, "nq_bench_tm_hop"));
  Guard G(c);
  int sp, lk;
  NQ_TRY(lk_of_prec(prec, &sp, &lk));
  TmSetup S;
  NQ_TRY(tm_setup(c, g, lk, nullptr, 1.0, &S));
  SpinorView vi = view_of(in, 0), vo = view_of(out, 0);
  if (sp != SP_F64) {
    NQ_TRY(alloc_spinor(c, "bench.in", sp, 12, &vi));
    NQ_TRY(alloc_spinor(c, "bench.out", sp, 12, &vo));
    NQ_TRY(launch_convert(c, vi, view_of(in, 0)));
  }
  HopEpi e;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, vo, vi, S.U, S.anis, e));  // warm-up
  Timer T(c->stream);
  for (int i = 0; i < napply; i++) NQ_TRY(launch_tm_hop(c, NQ_EVEN, vo, vi, S.U, S.anis, e));
  const double sec = T.stop();
  if (ms) *ms = sec * 1e3 / napply;
  if (bytes) *bytes = tm_hop_bytes(sp, lk, e) * (double)c->geo.Vh;
  if (sp != SP_F64) NQ_TRY(launch_convert(c, view_of(out, 0), vo));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_bench_tm_hop_multi(nq_ctx* c, const nq_gauge* g, int prec, int nrhs, int napply, const nq_field* in,
                          double* ms, double* bytes)
{
  if (!c || napply <= 0 || nrhs < 1 || nrhs > 16) return fail(NQ_EINVAL, "bad argument");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_bench_tm_hop_multi"));
  Guard G(c);
  int sp, lk;
  NQ_TRY(lk_of_prec(prec, &sp, &lk));
  TmSetup S;
  NQ_TRY(tm_setup(c, g, lk, nullptr, 1.0, &S));
  SpinorView vi, vo;
  NQ_TRY(alloc_spinor_nr(c, "bench.min", sp, 12, nrhs, &vi));
  NQ_TRY(alloc_spinor_nr(c, "bench.mout", sp, 12, nrhs, &vo));
  for (int r = 0; r < nrhs; r++) NQ_TRY(launch_convert(c, slot_view(vi, r), view_of(in, 0)));
  HopEpi e;
  e.nrhs = nrhs;
  e.batched = 1;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, vo, vi, S.U, S.anis, e));  // warm-up
  Timer T(c->stream);
  for (int i = 0; i < napply; i++) NQ_TRY(launch_tm_hop(c, NQ_EVEN, vo, vi, S.U, S.anis, e));
  const double sec = T.stop();
  if (ms) *ms = sec * 1e3 / napply;
  if (bytes) *bytes = tm_hop_bytes(sp, lk, e) * (double)c->geo.Vh;
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

}  // extern "C"

// ==================================================================== twisted clover
// Twisted-clover Wilson operator (src/dirac_operators/tmclovD_eoprec/,
// src/dirac_operators/tmclovQ/dirac_operator_tmclovQ.cpp,
// src/inverters/twisted_clover/cg_invert_tmclovD_eoprec.cpp): the tm hop kernels
// with the site-local clover blocks of clover.hip in place of the diagonal
// D_ee/D_oo.  K(dag) = g5 (D_oo(dag) - 1/4 H_oe D_ee(dag)^-1 H_eo): a hop whose
// epilogue applies the inverse blocks and a second hop whose epilogue applies the
// D_oo blocks and combines; K K^+ = K(0) K(1) (tmclovDkern_eoprec_square_eos).

namespace nq {

// CG on K K^+ (cg_invert semantics, cg_invert_template_threaded.hpp:19-150); x holds
// the guess when use_guess
static int clov_cg(const ClovSetup& s, int niter, double residue, const SpinorView& x, bool use_guess,
                   const SpinorView& b, nq_solve_stats* out)
{
  nq_ctx* ctx = s.S.ctx;
  SpinorView r, p, t1, sv;
  NQ_TRY(alloc_spinor(ctx, "clov.r", SP_F64, 12, &r));
  NQ_TRY(alloc_spinor(ctx, "clov.p", SP_F64, 12, &p));
  NQ_TRY(alloc_spinor(ctx, "clov.t1", SP_F64, 12, &t1));
  NQ_TRY(alloc_spinor(ctx, "clov.s", SP_F64, 12, &sv));
  Timer T(ctx->stream);
  double b2 = 0;
  NQ_TRY(reduce_norm2(ctx, b, &b2));
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgState s0 = {};
  s0.b2 = b2;
  s0.target = residue * b2;
  s0.niter = niter;
  NQ_TRY(put_state(ctx, 0, s0));
  CgState* st = ctx->state;
  SpinorView none;
  none.ncomp = 12;
  if (use_guess) {
    NQ_TRY(clov_kern(s, 1, t1, x, HopEpi()));
    NQ_TRY(clov_kern(s, 0, sv, t1, HopEpi()));
    NQ_TRY(launch_lincomb(ctx, r, 1.0, b, -1.0, sv, 0, none, FIN_RR_INIT, st));
  } else {
    NQ_TRY(launch_zero(ctx, x));
    NQ_TRY(launch_lincomb(ctx, r, 1.0, b, 0, none, 0, none, FIN_RR_INIT, st));
  }
  NQ_TRY(launch_convert(ctx, p, r));
  NQ_TRY(get_states(ctx, 0, 1));
  const double rr0 = ctx->state_host[0].rr;
  int chunk = 8, launched = 0;
  while (!ctx->state_host[0].done) {
    for (int i = 0; i < chunk; i++) {
      // <p, K K^+ p> = |K^+ p|^2 = |K(1) p|^2
      HopEpi e1;
      e1.mode = 1;
      e1.st = st;
      e1.fin = FIN_PAP;
      NQ_TRY(clov_kern(s, 1, t1, p, e1));
      HopEpi e2;
      e2.mode = 2;
      e2.st = st;
      e2.fin = FIN_RR;
      e2.x = x;
      e2.p = p;
      e2.r = r;
      NQ_TRY(clov_kern(s, 0, none, t1, e2));
      NQ_TRY(launch_cg_pupdate(ctx, p, r, st));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, 1));
    const CgState& h = ctx->state_host[0];
    chunk = next_chunk(rr0, h.rr, h.target, h.iter, chunk);
    if (launched > niter + 64) break;
  }
  const CgState h = ctx->state_host[0];
  if (h.nan) return fail(NQ_ENAN, "nanned");
  // last calculation of the residual (cg_invert_template_threaded.hpp:131-135)
  NQ_TRY(clov_kern(s, 1, t1, x, HopEpi()));
  NQ_TRY(clov_kern(s, 0, sv, t1, HopEpi()));
  NQ_TRY(put_state(ctx, ctx->nstate - 1, CgState{}));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, -1.0, sv, 0, none, FIN_STORE0, scratch_state(ctx)));
  NQ_TRY(get_states(ctx, ctx->nstate - 1, 1));
  const double true_rr = ctx->state_host[ctx->nstate - 1].extra[0];
  out->iters = h.iter;
  out->reliable_updates = 0;
  out->rel_res_sq = true_rr / b2;
  out->converged = h.iter < niter;
  out->seconds = T.stop();
  if (!out->converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations)", h.iter);
  return NQ_OK;
}

// Mixed precision: the CG iterates on the FP32 operator (s32) with the FP32 x/r of the
// tm mixed solver and FP64 reliable updates every reliable_delta drop of |r| (x64 += xs +
// alpha p; r64 = b - K K^+ x64 on the FP64 operator s64; rs = r64; p = rs + beta p).
static int clov_cg_mixed(const ClovSetup& s64, const ClovSetup& s32, int sp, int niter, double residue,
                         const SpinorView& x64, bool use_guess, const SpinorView& b, nq_solve_stats* out)
{
  nq_ctx* ctx = s64.S.ctx;
  SpinorView r64, t1, sv, xs, rs, p, t1s;
  NQ_TRY(alloc_spinor(ctx, "clov.r", SP_F64, 12, &r64));
  NQ_TRY(alloc_spinor(ctx, "clov.t1", SP_F64, 12, &t1));
  NQ_TRY(alloc_spinor(ctx, "clov.s", SP_F64, 12, &sv));
  NQ_TRY(alloc_spinor(ctx, "clov.xs32", SP_F32, 12, &xs));
  NQ_TRY(alloc_spinor(ctx, "clov.rs32", SP_F32, 12, &rs));
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "clov.p16" : "clov.p32", sp, 12, &p));
  // two search-direction buffers for the paired x updates (as in the twisted-mass CG)
  SpinorView pb;
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "clov.p16b" : "clov.p32b", sp, 12, &pb));
  const SpinorView pv[2] = {p, pb};
  NQ_TRY(alloc_spinor(ctx, sp == SP_F16 ? "clov.t1s16" : "clov.t1s32", sp, 12, &t1s));
  Timer T(ctx->stream);
  double b2 = 0;
  NQ_TRY(reduce_norm2(ctx, b, &b2));
  if (b2 == 0 || std::isnan(b2)) return fail(NQ_ENAN, "invalid norm: %g", b2);
  CgState s0 = {};
  s0.b2 = b2;
  s0.target = residue * b2;
  s0.niter = niter;
  s0.mixed = 1;
  s0.delta = reliable_delta(sp, residue);
  NQ_TRY(put_state(ctx, 0, s0));
  CgState* st = ctx->state;
  SpinorView none;
  none.ncomp = 12;
  auto residual = [&](int fin) -> int {
    NQ_TRY(clov_kern(s64, 1, t1, x64, HopEpi()));
    NQ_TRY(clov_kern(s64, 0, sv, t1, HopEpi()));
    return launch_lincomb(ctx, r64, 1.0, b, -1.0, sv, 0, none, fin, st);
  };
  if (use_guess) {
    NQ_TRY(residual(FIN_RR_INIT));
  } else {
    NQ_TRY(launch_zero(ctx, x64));
    NQ_TRY(launch_lincomb(ctx, r64, 1.0, b, 0, none, 0, none, FIN_RR_INIT, st));
  }
  NQ_TRY(launch_convert(ctx, rs, r64));
  NQ_TRY(launch_convert(ctx, p, rs));
  NQ_TRY(launch_zero(ctx, xs));
  NQ_TRY(get_states(ctx, 0, 1));
  const double rr_init = ctx->state_host[0].rr;
  int chunk = 8, launched = 0;
  while (!ctx->state_host[0].done) {
    const int par = ctx->state_host[0].iter & 1;  // iteration j reads p[j % 2]
    for (int i = 0; i < chunk; i++) {
      const SpinorView& p = pv[(par + i) & 1];
      const SpinorView& pn = pv[(par + i + 1) & 1];
      // <p, K K^+ p> reduced by the second hop, finalised by an extra workgroup of the
      // third (which does not need alpha)
      HopEpi e1;
      e1.mode = 1;
      e1.st = st;
      e1.fin = FIN_PARTIAL;
      e1.pred_reliable = 2;
      NQ_TRY(clov_kern(s32, 1, t1s, p, e1));
      HopEpi e2;
      e2.mode = 2;
      e2.st = st;
      e2.fin = FIN_RR_SLOPPY;
      e2.pred_reliable = 2;
      e2.p = p;
      e2.r = rs;
      NQ_TRY(clov_kern(s32, 0, none, t1s, e2, FIN_PAP));
      NQ_TRY(launch_cg_pupdate_pair(ctx, p, pn, rs, st, xs));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, 1));
    if (ctx->state_host[0].reliable && !ctx->state_host[0].done) {
      const int j = ctx->state_host[0].iter - 1;  // the flagged iteration, its p update skipped
      NQ_TRY(launch_reliable_accum(ctx, x64, xs, st, 1, pv[j & 1], pv[(j + 1) & 1]));
      NQ_TRY(residual(FIN_RR_TRUE));
      NQ_TRY(launch_convert(ctx, rs, r64));
      NQ_TRY(launch_cg_pupdate_pair(ctx, pv[j & 1], pv[(j + 1) & 1], rs, st));
      NQ_TRY(get_states(ctx, 0, 1));
    }
    const CgState& h = ctx->state_host[0];
    if (h.nan || h.done) break;
    const double goal = std::max(h.delta * h.delta * h.r0 * h.r0, h.target);
    const double slope = h.iter > 0 && h.rr < rr_init ? std::log(h.rr / rr_init) / h.iter : -0.3;
    const int need = (int)std::ceil(std::log(goal / h.rr) / slope);
    chunk = std::max(1, std::min(32, need));
    if (launched > niter + 64) break;
  }
  const CgState h = ctx->state_host[0];
  if (h.nan) return fail(NQ_ENAN, "nanned");
  out->iters = h.iter;
  out->reliable_updates = h.nreliable;
  out->rel_res_sq = h.true_rr / b2;
  out->converged = h.true_rr < h.target;
  out->seconds = T.stop();
  if (!out->converged) return fail(NQ_ENOCONV, "exit without converging (%d iterations, rel res %g)", h.iter, h.true_rr / b2);
  return NQ_OK;
}

// CG on K K^+ in the requested precision (NQ_PREC_F64, or NQ_PREC_F32 / NQ_PREC_F16
// inner iterations with FP64 reliable updates)
static int clov_cg_any(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                       int prec, const ClovSetup& s64, int niter, double residue, const SpinorView& x, bool use_guess,
                       const SpinorView& b, nq_solve_stats* st)
{
  if (prec == NQ_PREC_F64) return clov_cg(s64, niter, residue, x, use_guess, b, st);
  int sp = 0, lk = 0;
  NQ_TRY(lk_of_prec(prec, &sp, &lk));
  ClovSetup ss;
  NQ_TRY(clov_setup_sloppy(c, g, cl, kappa, anis, mu, sp, lk, &ss));
  return clov_cg_mixed(s64, ss, sp, niter, residue, x, use_guess, b, st);
}

}  // namespace nq

extern "C" {

int nq_clover_create(nq_ctx* c, const nq_gauge* g, double csw, nq_clover** out)
{
  if (!c || !g || !out) return fail(NQ_EINVAL, "null argument");
  if (g->ctx != c) return fail(NQ_EINVAL, "gauge handle belongs to another context");
  if (std::isnan(csw)) return fail(NQ_EINVAL, "bad cSW");
  Guard G(c);
  auto* cl = new nq_clover;
  cl->ctx = c;
  cl->csw = csw;
  int rc = NQ_OK;
  for (int p = 0; p < 2 && rc == NQ_OK; p++) rc = dev_alloc(cl->cl[p], (size_t)36 * 16 * c->geo.stride);
  for (int p = 0; p < 2 && rc == NQ_OK; p++)
    rc = launch_clover_build(c, (const double*)g->aos[0].p, (const double*)g->aos[1].p, p, csw, clov_view(cl, p));
  if (rc == NQ_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(NQ_EDEVICE, "clover build failed");
  if (rc != NQ_OK) {
    nq_clover_destroy(cl);
    return rc;
  }
  *out = cl;
  return NQ_OK;
}

int nq_clover_download_lx(const nq_clover* cl, double* cl_lx)
{
  if (!cl || !cl_lx) return fail(NQ_EINVAL, "null argument");
  nq_ctx* c = cl->ctx;
  Guard G(c);
  DevBuf d;
  NQ_TRY(dev_alloc(d, sizeof(double) * 72 * 2 * c->geo.Vh));
  int rc = NQ_OK;
  for (int p = 0; p < 2 && rc == NQ_OK; p++) rc = launch_soa_to_aos(c, clov_view(cl, p), 36, p, 1, (double*)d.p);
  if (rc == NQ_OK && hipMemcpyAsync(cl_lx, d.p, sizeof(double) * 72 * 2 * c->geo.Vh, hipMemcpyDeviceToHost,
                                    c->stream) != hipSuccess)
    rc = fail(NQ_EDEVICE, "download failed");
  if (rc == NQ_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(NQ_EDEVICE, "download failed");
  dev_free(d);
  return rc;
}

void nq_clover_destroy(nq_clover* cl)
{
  if (!cl) return;
  {
    Guard G(cl->ctx);
    for (int p = 0; p < 2; p++) {
      dev_free(cl->cl[p]);
      dev_free(cl->inv[p]);
      for (int i = 0; i < 2; i++) {
        dev_free(cl->cls[i][p]);
        dev_free(cl->invs[i][p]);
      }
    }
  }
  delete cl;
}

int nq_tmclov_Dee(nq_ctx* c, nq_clover* cl, double kappa, double mu, int dag, nq_field* out, const nq_field* in)
{
  if (!c || !cl || !out || !in) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  if (in->coverage == NQ_LX || out->coverage != in->coverage) return fail(NQ_EINVAL, "nq_tmclov_Dee: parity mismatch");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, in->coverage, "nq_tmclov_Dee"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, in->coverage, "nq_tmclov_Dee"));
  if (in == out) return fail(NQ_EINVAL, "in==out!");
  if (!(kappa != 0)) return fail(NQ_EINVAL, "bad kappa");
  Guard G(c);
  SpinorView none;
  NQ_TRY(launch_clover_apply(c, 0, view_of(out, 0), view_of(in, 0), clov_view(cl, in->coverage), none, kappa, mu, dag,
                             kOne, nullptr));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tmclov_inv_Dee(nq_ctx* c, nq_clover* cl, double kappa, double mu, int dag, nq_field* out, const nq_field* in)
{
  if (!c || !cl || !out || !in) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  if (in->coverage == NQ_LX || out->coverage != in->coverage)
    return fail(NQ_EINVAL, "nq_tmclov_inv_Dee: parity mismatch");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, in->coverage, "nq_tmclov_inv_Dee"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, in->coverage, "nq_tmclov_inv_Dee"));
  if (in == out) return fail(NQ_EINVAL, "in==out!");
  if (!(kappa != 0)) return fail(NQ_EINVAL, "bad kappa");
  Guard G(c);
  SpinorView inv, none;
  NQ_TRY(clov_inv_view(cl, in->coverage, kappa, mu, &inv));
  NQ_TRY(launch_clover_apply(c, 1, view_of(out, 0), view_of(in, 0), inv, none, kappa, mu, dag, kOne, nullptr));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tmclov_kern(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu, int dag,
                   nq_field* out, const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_kern"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_kern"));
  if (in == out) return fail(NQ_EINVAL, "in==out!");
  Guard G(c);
  ClovSetup s;
  NQ_TRY(clov_setup(c, g, cl, kappa, anis, mu, &s));
  NQ_TRY(clov_kern(s, dag ? 1 : 0, view_of(out, 0), view_of(in, 0), HopEpi()));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_tmclov_kern_square(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                          nq_field* out, const nq_field* in)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_kern_square"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_kern_square"));
  if (in == out) return fail(NQ_EINVAL, "in==out!");
  Guard G(c);
  ClovSetup s;
  NQ_TRY(clov_setup(c, g, cl, kappa, anis, mu, &s));
  SpinorView t1;
  NQ_TRY(alloc_spinor(c, "clov.t1", SP_F64, 12, &t1));
  NQ_TRY(clov_kern(s, 1, t1, view_of(in, 0), HopEpi()));
  NQ_TRY(clov_kern(s, 0, view_of(out, 0), t1, HopEpi()));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

// Q = g5 D, D = 1/(2 kappa) + i mu g5 + Cl - 1/2 hop on each parity
// apply_tmclovQ on lx views: per parity, out = gamma5 (D_pp-block in_p - 1/2 hop in_!p)
static int clov_apply_Q_views(const TmSetup& S, nq_clover* cl, double mu, const nq_field* out, const nq_field* in)
{
  nq_ctx* c = S.ctx;
  SpinorView u, none;
  NQ_TRY(alloc_spinor(c, "clov.uo", SP_F64, 12, &u));
  for (int p = 0; p < 2; p++) {
    NQ_TRY(launch_clover_apply(c, 0, u, view_of(in, p), clov_view(cl, p), none, S.kappa, mu, 0, kOne, nullptr));
    HopEpi e;
    setc(e.c1, -0.5, 0.5);
    setc(e.c2, 1.0, -1.0);
    e.aux = u;
    NQ_TRY(launch_tm_hop(c, p, view_of(out, p), view_of(in, !p), S.U, S.anis, e));
  }
  return NQ_OK;
}

int nq_tmclov_apply_Q(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                      nq_field* out, const nq_field* in)
{
  if (!c || !cl) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_apply_Q"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_apply_Q"));
  if (in == out) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  NQ_TRY(clov_apply_Q_views(S, cl, mu, out, in));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

// apply_tmclovQ2 (dirac_operators/tmclovQ2/dirac_operator_tmclovQ2.hpp:9-20):
// Q(+mu) into a temporary, then Q(-mu)
int nq_tmclov_apply_Q2(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                       nq_field* out, const nq_field* in)
{
  if (!c || !cl) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  NQ_TRY(check_field(c, in, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_apply_Q2"));
  NQ_TRY(check_field(c, out, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_apply_Q2"));
  if (in == out) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  SpinorView tv;
  NQ_TRY(alloc_spinor_nr(c, "clov.q2t", SP_F64, 12, 2, &tv));
  nq_field ft;
  ft.ctx = c;
  ft.kind = NQ_SPINCOLOR;
  ft.coverage = NQ_LX;
  ft.buf.p = tv.data;
  struct Release {
    nq_field& f;
    ~Release() { f.buf.p = nullptr; }
  } rel{ft};
  NQ_TRY(clov_apply_Q_views(S, cl, +mu, &ft, in));
  NQ_TRY(clov_apply_Q_views(S, cl, -mu, out, &ft));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

// inv_tmclovQ2_cg (inverters/twisted_clover/cg_invert_tmclovQ2.hpp:9-24 ->
// cg_64_invert_tmclovQ2_portable.hpp:14-40): cg_invert on apply_tmclovQ2 over lx fields
int nq_tmclovQ2_cg(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                   int niter, double residue, nq_field* sol, const nq_field* guess, const nq_field* src,
                   nq_solve_stats* st)
{
  if (!c || !cl) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmclovQ2_cg"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_LX, "nq_tmclovQ2_cg"));
  if (guess) NQ_TRY(check_field(c, guess, NQ_SPINCOLOR, NQ_LX, "nq_tmclovQ2_cg"));
  if (sol == src) return fail(NQ_EINVAL, "sol==src");
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  nq_solve_stats local;
  if (!st) st = &local;
  std::memset(st, 0, sizeof *st);
  const ApplyQ Q = [&](double m, const nq_field* o, const nq_field* i) { return clov_apply_Q_views(S, cl, m, o, i); };
  return q2_cg_f64(c, Q, mu, niter, residue, sol, guess, src, st);
}

// inv_tmclovQ2_cgm (inverters/twisted_clover/cgm_invert_tmclovQ2.cpp:8-28): cgm_invert
// with shifts m_k^2 on apply_tmclovQ2_m2 (Q(0)^2 + m_k^2), lx fields, FP64
int nq_tmclovQ2_cgm(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, int nmass,
                    const double* masses, int niter, const double* residues, nq_field** sols, const nq_field* src,
                    nq_solve_stats* st, double* true_res)
{
  if (!c || !cl || !masses || !residues || !sols) return fail(NQ_EINVAL, "null argument");
  if (cl->ctx != c) return fail(NQ_EINVAL, "clover handle belongs to another context");
  if (nmass < 1 || nmass > kMaxShift) return fail(NQ_EINVAL, "need 1..%d masses", kMaxShift);
  if (niter <= 0) return fail(NQ_EINVAL, "bad niter");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmclovQ2_cgm"));
  for (int k = 0; k < nmass; k++) {
    NQ_TRY(check_field(c, sols[k], NQ_SPINCOLOR, NQ_LX, "nq_tmclovQ2_cgm"));
    if (sols[k] == src) return fail(NQ_EINVAL, "sol==src");
  }
  Guard G(c);
  TmSetup S;
  NQ_TRY(tm_setup(c, g, LK_F64_18, anis, kappa, &S));
  const ApplyQ Q = [&](double m, const nq_field* o, const nq_field* i) { return clov_apply_Q_views(S, cl, m, o, i); };
  return q2_cgm_f64(c, Q, nmass, masses, niter, residues, sols, src, st, true_res);
}

int nq_tmclov_cg_kernsq(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                        int niter, double residue, int prec, nq_field* sol, const nq_field* guess, const nq_field* src,
                        nq_solve_stats* st)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_cg_kernsq"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_cg_kernsq"));
  if (guess) NQ_TRY(check_field(c, guess, NQ_SPINCOLOR, NQ_ODD, "nq_tmclov_cg_kernsq"));
  if (sol == src) return fail(NQ_EINVAL, "sol==src");
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  Guard G(c);
  ClovSetup s;
  NQ_TRY(clov_setup(c, g, cl, kappa, anis, mu, &s));
  nq_solve_stats local;
  if (!st) st = &local;
  std::memset(st, 0, sizeof *st);
  if (guess && guess != sol) NQ_TRY(launch_convert(c, view_of(sol, 0), view_of(guess, 0)));
  const int rc = clov_cg_any(c, g, cl, kappa, anis, mu, prec, s, niter, residue, view_of(sol, 0), guess != nullptr,
                             view_of(src, 0), st);
  NQ_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

// inv_tmclovD_cg_eoprec (cg_invert_tmclovD_eoprec.cpp:56-128 and the check of :179-232)
int nq_tmclov_solve_eoprec(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                           int niter, double residue, int prec, nq_field* sol, const nq_field* src, nq_solve_stats* st,
                           double* check)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  NQ_TRY(check_field(c, src, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_solve_eoprec"));
  NQ_TRY(check_field(c, sol, NQ_SPINCOLOR, NQ_LX, "nq_tmclov_solve_eoprec"));
  if (sol == src) return fail(NQ_EINVAL, "sol==src");
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  Guard G(c);
  ClovSetup s;
  NQ_TRY(clov_setup(c, g, cl, kappa, anis, mu, &s));
  SpinorView se = view_of(src, 0), so = view_of(src, 1), xe = view_of(sol, 0), xo = view_of(sol, 1);
  SpinorView et, phi, y, none;
  none.ncomp = 12;
  NQ_TRY(alloc_spinor(c, "eo.et", SP_F64, 12, &et));
  NQ_TRY(alloc_spinor(c, "eo.phi", SP_F64, 12, &phi));
  NQ_TRY(alloc_spinor(c, "eo.y", SP_F64, 12, &y));
  // eq. 8a: et = D_ee^-1 eta_e
  NQ_TRY(launch_clover_apply(c, 1, et, se, s.inv_e, none, kappa, mu, 0, kOne, nullptr));
  // eq. 8b: phi = g5 (eta_o + 1/2 H_oe et)
  HopEpi e8;
  setc(e8.c1, 0.5, -0.5);
  setc(e8.c2, 1.0, -1.0);
  e8.aux = so;
  NQ_TRY(launch_tm_hop(c, NQ_ODD, phi, et, s.S.U, s.S.anis, e8));
  // eq. 9
  nq_solve_stats local;
  if (!st) st = &local;
  std::memset(st, 0, sizeof *st);
  int rc = clov_cg_any(c, g, cl, kappa, anis, mu, prec, s, niter, residue, y, false, phi, st);
  if (rc != NQ_OK && rc != NQ_ENOCONV) return rc;
  // eq. 10: sol_o = K(dag) y
  NQ_TRY(clov_kern(s, 1, xo, y, HopEpi()));
  // eq. 11: sol_e = D_ee^-1 (eta_e + 1/2 H_eo sol_o)
  HopEpi e11;
  setc(e11.c1, 0.5, 0.5);
  setc(e11.c2, 1.0, 1.0);
  e11.aux = se;
  NQ_TRY(launch_tm_hop(c, NQ_EVEN, et, xo, s.S.U, s.S.anis, e11));
  NQ_TRY(launch_clover_apply(c, 1, xe, et, s.inv_e, none, kappa, mu, 0, kOne, nullptr));
  if (check) {
    // |D sol - src|^2 / |src|^2 per parity: (z + Cl) sol_p - 1/2 hop sol_!p - src_p
    double n2[2] = {0, 0};
    for (int p = 0; p < 2; p++) {
      NQ_TRY(launch_clover_apply(c, 0, s.uo, view_of(sol, p), clov_view(cl, p), none, kappa, mu, 0, kOne, nullptr));
      NQ_TRY(put_state(c, c->nstate - 1, CgState{}));
      HopEpi ec;
      setc(ec.c1, -0.5, -0.5);
      setc(ec.c2, 1.0, 1.0);
      setc(ec.c3, -1.0, -1.0);
      ec.aux = s.uo;
      ec.aux2 = view_of(src, p);
      ec.mode = 1;
      ec.write = 0;
      ec.st = scratch_state(c);
      ec.fin = FIN_STORE0;
      NQ_TRY(launch_tm_hop(c, p, none, view_of(sol, !p), s.S.U, s.S.anis, ec));
      NQ_TRY(get_states(c, c->nstate - 1, 1));
      n2[p] = c->state_host[c->nstate - 1].extra[0];
    }
    double s2 = 0;
    NQ_TRY(reduce_norm2(c, view_lx_flat(src), &s2));
    *check = (n2[0] + n2[1]) / s2;
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  return rc == NQ_ENOCONV ? fail(NQ_ENOCONV, "exit without converging") : NQ_OK;
}

// inv_tmclovD_cg_eoprec for several sources at one (kappa, mu, cSW) (the propagator
// driver's 12 spin-colour sources, projects/ib/prop.cpp:93-96): the K K^+ CG batched
// over up to 16 sources per group, sources streaming through the slots
int nq_tmclov_solve_eoprec_multi(nq_ctx* c, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                                 double mu, int nrhs, int niter, double residue, int prec, nq_field** sols,
                                 const nq_field* const* srcs, nq_solve_stats* st)
{
  if (!c || !cl || !sols || !srcs) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_multi(c, nrhs, niter, residue, sols, nullptr, srcs, "nq_tmclov_solve_eoprec_multi"));
  Guard G(c);
  TmBatch B;
  NQ_TRY(B.setup(c, g, kappa, anis, prec, 0, cl, mu));
  std::vector<double> mus(nrhs, mu);
  return run_batch_solves(c, B, nrhs, mus.data(), niter, residue, sols, nullptr, srcs, st);
}

}  // extern "C"
