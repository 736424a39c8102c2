// Staggered operators and solvers on the device.
//
// Reference: src/dirac_operators/stD/dirac_operator_stD.hpp (apply_st2Doe,
// apply_stDeo_half, apply_stD2ee_m2), src/inverters/staggered/cg_invert_stD2ee_m2.hpp
// (CG on m2 - 1/4 Deo Doe) and cgm_invert_template_threaded.hpp with
// ApplyStD2eeM2Functor (multi-shift CG, one Krylov space per source for all shifts).
//
// The multi-shift solve is batched over sources: one launch of each CGM kernel
// serves every source, so each loaded link is reused across right-hand sides.
// Mixed precision (NQ_PREC_F32): the CGM runs on FP32 fields and links to a
// residual the FP32 iteration can reach, then every (source, shift) system is
// refined by a batched CG with FP32 inner iterations and FP64 reliable updates
// that is warm-started from the CGM solution; convergence is declared on the
// FP64 true residual of each shifted system.

#include <algorithm>
#include <map>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nq_internal.hpp"

namespace nq {

namespace {
struct Guard {
  std::lock_guard<std::recursive_mutex> lk;
  explicit Guard(nq_ctx* c) : lk(c->mu) { (void)hipSetDevice(c->device); }
};
int check_color(const nq_ctx* c, const nq_field* f, int cov, const char* what)
{
  if (!f) return fail(NQ_EINVAL, "%s: null field", what);
  if (f->ctx != c) return fail(NQ_EINVAL, "%s: field belongs to another context", what);
  if (f->kind != NQ_COLOR || f->coverage != cov) return fail(NQ_EINVAL, "%s: wrong field kind/coverage", what);
  return NQ_OK;
}
int put_states(nq_ctx* ctx, int i0, int n, const CgState* s)
{
  std::memcpy(ctx->state_host + i0, s, sizeof(CgState) * n);
  NQ_HIP(hipMemcpyAsync(ctx->state + i0, ctx->state_host + i0, sizeof(CgState) * n, hipMemcpyHostToDevice, ctx->stream));
  return NQ_OK;
}
int get_states(nq_ctx* ctx, int i0, int n)
{
  NQ_HIP(hipMemcpyAsync(ctx->state_host + i0, ctx->state + i0, sizeof(CgState) * n, hipMemcpyDeviceToHost, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  return NQ_OK;
}
constexpr int kStateInts = (int)(sizeof(CgState) / sizeof(int));
SpinorView shifted(const SpinorView& v, int64_t elems)
{
  SpinorView o = v;
  const int eb = v.prec == SP_F64 ? 16 : (v.prec == SP_F32 ? 8 : 4);
  o.data = (char*)v.data + elems * eb;
  return o;
}
}  // namespace

// A(m2) x = m2 x - 1/4 Deo Doe x on nr right-hand sides (t: odd scratch)
static int apply_DDee(nq_ctx* ctx, const LinkView& U, double mass2, const double* coef, const SpinorView& out,
                      const SpinorView& in, const SpinorView& t, int nr)
{
  SpinorView none;
  none.ncomp = 3;
  NQ_TRY(launch_st_hop(ctx, NQ_ODD, t, in, U, 1.0, 0.0, nullptr, none, 0, 1, nr, FIN_NONE, nullptr));
  const bool aux = coef || mass2 != 0;
  return launch_st_hop(ctx, NQ_EVEN, out, t, U, -0.25, mass2, coef, aux ? in : none, 0, 1, nr, FIN_NONE, nullptr);
}

// Batched CG on A(sigma_i) x_i = b_i for nr systems (per-system shift in coef[2i+1]).
// prec NQ_PREC_F64: plain CG (cg_invert semantics); NQ_PREC_F32: FP32 inner iteration
// with FP64 reliable updates.  x64 holds the guess on entry (use_guess) and the
// solution on exit; residues[i] is the target of |r_i|^2/|b_i|^2.  States: ctx->state[0..nr).
// Smallest zeta of the reference shift for which the mixed multi-shift CG replaces
// its Krylov residual by the FP64 residual of that shift divided by zeta.
constexpr double kMinZetaReplace = 1e-2;

static int st_cg_batched(nq_ctx* ctx, const nq_gauge* g, int nr, const double* shifts, int niter,
                         const double* residues, int prec, const SpinorView& x64, const SpinorView& b, bool use_guess,
                         nq_solve_stats* st)
{
  LinkView U64, Us;
  NQ_TRY(link_view(g, LK_F64_18, &U64));
  const bool mixed = prec != NQ_PREC_F64;
  if (mixed) NQ_TRY(link_view(g, LK_F32_18, &Us));
  double* coef = nullptr;
  NQ_TRY(ctx_temp(ctx, "stcg.coef", sizeof(double) * 2 * 16, (void**)&coef));
  std::vector<double> hc(2 * nr);
  for (int i = 0; i < nr; i++) { hc[2 * i] = -0.25; hc[2 * i + 1] = shifts[i]; }
  NQ_HIP(hipMemcpyAsync(coef, hc.data(), sizeof(double) * 2 * nr, hipMemcpyHostToDevice, ctx->stream));
  SpinorView r64, t64, s64, xs, rs, p, t;
  NQ_TRY(alloc_spinor_nr(ctx, "stcg.r", SP_F64, 3, nr, &r64));
  NQ_TRY(alloc_spinor_nr(ctx, "stcg.t", SP_F64, 3, nr, &t64));
  NQ_TRY(alloc_spinor_nr(ctx, "stcg.s", SP_F64, 3, nr, &s64));
  if (mixed) {
    NQ_TRY(alloc_spinor_nr(ctx, "stcg.xs", SP_F32, 3, nr, &xs));
    NQ_TRY(alloc_spinor_nr(ctx, "stcg.rs", SP_F32, 3, nr, &rs));
    NQ_TRY(alloc_spinor_nr(ctx, "stcg.p32", SP_F32, 3, nr, &p));
    NQ_TRY(alloc_spinor_nr(ctx, "stcg.t32", SP_F32, 3, nr, &t));
  } else {
    NQ_TRY(alloc_spinor_nr(ctx, "stcg.p", SP_F64, 3, nr, &p));
    t = t64;
    xs = x64;
    rs = r64;
  }
  SpinorView none;
  none.ncomp = 3;
  // |b|^2 per system
  std::vector<CgState> h0(nr);
  for (int i = 0; i < nr; i++) h0[i] = CgState{};
  NQ_TRY(put_states(ctx, 0, nr, h0.data()));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, 0, none, 0, none, FIN_STORE0, ctx->state, nr));
  NQ_TRY(get_states(ctx, 0, nr));
  std::vector<double> b2(nr);
  for (int i = 0; i < nr; i++) {
    b2[i] = ctx->state_host[i].extra[0];
    if (b2[i] == 0 || std::isnan(b2[i])) return fail(NQ_ENAN, "invalid norm: %g", b2[i]);
    CgState s0 = {};
    s0.b2 = b2[i];
    s0.target = residues[i] * b2[i];
    s0.niter = niter;
    s0.mixed = mixed;
    s0.delta = 1e-2;
    h0[i] = s0;
  }
  NQ_TRY(put_states(ctx, 0, nr, h0.data()));
  CgState* S = ctx->state;
  // r = b - A x
  if (use_guess) {
    NQ_TRY(apply_DDee(ctx, U64, 0, coef, s64, x64, t64, nr));
    NQ_TRY(launch_lincomb(ctx, r64, 1.0, b, -1.0, s64, 0, none, FIN_RR_INIT, S, nr));
  } else {
    NQ_TRY(launch_zero(ctx, x64, nr));
    NQ_TRY(launch_lincomb(ctx, r64, 1.0, b, 0, none, 0, none, FIN_RR_INIT, S, nr));
  }
  if (mixed) {
    NQ_TRY(launch_convert(ctx, rs, r64, nr));
    NQ_TRY(launch_zero(ctx, xs, nr));
  }
  NQ_TRY(launch_convert(ctx, p, rs, nr));
  NQ_TRY(get_states(ctx, 0, nr));
  const LinkView& Ui = mixed ? Us : U64;
  int chunk = 8, launched = 0;
  auto all_done = [&]() {
    for (int i = 0; i < nr; i++)
      if (!ctx->state_host[i].done) return false;
    return true;
  };
  while (!all_done()) {
    for (int it = 0; it < chunk; it++) {
      // t = Doe p with <p, A p> = sigma |p|^2 + |t|^2/4 (Deo^+ = -Doe) -> alpha;
      // s = sigma p - 1/4 Deo t fused with x += alpha p, r -= alpha s, |r|^2 -> beta;
      // p = r + beta p (idle systems skipped throughout)
      NQ_TRY(launch_st_hop(ctx, NQ_ODD, t, p, Ui, 1.0, 0.0, coef, none, 6, 1, nr, FIN_PAP, S, &S->done, kStateInts,
                           2));
      NQ_TRY(launch_st_cgtail(ctx, t, Ui, coef, p, xs, rs, S, mixed ? FIN_RR_SLOPPY : FIN_RR, nr));
      NQ_TRY(launch_cg_pupdate(ctx, p, rs, S, nr));
    }
    launched += chunk;
    NQ_TRY(get_states(ctx, 0, nr));
    bool any_rel = false;
    for (int i = 0; i < nr; i++) any_rel |= ctx->state_host[i].reliable && !ctx->state_host[i].done;
    if (mixed && any_rel) {
      // reliable update of the flagged systems: x64 += xs, xs = 0; r64 = b - A x64 in
      // FP64; rs = r64; p = rs + beta p (beta from the true residual); then clear flags
      NQ_TRY(launch_reliable_accum(ctx, x64, xs, S, nr));
      NQ_TRY(apply_DDee(ctx, U64, 0, coef, s64, x64, t64, nr));
      NQ_TRY(launch_lincomb(ctx, r64, 1.0, b, -1.0, s64, 0, none, FIN_RR_TRUE_HOLD, S, nr, 2));
      NQ_TRY(launch_convert(ctx, rs, r64, nr, S, 2));
      NQ_TRY(launch_cg_pupdate(ctx, p, rs, S, nr, 2));
      NQ_TRY(launch_clear_reliable(ctx, S, nr));
      NQ_TRY(get_states(ctx, 0, nr));
    }
    bool stop = true;
    int need = 2;
    for (int i = 0; i < nr; i++) {
      const CgState& h = ctx->state_host[i];
      if (h.nan) return fail(NQ_ENAN, "nanned");
      if (h.done) continue;
      stop = false;
      const double goal = mixed ? std::max(h.delta * h.delta * h.r0 * h.r0, h.target) : h.target;
      const double slope = h.iter > 0 && h.rr < h.b2 ? std::log(h.rr / h.b2) / h.iter : -0.1;
      need = std::max(need, (int)std::ceil(std::log(goal / h.rr) / slope));
    }
    if (stop) break;
    chunk = std::max(2, std::min(64, need));
    if (launched > niter + 64) break;
  }
  int rc = NQ_OK;
  for (int i = 0; i < nr; i++) {
    const CgState& h = ctx->state_host[i];
    st[i].iters = h.iter;
    st[i].reliable_updates = h.nreliable;
    st[i].converged = mixed ? h.true_rr < h.target : h.rr < h.target;
    if (!st[i].converged) rc = NQ_ENOCONV;
  }
  // FP64 true residual of every system
  std::vector<CgState> z(nr);
  NQ_TRY(put_states(ctx, 0, nr, z.data()));
  NQ_TRY(apply_DDee(ctx, U64, 0, coef, s64, x64, t64, nr));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, -1.0, s64, 0, none, FIN_STORE0, S, nr));
  NQ_TRY(get_states(ctx, 0, nr));
  for (int i = 0; i < nr; i++) st[i].rel_res_sq = ctx->state_host[i].extra[0] / b2[i];
  if (rc != NQ_OK) return fail(rc, "exit without converging");
  return NQ_OK;
}

// Batched multi-shift CG (cgm_invert) for nr <= 16 sources; sol/ps laid out
// [source][shift][colour][site].  States: device CgmState[nr] in a ctx temp.
// With x64 (mixed precision): sol is the FP32 accumulator since the last reliable
// update and x64 the FP64 solution; b64/U64 give the FP64 residual of the update.
static int st_cgm_batched(nq_ctx* ctx, const LinkView& U, int nr, int nshift, const double* shifts, int niter,
                          const double* residues, const SpinorView& b, const SpinorView& sol, int sp,
                          std::vector<int>* iters, const SpinorView* x64 = nullptr, const SpinorView* b64 = nullptr,
                          const LinkView* U64 = nullptr, std::vector<int>* nrel = nullptr)
{
  const bool mixed = x64 != nullptr;
  const int64_t csz = 3 * ctx->geo.stride;
  SpinorView p, t, r, ps;
  NQ_TRY(alloc_spinor_nr(ctx, sp == SP_F64 ? "cgm.p" : "cgm.p32", sp, 3, nr, &p));
  NQ_TRY(alloc_spinor_nr(ctx, sp == SP_F64 ? "cgm.t" : "cgm.t32", sp, 3, nr, &t));
  NQ_TRY(alloc_spinor_nr(ctx, sp == SP_F64 ? "cgm.r" : "cgm.r32", sp, 3, nr, &r));
  NQ_TRY(alloc_spinor_nr(ctx, sp == SP_F64 ? "cgm.ps" : "cgm.ps32", sp, 3, nr * nshift, &ps));
  ps.rstride = csz * nshift;
  CgmState* dst = nullptr;
  CgmState* hst = nullptr;
  NQ_TRY(ctx_temp(ctx, "cgm.state", sizeof(CgmState) * 16, (void**)&dst));
  NQ_HIP(hipHostMalloc((void**)&hst, sizeof(CgmState) * 16, hipHostMallocDefault));
  struct HostFree {
    CgmState* p;
    ~HostFree() { (void)hipHostFree(p); }
  } hf{hst};
  SpinorView none;
  none.ncomp = 3;
  // |b|^2 per source (CgState scratch slots)
  std::vector<CgState> z(nr);
  NQ_TRY(put_states(ctx, 0, nr, z.data()));
  NQ_TRY(launch_lincomb(ctx, none, 1.0, b, 0, none, 0, none, FIN_STORE0, ctx->state, nr));
  NQ_TRY(get_states(ctx, 0, nr));
  for (int i = 0; i < nr; i++) {
    CgmState s = {};
    s.b2 = ctx->state_host[i].extra[0];
    if (s.b2 == 0 || std::isnan(s.b2)) return fail(NQ_ENAN, "invalid norm: %g", s.b2);
    s.rr = s.b2;
    s.betaa = 1;
    s.alpha = 0;
    s.nshift = nshift;
    s.nrun = nshift;
    s.niter = niter;
    s.mixed = mixed;
    s.rel_ref = 1.0;
    s.delta2 = 1e-4;
    if (const char* e = getenv("NQ_CGM_DELTA2")) s.delta2 = atof(e);
    for (int k = 0; k < nshift; k++) {
      s.shift[k] = shifts[k];
      s.target[k] = residues[k];
      s.zps[k] = s.zas[k] = 1;
      s.alphas[k] = 0;
      s.run[k] = 1;
    }
    hst[i] = s;
  }
  NQ_HIP(hipMemcpyAsync(dst, hst, sizeof(CgmState) * nr, hipMemcpyHostToDevice, ctx->stream));
  // sol_k = 0 ; ps_k = b ; p = r = b
  for (int k = 0; k < nshift; k++) {
    NQ_TRY(launch_zero(ctx, shifted(sol, k * csz), nr));
    NQ_TRY(launch_convert(ctx, shifted(ps, k * csz), b, nr));
  }
  NQ_TRY(launch_convert(ctx, p, b, nr));
  NQ_TRY(launch_convert(ctx, r, b, nr));
  SpinorView t64, s64;
  if (mixed) {
    for (int k = 0; k < nshift; k++) NQ_TRY(launch_zero(ctx, shifted(*x64, k * csz), nr));
    NQ_TRY(alloc_spinor(ctx, "cgm.t64", SP_F64, 3, &t64));
    NQ_TRY(alloc_spinor(ctx, "cgm.s64", SP_F64, 3, &s64));
  }
  // Chunks of 2^k iterations are captured once into HIP graphs and replayed: every
  // launch of an iteration has fixed arguments (this solve's vectors and states; the
  // scalars live on the device).  The first chunk runs eagerly, and so does every chunk
  // while the launch ledger is on; NQ_GRAPHS=0 switches the graphs off.
  struct Graphs {
    std::map<int, hipGraphExec_t> ex;
    ~Graphs()
    {
      for (auto& kv : ex) (void)hipGraphExecDestroy(kv.second);
    }
  } graphs;
  const char* genv = getenv("NQ_GRAPHS");
  const bool use_graphs = !(genv && genv[0] == '0');
  // NQ_CGM_PERSIST=1: each chunk of iterations runs as one cooperative launch
  // (launch_cgm_persistent; same bits); NQ_CGM_SPIN bounds its barrier waits
  const char* penv = getenv("NQ_CGM_PERSIST");
  const bool persist = penv && penv[0] == '1';
  unsigned spin_limit = 1u << 22;
  if (const char* e = getenv("NQ_CGM_SPIN")) spin_limit = (unsigned)strtoul(e, nullptr, 10);
  void* bar = nullptr;
  int* abort_host = nullptr;
  if (persist) {
    NQ_TRY(ctx_temp(ctx, "cgm.bar", (size_t)cgm_grid_bar_bytes(), &bar));
    NQ_HIP(hipHostMalloc((void**)&abort_host, sizeof(int), hipHostMallocDefault));
    *abort_host = 0;
  }
  struct HostFreeInt {
    int* p;
    ~HostFreeInt()
    {
      if (p) (void)hipHostFree(p);
    }
  } hfi{abort_host};
  int chunk = ctx->profiling ? 2 : 16, launched = 0;
  for (;;) {
    int nbusy = 0, nrun = 0;  // as of the last poll (the ledger's byte count)
    for (int i = 0; i < nr; i++)
      if (!hst[i].done) {
        nbusy++;
        for (int k = 0; k < nshift; k++) nrun += hst[i].run[k] != 0;
      }
    if (persist) {
      NQ_TRY(launch_cgm_persistent(ctx, U, p, t, r, sol, ps, csz, dst, nr, chunk, bar, spin_limit));
      NQ_HIP(hipMemcpyAsync(abort_host, (char*)bar + 2 * sizeof(unsigned), sizeof(int), hipMemcpyDeviceToHost,
                            ctx->stream));
    } else if (use_graphs && launched > 0 && !ctx->profiling) {
      for (int k = 7, left = chunk; k >= 0; k--) {
        const int n = 1 << k;
        for (; left >= n; left -= n) {
          auto it = graphs.ex.find(n);
          if (it == graphs.ex.end()) {
            hipGraph_t gr = nullptr;
            NQ_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            ctx->capturing = true;
            int rc = NQ_OK;
            for (int i = 0; i < n && rc == NQ_OK; i++)
              rc = launch_cgm_iteration(ctx, U, p, t, r, sol, ps, csz, dst, nr, nbusy, nrun);
            ctx->capturing = false;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
            if (rc != NQ_OK) {
              if (gr) (void)hipGraphDestroy(gr);
              return rc;
            }
            if (e != hipSuccess) return fail(NQ_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) return fail(NQ_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
            it = graphs.ex.emplace(n, ex).first;
          }
          NQ_HIP(hipGraphLaunch(it->second, ctx->stream));
        }
      }
    } else {
      for (int it = 0; it < chunk; it++)
        NQ_TRY(launch_cgm_iteration(ctx, U, p, t, r, sol, ps, csz, dst, nr, nbusy, nrun));
    }
    launched += chunk;
    NQ_HIP(hipMemcpyAsync(hst, dst, sizeof(CgmState) * nr, hipMemcpyDeviceToHost, ctx->stream));
    NQ_HIP(hipStreamSynchronize(ctx->stream));
    if (persist && *abort_host) return fail(NQ_EDEVICE, "persistent CGM: a grid barrier waited past its limit");
    // Shifts that stopped, and sources that ended, right after a C2 that left its
    // sol_k terms to the next one: apply those terms before anything reads sol.
    bool owes = false, flags = false;
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nshift; k++) {
        owes = owes || hst[i].owe[k] == 1;
        flags = flags || hst[i].owe[k] != 0;
      }
    if (owes) NQ_TRY(launch_cgm_settle(ctx, sol, ps, csz, dst, nr));
    if (flags) {
      for (int i = 0; i < nr; i++)
        for (int k = 0; k < nshift; k++) hst[i].owe[k] = 0;
      NQ_HIP(hipMemcpyAsync(dst, hst, sizeof(CgmState) * nr, hipMemcpyHostToDevice, ctx->stream));
      // the host edits hst below (reliable updates): the copy must have read it first
      NQ_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<int> upd;
    for (int i = 0; i < nr; i++)
      if (mixed && hst[i].reliable && !hst[i].done) upd.push_back(i);
    if (!upd.empty()) {
      // reliable update: flush the FP32 accumulators into FP64; replace the Krylov
      // residual by the FP64 residual of the hardest running shift kref,
      // r = (b - A(sigma_kref) x_kref) / zeta_kref, which it equals in exact arithmetic
      NQ_TRY(launch_cgm_flush(ctx, *x64, sol, csz, nshift, dst, nr, 1));
      std::vector<int> rep;
      for (int i : upd) {
        const int kr = hst[i].kref;
        const double zeta = hst[i].zas[kr];
        if (zeta < kMinZetaReplace) {
          // The Krylov residual is that of the unshifted base system.  When it lags
          // this far behind the hardest running shift (a near-zero mode of
          // -1/4 Deo Doe), the replacement would amplify the FP32 error of x_kref
          // by 1/zeta and the FP32 recursion breaks down (measured: |r|^2 growing
          // to overflow on 8x4x2x2).  The source leaves the multi-shift iteration
          // with its flushed FP64 solutions; its unconverged shifts go to the
          // per-shift refinement below.
          hst[i].done = 1;
          hst[i].reliable = 0;
          hst[i].nreliable += 1;
          continue;
        }
        rep.push_back(i);
        const SpinorView xk = shifted(*x64, (int64_t)i * x64->rstride + kr * csz);
        NQ_TRY(launch_st_hop(ctx, NQ_ODD, t64, xk, *U64, 1.0, 0.0, nullptr, none, 0, 1, 1, FIN_NONE, nullptr));
        NQ_TRY(launch_st_hop(ctx, NQ_EVEN, s64, t64, *U64, -0.25, shifts[kr], nullptr, xk, 0, 1, 1, FIN_NONE, nullptr));
        CgState z0 = {};
        NQ_TRY(put_states(ctx, i, 1, &z0));
        NQ_TRY(launch_lincomb(ctx, shifted(r, (int64_t)i * r.rstride), 1.0 / zeta, shifted(*b64, (int64_t)i * b64->rstride),
                              -1.0 / zeta, s64, 0, none, FIN_STORE0, ctx->state + i, 1));
      }
      NQ_TRY(get_states(ctx, 0, nr));
      for (int i : rep) {
        CgmState& h = hst[i];
        h.rr = ctx->state_host[i].extra[0];
        for (int k = 0; k < nshift; k++)
          if (h.run[k]) h.res[k] = h.rr * h.zas[k] * h.zas[k] / h.b2;
        h.rel_ref = h.res[h.kref];
        h.reliable = 0;
        h.nreliable += 1;
        if (std::isnan(h.rr)) return fail(NQ_ENAN, "nanned");
      }
      NQ_HIP(hipMemcpyAsync(dst, hst, sizeof(CgmState) * nr, hipMemcpyHostToDevice, ctx->stream));
    }
    bool all = true;
    int need = 4;
    for (int i = 0; i < nr; i++) {
      if (hst[i].nan) return fail(NQ_ENAN, "nanned");
      if (hst[i].done) continue;
      all = false;
      double worst = 0;
      for (int k = 0; k < nshift; k++)
        if (hst[i].run[k]) worst = std::max(worst, std::log(hst[i].res[k] / hst[i].target[k]));
      const double slope = hst[i].iter > 0 ? std::log(hst[i].rr / hst[i].b2) / hst[i].iter : -0.05;
      if (slope < 0) need = std::max(need, (int)std::ceil(-worst / slope));
    }
    if (all || launched > niter + 64) break;
    // with the launch ledger on, short chunks keep its count of running shifts (taken
    // at each poll) close to the shifts that actually run
    chunk = ctx->profiling ? 2 : std::max(4, std::min(128, need));
  }
  if (mixed) NQ_TRY(launch_cgm_flush(ctx, *x64, sol, csz, nshift, dst, nr, 0));
  if (iters) {
    iters->resize(nr);
    for (int i = 0; i < nr; i++) (*iters)[i] = hst[i].iter;
  }
  if (nrel) {
    nrel->resize(nr);
    for (int i = 0; i < nr; i++) (*nrel)[i] = hst[i].nreliable;
  }
  for (int i = 0; i < nr; i++)
    if (!hst[i].done || hst[i].iter >= niter) return fail(NQ_ENOCONV, "exit without converging");
  return NQ_OK;
}

}  // namespace nq

using namespace nq;

extern "C" {

int nq_st_2Doe(nq_ctx* c, const nq_gauge* g, nq_field* out, const nq_field* in)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_color(c, in, NQ_EVEN, "nq_st_2Doe"));
  NQ_TRY(check_color(c, out, NQ_ODD, "nq_st_2Doe"));
  Guard G(c);
  LinkView U;
  NQ_TRY(link_view(g, LK_F64_18, &U));
  SpinorView none;
  none.ncomp = 3;
  NQ_TRY(launch_st_hop(c, NQ_ODD, view_of(out, 0), view_of(in, 0), U, 1.0, 0.0, nullptr, none, 0, 1, 1, FIN_NONE, nullptr));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_st_Deo_half(nq_ctx* c, const nq_gauge* g, nq_field* out, const nq_field* in)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_color(c, in, NQ_ODD, "nq_st_Deo_half"));
  NQ_TRY(check_color(c, out, NQ_EVEN, "nq_st_Deo_half"));
  Guard G(c);
  LinkView U;
  NQ_TRY(link_view(g, LK_F64_18, &U));
  SpinorView none;
  none.ncomp = 3;
  NQ_TRY(launch_st_hop(c, NQ_EVEN, view_of(out, 0), view_of(in, 0), U, 0.25, 0.0, nullptr, none, 0, 1, 1, FIN_NONE,
                       nullptr));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_st_DDee(nq_ctx* c, const nq_gauge* g, double mass2, nq_field* out, const nq_field* in)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_color(c, in, NQ_EVEN, "nq_st_DDee"));
  NQ_TRY(check_color(c, out, NQ_EVEN, "nq_st_DDee"));
  if (out == in) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  LinkView U;
  NQ_TRY(link_view(g, LK_F64_18, &U));
  SpinorView t;
  NQ_TRY(alloc_spinor(c, "st.t", SP_F64, 3, &t));
  NQ_TRY(apply_DDee(c, U, mass2, nullptr, view_of(out, 0), view_of(in, 0), t, 1));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_st_cg(nq_ctx* c, const nq_gauge* g, double mass2, int niter, double residue, nq_field* sol,
             const nq_field* guess, const nq_field* src, nq_solve_stats* st)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_color(c, src, NQ_EVEN, "nq_st_cg"));
  NQ_TRY(check_color(c, sol, NQ_EVEN, "nq_st_cg"));
  if (guess) NQ_TRY(check_color(c, guess, NQ_EVEN, "nq_st_cg"));
  if (niter <= 0 || !(residue > 0)) return fail(NQ_EINVAL, "bad niter/residue");
  Guard G(c);
  nq_solve_stats local;
  if (!st) st = &local;
  std::memset(st, 0, sizeof *st);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, c->stream);
  if (guess && guess != sol) NQ_TRY(launch_convert(c, view_of(sol, 0), view_of(guess, 0)));
  const int rc = st_cg_batched(c, g, 1, &mass2, niter, &residue, NQ_PREC_F64, view_of(sol, 0), view_of(src, 0),
                               guess != nullptr, st);
  (void)hipEventRecord(b, c->stream);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  st->seconds = ms * 1e-3;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

int nq_st_cgm(nq_ctx* c, const nq_gauge* g, int nrhs, int nshift, const double* shifts, int niter,
              const double* residues, int prec, nq_field** sols, const nq_field* const* srcs, nq_solve_stats* st,
              double* true_res_sq)
{
  if (!c || !g || !shifts || !residues || !sols || !srcs) return fail(NQ_EINVAL, "null argument");
  if (nrhs < 1 || nshift < 1 || nshift > kMaxShift) return fail(NQ_EINVAL, "need 1..%d shifts and >= 1 source", kMaxShift);
  if (prec != NQ_PREC_F64 && prec != NQ_PREC_F32) return fail(NQ_EINVAL, "CGM supports NQ_PREC_F64 and NQ_PREC_F32");
  for (int i = 0; i < nrhs; i++) {
    NQ_TRY(check_color(c, srcs[i], NQ_EVEN, "nq_st_cgm"));
    for (int k = 0; k < nshift; k++) NQ_TRY(check_color(c, sols[i * nshift + k], NQ_EVEN, "nq_st_cgm"));
  }
  Guard G(c);
  hipEvent_t ea, eb;
  (void)hipEventCreate(&ea);
  (void)hipEventCreate(&eb);
  (void)hipEventRecord(ea, c->stream);
  LinkView U64, U32;
  NQ_TRY(link_view(g, LK_F64_18, &U64));
  const bool mixed = prec == NQ_PREC_F32;
  if (mixed) NQ_TRY(link_view(g, LK_F32_18, &U32));
  const int64_t csz = 3 * c->geo.stride;
  int rc = NQ_OK;
  for (int r0 = 0; r0 < nrhs; r0 += 16) {
    const int nr = std::min(16, nrhs - r0);
    SpinorView b, sol;
    NQ_TRY(alloc_spinor_nr(c, "cgm.b", SP_F64, 3, nr, &b));
    for (int i = 0; i < nr; i++) NQ_TRY(launch_convert(c, shifted(b, i * csz), view_of(srcs[r0 + i], 0)));
    const int sp = mixed ? SP_F32 : SP_F64;
    SpinorView bs = b;
    if (mixed) {
      NQ_TRY(alloc_spinor_nr(c, "cgm.b32", SP_F32, 3, nr, &bs));
      NQ_TRY(launch_convert(c, bs, b, nr));
    }
    NQ_TRY(alloc_spinor_nr(c, mixed ? "cgm.sol32" : "cgm.sol", sp, 3, nr * nshift, &sol));
    sol.rstride = csz * nshift;
    SpinorView x64 = sol;
    if (mixed) {
      NQ_TRY(alloc_spinor_nr(c, "cgm.x64", SP_F64, 3, nr * nshift, &x64));
      x64.rstride = csz * nshift;
    }
    std::vector<int> its, nrel;
    int rci = st_cgm_batched(c, mixed ? U32 : U64, nr, nshift, shifts, niter, residues, bs, sol, sp, &its,
                             mixed ? &x64 : nullptr, &b, &U64, &nrel);
    if (rci != NQ_OK && rci != NQ_ENOCONV) return rci;
    if (rci == NQ_ENOCONV) rc = rci;
    // copy out; FP64 solutions per (source, shift)
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nshift; k++)
        NQ_TRY(launch_convert(c, view_of(sols[(r0 + i) * nshift + k], 0), shifted(x64, i * csz * nshift + k * csz)));
    for (int i = 0; i < nr; i++)
      if (st) {
        st[r0 + i] = nq_solve_stats{};
        st[r0 + i].iters = its.empty() ? 0 : its[i];
        st[r0 + i].reliable_updates = nrel.empty() ? 0 : nrel[i];
      }
    // FP64 check |A(sigma) x - b|^2 / |b|^2 of every (source, shift) system, 16 at a
    // time; in mixed precision the systems that miss their residue are refined by a
    // batched FP32 CG with FP64 reliable updates (warm start) and checked again
    std::vector<double> tr((size_t)nr * nshift, 0.0);
    auto check = [&](const std::vector<int>& sys) -> int {
      for (size_t s0 = 0; s0 < sys.size(); s0 += 16) {
        const int ns = (int)std::min<size_t>(16, sys.size() - s0);
        SpinorView xb, bb, tt, ss;
        NQ_TRY(alloc_spinor_nr(c, "cgm.xb", SP_F64, 3, ns, &xb));
        NQ_TRY(alloc_spinor_nr(c, "cgm.bb", SP_F64, 3, ns, &bb));
        NQ_TRY(alloc_spinor_nr(c, "cgm.tt", SP_F64, 3, ns, &tt));
        NQ_TRY(alloc_spinor_nr(c, "cgm.ss", SP_F64, 3, ns, &ss));
        double* coef = nullptr;
        NQ_TRY(ctx_temp(c, "cgm.coef", sizeof(double) * 32, (void**)&coef));
        std::vector<double> hc(2 * ns);
        for (int j = 0; j < ns; j++) {
          const int q = sys[s0 + j], i = q / nshift, k = q % nshift;
          hc[2 * j] = -0.25;
          hc[2 * j + 1] = shifts[k];
          NQ_TRY(launch_convert(c, shifted(xb, j * csz), view_of(sols[(r0 + i) * nshift + k], 0)));
          NQ_TRY(launch_convert(c, shifted(bb, j * csz), view_of(srcs[r0 + i], 0)));
        }
        NQ_HIP(hipMemcpyAsync(coef, hc.data(), sizeof(double) * 2 * ns, hipMemcpyHostToDevice, c->stream));
        NQ_TRY(apply_DDee(c, U64, 0, coef, ss, xb, tt, ns));
        std::vector<CgState> z(ns);
        NQ_TRY(put_states(c, 0, ns, z.data()));
        SpinorView none;
        none.ncomp = 3;
        NQ_TRY(launch_lincomb(c, none, 1.0, bb, -1.0, ss, 0, none, FIN_STORE0, c->state, ns));
        NQ_TRY(launch_lincomb(c, none, 1.0, bb, 0, none, 0, none, FIN_STORE1, c->state, ns));
        NQ_TRY(get_states(c, 0, ns));
        for (int j = 0; j < ns; j++) tr[sys[s0 + j]] = c->state_host[j].extra[0] / c->state_host[j].extra[1];
      }
      return NQ_OK;
    };
    std::vector<int> all;
    for (int q = 0; q < nr * nshift; q++) all.push_back(q);
    NQ_TRY(check(all));
    if (mixed) {
      // failing systems in shift-major order, so groups hold neighbouring shifts
      std::vector<int> bad;
      for (int k = 0; k < nshift; k++)
        for (int i = 0; i < nr; i++)
          if (!(tr[i * nshift + k] < residues[k])) bad.push_back(i * nshift + k);
      for (size_t s0 = 0; s0 < bad.size(); s0 += 16) {
        const int ns = (int)std::min<size_t>(16, bad.size() - s0);
        SpinorView xb, bb;
        NQ_TRY(alloc_spinor_nr(c, "cgm.rxb", SP_F64, 3, ns, &xb));
        NQ_TRY(alloc_spinor_nr(c, "cgm.rbb", SP_F64, 3, ns, &bb));
        std::vector<double> sh(ns), rs(ns);
        for (int j = 0; j < ns; j++) {
          const int q = bad[s0 + j], i = q / nshift, k = q % nshift;
          sh[j] = shifts[k];
          rs[j] = residues[k];
          NQ_TRY(launch_convert(c, shifted(xb, j * csz), view_of(sols[(r0 + i) * nshift + k], 0)));
          NQ_TRY(launch_convert(c, shifted(bb, j * csz), view_of(srcs[r0 + i], 0)));
        }
        std::vector<nq_solve_stats> sst(ns);
        const int rr = st_cg_batched(c, g, ns, sh.data(), niter, rs.data(), NQ_PREC_F32, xb, bb, true, sst.data());
        if (rr != NQ_OK && rr != NQ_ENOCONV) return rr;
        if (rr == NQ_ENOCONV) rc = rr;
        for (int j = 0; j < ns; j++) {
          const int q = bad[s0 + j], i = q / nshift, k = q % nshift;
          NQ_TRY(launch_convert(c, view_of(sols[(r0 + i) * nshift + k], 0), shifted(xb, j * csz)));
          if (st) {
            st[r0 + i].reliable_updates += sst[j].reliable_updates;
            st[r0 + i].iters += sst[j].iters;
          }
        }
      }
      if (!bad.empty()) NQ_TRY(check(bad));
    }
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nshift; k++) {
        const double t = tr[i * nshift + k];
        if (true_res_sq) true_res_sq[(r0 + i) * nshift + k] = t;
        if (st) st[r0 + i].rel_res_sq = std::max(st[r0 + i].rel_res_sq, t);
      }
  }
  (void)hipEventRecord(eb, c->stream);
  (void)hipEventSynchronize(eb);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ea, eb);
  if (st)
    for (int i = 0; i < nrhs; i++) {
      st[i].seconds = ms * 1e-3;
      st[i].converged = rc == NQ_OK;
    }
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  if (rc != NQ_OK) return fail(rc, "exit without converging");
  return NQ_OK;
}

// evn_apply_stD / odd_apply_stD (and their _dag, sign = -1) and apply_stD
// (dirac_operator_stD.hpp:16-41,254-291): D = m + 1/2 hop on each parity
// (2 * apply_stDeo_half = 1/2 hop on the even sites, 1/2 apply_st2Doe on the odd).
// A null output skips that parity.
int nq_st_apply_D(nq_ctx* c, const nq_gauge* g, double m, int dag, nq_field* out_evn, nq_field* out_odd,
                  const nq_field* in_evn, const nq_field* in_odd)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  NQ_TRY(check_color(c, in_evn, NQ_EVEN, "nq_st_apply_D"));
  NQ_TRY(check_color(c, in_odd, NQ_ODD, "nq_st_apply_D"));
  if (out_evn) NQ_TRY(check_color(c, out_evn, NQ_EVEN, "nq_st_apply_D"));
  if (out_odd) NQ_TRY(check_color(c, out_odd, NQ_ODD, "nq_st_apply_D"));
  if (out_evn == in_evn || out_odd == in_odd) return fail(NQ_EINVAL, "out==in!");
  Guard G(c);
  LinkView U;
  NQ_TRY(link_view(g, LK_F64_18, &U));
  const double h = dag ? -0.5 : 0.5;
  if (out_evn)
    NQ_TRY(launch_st_hop(c, NQ_EVEN, view_of(out_evn, 0), view_of(in_odd, 0), U, h, m, nullptr, view_of(in_evn, 0), 0,
                         1, 1, FIN_NONE, nullptr));
  if (out_odd)
    NQ_TRY(launch_st_hop(c, NQ_ODD, view_of(out_odd, 0), view_of(in_evn, 0), U, h, m, nullptr, view_of(in_odd, 0), 0,
                         1, 1, FIN_NONE, nullptr));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

// inv_evn_stD_cg / inv_stD_cg (cg_invert_evn_stD.hpp:17-31, cg_invert_stD.hpp:14-88):
// sol_e = (m^2 - 1/4 Deo Doe)^-1 (D^+ src)_e by CG, then (full solve, sol_odd given)
// sol_o = (src_o - 1/2 hop_o sol_e) / m and the check |D sol - src|^2 / |src|^2.
int nq_st_solve_D(nq_ctx* c, const nq_gauge* g, double m, int niter, double residue, nq_field* sol_evn,
                  nq_field* sol_odd, const nq_field* guess_evn, const nq_field* src_evn, const nq_field* src_odd,
                  nq_solve_stats* st, double* check_rel_res_sq)
{
  if (!c || !g) return fail(NQ_EINVAL, "null argument");
  if (!(m != 0)) return fail(NQ_EINVAL, "mass must be non-zero");
  NQ_TRY(check_color(c, src_evn, NQ_EVEN, "nq_st_solve_D"));
  NQ_TRY(check_color(c, src_odd, NQ_ODD, "nq_st_solve_D"));
  NQ_TRY(check_color(c, sol_evn, NQ_EVEN, "nq_st_solve_D"));
  if (sol_odd) NQ_TRY(check_color(c, sol_odd, NQ_ODD, "nq_st_solve_D"));
  Guard G(c);
  LinkView U;
  NQ_TRY(link_view(g, LK_F64_18, &U));
  // temp = evn_apply_stD_dag(src) = m src_e - 1/2 hop_e src_o
  nq_field tmp;
  tmp.ctx = c;
  tmp.kind = NQ_COLOR;
  tmp.coverage = NQ_EVEN;
  SpinorView tv;
  NQ_TRY(alloc_spinor(c, "stD.tmp", SP_F64, 3, &tv));
  tmp.buf.p = tv.data;  // borrowed scratch, owned by the context
  NQ_TRY(launch_st_hop(c, NQ_EVEN, tv, view_of(src_odd, 0), U, -0.5, m, nullptr, view_of(src_evn, 0), 0, 1, 1,
                       FIN_NONE, nullptr));
  int rc = nq_st_cg(c, g, m * m, niter, residue, sol_evn, guess_evn, &tmp, st);
  if (rc != NQ_OK && rc != NQ_ENOCONV) return rc;
  if (sol_odd) {
    NQ_TRY(launch_st_hop(c, NQ_ODD, view_of(sol_odd, 0), view_of(sol_evn, 0), U, -0.5 / m, 1.0 / m, nullptr,
                         view_of(src_odd, 0), 0, 1, 1, FIN_NONE, nullptr));
    if (check_rel_res_sq) {
      // |D sol - src|^2 per parity: (m sol_p + 1/2 hop_p sol_!p) - src_p
      SpinorView none;
      none.ncomp = 3;
      SpinorView de, dd;
      NQ_TRY(alloc_spinor(c, "stD.de", SP_F64, 3, &de));
      NQ_TRY(alloc_spinor(c, "stD.do", SP_F64, 3, &dd));
      NQ_TRY(launch_st_hop(c, NQ_EVEN, de, view_of(sol_odd, 0), U, 0.5, m, nullptr, view_of(sol_evn, 0), 0, 1, 1,
                           FIN_NONE, nullptr));
      NQ_TRY(launch_st_hop(c, NQ_ODD, dd, view_of(sol_evn, 0), U, 0.5, m, nullptr, view_of(sol_odd, 0), 0, 1, 1,
                           FIN_NONE, nullptr));
      std::vector<CgState> z(2);
      NQ_TRY(put_states(c, 0, 2, z.data()));
      NQ_TRY(launch_lincomb(c, none, 1.0, de, -1.0, view_of(src_evn, 0), 0, none, FIN_STORE0, c->state));
      NQ_TRY(launch_lincomb(c, none, 1.0, dd, -1.0, view_of(src_odd, 0), 0, none, FIN_STORE1, c->state));
      NQ_TRY(launch_lincomb(c, none, 1.0, view_of(src_evn, 0), 0, none, 0, none, FIN_STORE0, c->state + 1));
      NQ_TRY(launch_lincomb(c, none, 1.0, view_of(src_odd, 0), 0, none, 0, none, FIN_STORE1, c->state + 1));
      NQ_TRY(get_states(c, 0, 2));
      *check_rel_res_sq = (c->state_host[0].extra[0] + c->state_host[0].extra[1]) /
                          (c->state_host[1].extra[0] + c->state_host[1].extra[1]);
    }
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  if (rc != NQ_OK) return fail(rc, "exit without converging");
  return NQ_OK;
}

// summ_the_rootst_eoimpr_quark_force (hmc/fermions/rootst_eoimpr_quark_force.cpp:17-85).
// g holds the links as the solver sees them (with background field and staggered
// phases, which nissa adds at the start of the routine); u1b the background field
// alone, whose product with eta_mu(x) is the phase of each link in the force.
// F (host, nissa eo AoS) is summed into; the TA projection and the product with U
// stay with the caller, as in the reference.
int nq_st_rootst_force(nq_ctx* c, const nq_gauge* g, const double* u1b_evn, const double* u1b_odd, int ndeg,
                       const double* poles, const double* weights, int niter, double residue, int prec,
                       const nq_field* pf, double* F_evn, double* F_odd, nq_solve_stats* st)
{
  if (!c || !g || !u1b_evn || !u1b_odd || !poles || !weights || !F_evn || !F_odd) return fail(NQ_EINVAL, "null argument");
  if (ndeg < 1 || ndeg > kMaxShift) return fail(NQ_EINVAL, "need 1..%d poles", kMaxShift);
  NQ_TRY(check_color(c, pf, NQ_EVEN, "nq_st_rootst_force"));
  Guard G(c);
  const int64_t Vh = c->geo.Vh;
  std::vector<nq_field*> chi(ndeg, nullptr);
  DevBuf u1[2], F[2];
  int rc = NQ_OK;
  for (int k = 0; k < ndeg && rc == NQ_OK; k++) rc = nq_field_create(c, NQ_COLOR, NQ_EVEN, &chi[k]);
  for (int p = 0; p < 2 && rc == NQ_OK; p++) rc = dev_alloc(u1[p], sizeof(double) * 8 * Vh);
  for (int p = 0; p < 2 && rc == NQ_OK; p++) rc = dev_alloc(F[p], sizeof(double) * 72 * Vh);
  auto cleanup = [&]() {
    for (auto* f : chi) nq_field_destroy(f);
    for (int p = 0; p < 2; p++) {
      dev_free(u1[p]);
      dev_free(F[p]);
    }
  };
  if (rc != NQ_OK) {
    cleanup();
    return rc;
  }
  // the pseudofermion's multi-shift solve on the poles
  std::vector<double> res(ndeg, residue), tr(ndeg);
  nq_solve_stats local;
  if (!st) st = &local;
  const nq_field* srcs[1] = {pf};
  rc = nq_st_cgm(c, g, 1, ndeg, poles, niter, res.data(), prec, chi.data(), srcs, st, tr.data());
  if (rc != NQ_OK && rc != NQ_ENOCONV) {
    cleanup();
    return rc;
  }
  const int rc_solve = rc;
  LinkView U;
  SpinorView v, none;
  none.ncomp = 3;
  rc = link_view(g, LK_F64_18, &U);
  if (rc == NQ_OK) rc = alloc_spinor(c, "force.v", SP_F64, 3, &v);
  const double* U1h[2] = {u1b_evn, u1b_odd};
  for (int p = 0; p < 2 && rc == NQ_OK; p++) {
    if (hipMemcpyAsync(u1[p].p, U1h[p], sizeof(double) * 8 * Vh, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(F[p].p, 0, sizeof(double) * 72 * Vh, c->stream) != hipSuccess)
      rc = fail(NQ_EDEVICE, "force buffers");
  }
  for (int k = 0; k < ndeg && rc == NQ_OK; k++) {
    // v_k = Doe chi_k = 1/2 (2 Doe) chi_k (apply_stDoe, dirac_operator_stD.hpp:80-92)
    rc = launch_st_hop(c, NQ_ODD, v, view_of(chi[k], 0), U, 0.5, 0.0, nullptr, none, 0, 1, 1, FIN_NONE, nullptr);
    if (rc == NQ_OK) rc = launch_st_force(c, NQ_EVEN, v, view_of(chi[k], 0), (const double*)u1[0].p, weights[k], (double*)F[0].p);
    if (rc == NQ_OK) rc = launch_st_force(c, NQ_ODD, view_of(chi[k], 0), v, (const double*)u1[1].p, -weights[k], (double*)F[1].p);
  }
  double* Fh[2] = {F_evn, F_odd};
  std::vector<double> h(72 * Vh);
  for (int p = 0; p < 2 && rc == NQ_OK; p++) {
    if (hipMemcpyAsync(h.data(), F[p].p, sizeof(double) * 72 * Vh, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = fail(NQ_EDEVICE, "force download");
      break;
    }
    for (int64_t i = 0; i < 72 * Vh; i++) Fh[p][i] += h[i];
  }
  cleanup();
  if (rc != NQ_OK) return rc;
  return rc_solve == NQ_ENOCONV ? fail(NQ_ENOCONV, "exit without converging") : NQ_OK;
}

int nq_st_cgm_rat_sum(nq_ctx* c, const nq_gauge* g, int nrhs, int degree, double cons, const double* poles,
                      const double* weights, int niter, double residue, int prec, nq_field** chis,
                      const nq_field* const* srcs, nq_solve_stats* st)
{
  if (!c || !g || !poles || !weights || !chis || !srcs) return fail(NQ_EINVAL, "null argument");
  if (nrhs < 1 || degree < 1 || degree > kMaxShift) return fail(NQ_EINVAL, "need 1..%d poles and >= 1 source", kMaxShift);
  for (int i = 0; i < nrhs; i++) {
    NQ_TRY(check_color(c, srcs[i], NQ_EVEN, "nq_st_cgm_rat_sum"));
    NQ_TRY(check_color(c, chis[i], NQ_EVEN, "nq_st_cgm_rat_sum"));
    for (int j = 0; j < nrhs; j++)
      if (chis[i] == srcs[j]) return fail(NQ_EINVAL, "chi aliases a source");
  }
  Guard G(c);
  // the CGM solutions of every (source, pole) live in context temporaries
  std::vector<nq_field> xs((size_t)nrhs * degree);
  std::vector<nq_field*> xp(xs.size());
  SpinorView xbuf;
  NQ_TRY(alloc_spinor_nr(c, "rat.x", SP_F64, 3, (int)xs.size(), &xbuf));
  for (size_t j = 0; j < xs.size(); j++) {
    xs[j].ctx = c;
    xs[j].kind = NQ_COLOR;
    xs[j].coverage = NQ_EVEN;
    xs[j].buf.p = (char*)xbuf.data + j * xbuf.rstride * 16;
    xp[j] = &xs[j];
  }
  struct Release {
    std::vector<nq_field>& v;
    ~Release()
    {
      for (auto& f : v) f.buf.p = nullptr;
    }
  } rel{xs};
  std::vector<double> res(degree, residue);
  const int rc = nq_st_cgm(c, g, nrhs, degree, poles, niter, res.data(), prec, xp.data(), srcs, st, nullptr);
  if (rc != NQ_OK && rc != NQ_ENOCONV) return rc;
  SpinorView none;
  none.ncomp = 3;
  for (int i = 0; i < nrhs; i++) {
    const SpinorView chi = view_of(chis[i], 0);
    NQ_TRY(launch_lincomb(c, chi, cons, view_of(srcs[i], 0), 0, none, 0, none, FIN_NONE, nullptr));
    for (int k = 0; k < degree; k++)
      NQ_TRY(launch_lincomb(c, chi, 1.0, chi, weights[k], view_of(xp[(size_t)i * degree + k], 0), 0, none, FIN_NONE,
                            nullptr));
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  if (rc != NQ_OK) return fail(rc, "exit without converging");
  return NQ_OK;
}

}  // extern "C"
