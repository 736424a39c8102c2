// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nissa_amd.h"

namespace nq {

// ---------------------------------------------------------------- errors
int fail(int code, const char* fmt, ...);
#define NQ_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return ::nq::fail(NQ_EDEVICE, "%s failed: %s (%s:%d)", #call,     \
                                            hipGetErrorString(e_), __FILE__, __LINE__);  \
  } while (0)
#define NQ_TRY(call)            \
  do {                          \
    int rc_ = (call);           \
    if (rc_ != NQ_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------- storage kinds
enum SpPrec { SP_F64 = 0, SP_F32 = 1, SP_F16 = 2 };
enum LkPrec { LK_F64_18 = 0, LK_F32_18 = 1, LK_F32_12 = 2, LK_F16_12 = 3, LK_F64_12 = 4 };
// complex elements stored per link (12-real: rows 0 and 1) and bytes per element
constexpr int lk_nc(int lk) { return (lk == LK_F64_18 || lk == LK_F32_18) ? 9 : 6; }
constexpr int lk_eb(int lk) { return (lk == LK_F64_18 || lk == LK_F64_12) ? 16 : (lk == LK_F16_12 ? 4 : 8); }

// Device geometry of one lattice (all extents even).
// Device site order: t is split as t = th*TB + tl (TB in {2,4,8} dividing T) and
// sites are ordered lexicographically in (th, x, y, z, tl), tl fastest; within a
// parity the device index is that order's rank >> 1 (TB even).  Compared with
// nissa's (t, x, y, z) order this brings the most frequent neighbour distance
// down from LX*LY*LZ/2 to TB*LY*LZ/2 sites, so neighbour spinors are re-read from
// L2 instead of the fabric.  Host buffers keep nissa's order; the conversion
// kernels translate.
// Device site order: (th, xt, yt, xi, yi, z, tl) from slowest to fastest, with
// t = th*TB + tl, x = xt*TX + xi, y = yt*TY + yi; the eo index of a site is its rank
// in this order divided by two.  Tiling t keeps the +-t neighbours of a wavefront in
// the same rows; tiling x and y in TX x TY patches keeps the +-x/+-y neighbour rows
// close in processing order, so they are still in L2 when they are needed.
struct Geo {
  int T, LX, LY, LZ;
  int TB, NTH;         // t tile (even, divides T), number of tiles
  int TX, logTX, NXT;  // x tile, log2, number of tiles
  int TY, logTY, NYT;  // y tile, log2, number of tiles
  int64_t Vh;          // sites per parity
  int64_t stride;      // per-component stride of SoA fields (Vh rounded up to 64)
  int64_t Sz, Syi, Sxi, Syt, Sxt, Sth;  // device-order strides (tl stride is 1)
  // visit order of the register-blocked (multi-RHS) kernels: rb_bpc 256-site blocks
  // per (th, x, y) column and rb_ncol columns, visited z slab by z slab
  // (map_thread_rb); rb_bpc = 0 visits the blocks in storage order
  int64_t rb_bpc, rb_ncol;
};

// rank of lattice site (t, x, y, z) in the device order (both parities)
__host__ __device__ inline int64_t device_rank(const Geo& g, int t, int x, int y, int z)
{
  return (int64_t)(t / g.TB) * g.Sth + (int64_t)(x >> g.logTX) * g.Sxt + (int64_t)(y >> g.logTY) * g.Syt +
         (int64_t)(x & (g.TX - 1)) * g.Sxi + (int64_t)(y & (g.TY - 1)) * g.Syi + (int64_t)z * g.Sz + (t % g.TB);
}

// One parity of a spinor (ncomp=12) or colour (ncomp=3) field in SoA layout:
// component k of site e at data[k*stride + e] (complex element).  F16 keeps a
// per-site float norm; the stored halves are value/norm.
struct SpinorView {
  void* data = nullptr;
  float* norm = nullptr;   // F16 only: norm[r*stride + e]
  int prec = SP_F64;
  int ncomp = 12;
  int64_t rstride = 0;     // elements between consecutive right-hand sides
};

// Gauge links as seen from output sites of one parity: 8 slots per site,
// slot 2mu = U_mu(x), slot 2mu+1 = U_mu(x-mu)^+ (already daggered), each a 3x3
// complex matrix stored as 9 (18-real) or 6 (12-real: rows 0,1) complex SoA
// components: element (slot*nc + k)*stride + e.  For 12-real, row 2 is
// det * conj(row0 x row1) with det a per-slot constant phase.
struct LinkView {
  const void* data[2] = {nullptr, nullptr};  // indexed by output parity
  int prec = LK_F64_18;
  double det[2][8][2] = {};                  // [out parity][slot][re,im]
};

// Solver state living in device memory; updated by the finalisation kernel
// that follows each reducing launch.
struct CgState {
  double b2, target;
  double rr, rr_new, pAp, alpha, beta;
  double r0, delta;     // reliable-update reference |r| and threshold factor
  double true_rr;       // last FP64 residual
  double extra[6];
  double alpha_prev;    // alpha of the previous iteration (paired x updates)
  int iter, niter, done, reliable, nreliable, nan, mixed, flags;
  // paired x updates of the one-source mixed CG: xlag = the last p update left its
  // alpha p term to the next one; xstep = the coming p update applies both terms
  int xlag, xstep;
};

constexpr int kMaxShift = 16;

// Multi-shift CG state of one source (cgm_invert recurrences); lives in device
// memory, updated by the CGM finalisation kernels.
// Mixed precision (mixed = 1): the sloppy solutions are flushed into FP64 and the
// Krylov residual is replaced by the FP64 residual of the hardest running shift
// (kref, the smallest running shift) whenever that shift's squared residual has
// dropped by delta2 since the last such update.  reliable = 1: flagged at the end of an iteration
// whose p/ps update still runs; 2: idle until the host's update clears it.
struct CgmState {
  double b2, rr, betaa, betap, alpha, rfrf, pap, rel_ref, delta2, pad0;
  double shift[kMaxShift], target[kMaxShift], zps[kMaxShift], zas[kMaxShift], zfs[kMaxShift];
  double betas[kMaxShift], alphas[kMaxShift], res[kMaxShift];
  double betas_prev[kMaxShift];  // betas_k of the previous iteration (paired sol_k updates)
  int run[kMaxShift];
  // owe_k: the sol_k term betas_k ps_k of the last hop2 is still to be applied,
  // 1 by the host's settle kernel (the shift stopped or the source ended), 2 by
  // the update kernel (a reliable update follows)
  int owe[kMaxShift];
  int done, reliable;  // adjacent: the sloppy kernels skip a source when either is set
  int nshift, nrun, iter, niter, nan, mixed, kref, nreliable;
  // pend: the ps_k update of the iteration just finalised is left to the next
  // iteration's hop2 kernel (set by the rfrf finalisation when the source goes on
  // without a reliable update; the update kernel then updates p only)
  // lag: the last hop2 left its sol_k terms to the next one, which applies both
  // iterations' terms in order (so sol_k is read and written every other iteration)
  int pend, lag;
};

enum Fin {
  FIN_NONE = 0,
  FIN_STORE0 = 1,     // extra[0] = total
  FIN_PAP = 2,        // pAp = total; alpha = rr/pAp
  FIN_RR = 3,         // plain CG: beta, rr, iter, done
  FIN_RR_SLOPPY = 4,  // mixed CG: decide reliable update
  FIN_RR_TRUE = 5,    // mixed CG: after FP64 recomputation
  FIN_RR_INIT = 6,    // initial residual: rr = total, r0 = sqrt
  FIN_STORE1 = 7,     // extra[1] = total
  FIN_RR_TRUE_HOLD = 8,  // as FIN_RR_TRUE, but `reliable` stays raised (batched updates)
  FIN_INIT_B2 = 9,    // start of a solve on r = b: b2 = rr = total, target *= total
  FIN_PARTIAL = 10,   // leave the per-chunk partials for a caller's own finalisation
};

// Hop epilogue: out_h = c1_h*hop + c2_h*aux + c3_h*aux2 (h = upper/lower spins).
// mode 0: write out; 1: write out (if write) and reduce |out|^2;
// 2: CG tail (s=out, x+=alpha p, r-=alpha s, reduce |r|^2);
// 3: residual tail (out written to r64 = out, sloppy copy to rs, reduce |out|^2).
struct HopEpi {
  int mode = 0;
  int write = 1;
  int fin = FIN_NONE;
  int pred_reliable = 0;   // 0: skip when done; 1: run only when state->reliable;
                           // 2: run only when !state->reliable (sloppy iteration)
  int nrhs = 1;            // right-hand sides in the launch (views carry rstride)
  int batched = 0;         // use the multi-RHS kernel build even for nrhs == 1
  // c1, c2, c3 must stay contiguous: the kernel reads them as double[12]
  double c1[2][2] = {{1, 0}, {1, 0}};
  double c2[2][2] = {{0, 0}, {0, 0}};
  double c3[2][2] = {{0, 0}, {0, 0}};
  const double* coef = nullptr;  // device double[nrhs][12] overriding c1,c2,c3 per rhs
  SpinorView aux, aux2;
  SpinorView x, p, r, rs;
  CgState* st = nullptr;
  // finalisation of the previous launch's reduction (partials left with FIN_PARTIAL)
  // carried out by one extra workgroup of this launch (plain one-RHS hop, mode 0;
  // other kernels launch it separately): fold_fin on st with predicate fold_pred
  int fold_fin = FIN_NONE;
  int fold_pred = 0;
  // twisted clover blocks in the epilogue (clover_block.hpp): blk_mode 1 applies the
  // inverse blocks blk (72 complex per site) to the hop before c1; blk_mode 2 applies
  // the blocks z + clover term (blk: 36 complex per site, z = 1/(2 blk_kappa) +- i
  // blk_mu) to aux before c2; blk_dag selects the dagger
  SpinorView blk;
  int blk_mode = 0;
  int blk_dag = 0;
  double blk_kappa = 0, blk_mu = 0;
};

// ---------------------------------------------------------------- objects
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct KernelStat {
  int64_t launches = 0;
  double total_ms = 0;
  double total_bytes = 0;  // algorithmic bytes of all launches (the mean per launch is reported)
};

}  // namespace nq

struct nq_ctx {
  int device = 0;
  int ncu = 0;  // compute units of the device (persistent launches)
  nq::Geo geo{};
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  // reduction scratch
  double* partial = nullptr;  // per-(rhs, 64-site chunk) partials
  int64_t max_partials = 0;
  nq::CgState* state = nullptr;       // device, array of nstate
  nq::CgState* state_host = nullptr;  // pinned mirror
  int nstate = 0;
  // profiling ledger
  bool profiling = false;
  std::map<std::string, nq::KernelStat> stats;
  struct Pending { std::string name; hipEvent_t a, b; double bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t hop_ev[2] = {nullptr, nullptr};  // events of the hop launch in flight (ledger on)
  // cached temporaries keyed by role
  std::map<std::string, nq::DevBuf> temps;
  // autotuned launch choices (key: kernel family and shape), e.g. right-hand sides
  // per wavefront of the batched hop; filled on first use
  std::map<std::string, int> tuned;
  bool capturing = false;  // a HIP graph capture is open on the stream: no tuning, no ledger
};

struct nq_field {
  nq_ctx* ctx = nullptr;
  int kind = NQ_SPINCOLOR;  // NQ_SPINCOLOR / NQ_COLOR
  int coverage = NQ_EVEN;   // NQ_EVEN / NQ_ODD / NQ_LX
  nq::DevBuf buf;           // FP64 SoA; LX: even half then odd half
  int ncomp() const { return kind == NQ_SPINCOLOR ? 12 : 3; }
};

// clover term of a gauge field (FP64, 36 complex per site and parity) and the
// inverse twisted clover blocks (72 complex per site), cached per parity for the
// last (kappa, mu)
struct nq_clover {
  nq_ctx* ctx = nullptr;
  double csw = 0;
  nq::DevBuf cl[2];
  nq::DevBuf inv[2];
  double inv_kappa[2] = {0, 0}, inv_mu[2] = {0, 0};
  int inv_valid[2] = {0, 0};
  // packed FP32 ([0]) and FP16 ([1]) copies for sloppy solves (the inverse copies
  // follow inv)
  nq::DevBuf cls[2][2], invs[2][2];
  int cls_valid[2][2] = {{0, 0}, {0, 0}}, invs_valid[2][2] = {{0, 0}, {0, 0}};
};

struct nq_gauge {
  nq_ctx* ctx = nullptr;
  nq::DevBuf aos[2];                 // eo AoS FP64 copy (source for derived layouts)
  nq::DevBuf l64;                    // LK_F64_18, both output parities
  nq::DevBuf l32;                    // LK_F32_18
  nq::DevBuf l32c;                   // LK_F32_12
  nq::DevBuf l16c;                   // LK_F16_12
  nq::DevBuf l64c;                   // LK_F64_12
  bool compressible = false;         // per-direction constant det
  double det[4][2] = {};             // det(U_mu) when compressible
  // sampled fingerprints of the host conf the handle was made from (nq_gauge_matches_*)
  uint64_t fp_eo[2] = {0, 0};
  uint64_t fp_lx = 0;
  bool from_lx = false;
  int64_t bytes() const;
};

namespace nq {

// ---------------------------------------------------------------- memory
int dev_alloc(DevBuf& b, size_t bytes);
void dev_free(DevBuf& b);
int ctx_temp(nq_ctx* ctx, const std::string& role, size_t bytes, void** out);
int ctx_temp_release(nq_ctx* ctx, const std::string& prefix);
// sampled fingerprint of a host buffer of doubles (nq_host_conf_fingerprint)
uint64_t conf_fingerprint(const double* p, int64_t n);
int ensure_state(nq_ctx* ctx, int n);

// views
SpinorView view_of(const nq_field* f, int parity);
SpinorView view_lx_flat(const nq_field* f);  // lx field as 2*ncomp components over Vh sites
int spinor_bytes(int prec, int ncomp, const Geo& g, size_t* data_bytes, size_t* norm_bytes);
int alloc_spinor(nq_ctx* ctx, const std::string& role, int prec, int ncomp, SpinorView* v);
int alloc_spinor_nr(nq_ctx* ctx, const std::string& role, int prec, int ncomp, int nr, SpinorView* v);

// gauge
int gauge_build(nq_gauge* g, int lk);
int link_view(const nq_gauge* g, int lk, LinkView* v);

// ---------------------------------------------------------------- profiling
void prof_begin(nq_ctx* ctx, const char* name, double bytes, hipEvent_t* a);
void prof_end(nq_ctx* ctx, hipEvent_t a);
// ledger entry whose two events the caller hands to hipExtLaunchKernelGGL, which
// records the kernel's own start and end (as a profiler's kernel trace sees them);
// both stay null when the ledger is off.  prof_end_ext closes the entry.
void prof_begin_ext(nq_ctx* ctx, const char* name, double bytes, hipEvent_t* a, hipEvent_t* b);
void prof_end_ext(nq_ctx* ctx);
void prof_flush(nq_ctx* ctx);

// ---------------------------------------------------------------- launchers (device code)
int64_t blocks_for(int64_t n);
constexpr int kBlock = 256;
// launch shape of the (64-site chunk x right-hand side) mapping
// Launch shape of the field kernels: a workgroup covers W chunks of 64 sites for
// NRB right-hand sides (W*NRB <= 4 wavefronts); the NR right-hand sides of a launch
// are split into NRG groups of NRB, the groups of one chunk range in adjacent
// workgroups so that they read the same links close in time.
struct Shape {
  int W, NRB, NRG, NR;
};
struct Launch {
  Shape s;
  int64_t nchunks, blocks;
  int threads;
};
Launch launch_shape(const Geo& g, int nr);

// twisted-mass hop of output parity out_par with epilogue
int launch_tm_hop(nq_ctx* ctx, int out_par, const SpinorView& out, const SpinorView& in,
                  const LinkView& U, double anis, const HopEpi& epi);
double tm_hop_bytes(int sp, int lk, const HopEpi& epi, int ncomp = 12);

// staggered hop of output parity op: out = c1*hop + c2*aux (per-rhs (c1,c2) from
// coef_dev when given). mode 0 write; 1 |out|^2; 4 Re<aux,out> (reduced into st with fin);
// 6 out = c1*hop and sigma_r |in|^2 + |out|^2/4 with sigma_r = coef_dev[2r+1].
// skip_flags (optional): rhs r is skipped when any of skip_flags[r*skip_stride + i],
// i < skip_n, is set (skip_n = 2 on &CgState::done also skips while a reliable update
// is due, and the finalisation then leaves such states alone).
int launch_st_hop(nq_ctx* ctx, int op, const SpinorView& out, const SpinorView& in, const LinkView& U, double c1,
                  double c2, const double* coef_dev, const SpinorView& aux, int mode, int write, int nr, int fin,
                  CgState* st, const int* skip_flags = nullptr, int skip_stride = 0, int skip_n = 1);
// sol64 += sol ; sol = 0 for every shift of the sources (only those waiting for a
// reliable update when only_reliable)
int launch_cgm_flush(nq_ctx* ctx, const SpinorView& sol64, const SpinorView& sol, int64_t sstride, int nshift,
                     void* states, int nr, int only_reliable);
// generic multi-shift CG pieces for lx spincolor fields (one source, FP64):
// finalisation of the reduction just launched (fin 1: <p,A p> = pap_scale * sum,
// 2: |r|^2), v2: sol_k -= betas_k ps_k, r += betaa s, |r|^2 (+ finalisation),
// v3: p = r + alpha p, ps_k = zfs_k r + alphas_k ps_k
int launch_cgm_finalize(nq_ctx* ctx, void* states, int nr, int fin, double pap_scale);
int launch_cgm_vec2(nq_ctx* ctx, const SpinorView& s, const SpinorView& r, const SpinorView& sol, const SpinorView& ps,
                    int64_t sstride, void* states);
int launch_cgm_vec3(nq_ctx* ctx, const SpinorView& r, const SpinorView& p, const SpinorView& ps, int64_t sstride,
                    void* states);
// one multi-shift CG iteration for nr sources (states: device CgmState[nr]);
// sol/ps hold the shifts of each source at element offset k*sstride; nbusy and nrun
// (sources still iterating and their running shifts, as the host last saw them) only
// size the ledger's byte count
// batched staggered CG: s = coef[2r] hop(t) + coef[2r+1] p at the even sites fused with
// x += alpha p, r -= alpha s and |r|^2 (finalised with fin; done or reliable-pending
// systems idle)
int launch_st_cgtail(nq_ctx* ctx, const SpinorView& t, const LinkView& U, const double* coef_dev, const SpinorView& p,
                     const SpinorView& x, const SpinorView& r, CgState* st, int fin, int nr);
// niters multi-shift iterations of nr sources in one cooperative launch (the same
// block code and bits as launch_cgm_iteration); bar: device scratch of
// cgm_grid_bar_bytes() whose `abort` word (third int) is nonzero after the launch when
// a grid barrier waited more than spin_limit polls
int launch_cgm_persistent(nq_ctx* ctx, const LinkView& U, const SpinorView& p, const SpinorView& t,
                          const SpinorView& r, const SpinorView& sol, const SpinorView& ps, int64_t sstride,
                          void* states, int nr, int niters, void* bar, unsigned spin_limit);
int cgm_grid_bar_bytes();
// sol_k -= betas_k ps_k for the (source, shift) pairs whose CgmState owe_k is 1
int launch_cgm_settle(nq_ctx* ctx, const SpinorView& sol, const SpinorView& ps, int64_t sstride, void* states, int nr);
int launch_cgm_iteration(nq_ctx* ctx, const LinkView& U, const SpinorView& p, const SpinorView& t,
                         const SpinorView& r, const SpinorView& sol, const SpinorView& ps, int64_t sstride,
                         void* states, int nr, int nbusy, int nrun);

// layout conversions
// host buffer in nissa order (host_lx: lx-indexed buffer, else eo-indexed) <->
// device parity `par` of a field
// rooted-staggered force of one pole on the links of parity par (st_kernels.hip)
int launch_st_force(nq_ctx* ctx, int par, const SpinorView& a, const SpinorView& b, const double* u1b, double wsign,
                    double* F);
// twisted clover (clover.hip): mode 0 = (z + C) per chirality, 1 = inverse blocks;
// out = a_h (that) + b_h aux
int launch_clover_build(nq_ctx* ctx, const double* aos_e, const double* aos_o, int par, double csw,
                        const SpinorView& out);
int launch_clover_invert(nq_ctx* ctx, const SpinorView& cl, const SpinorView& inv, double kappa, double mass);
// Re tr sums of the temporal (ts[0]) and spatial (ts[1]) plaquettes over the lattice
int launch_plaquette(nq_ctx* ctx, const double* aos_e, const double* aos_o, double ts[2]);
int launch_blocks_sloppy(nq_ctx* ctx, int ncomp, const SpinorView& in, const SpinorView& out);
// the block (mode 0: z + clover term, 1: inverse) of every site applied to nr right-hand
// sides (views with rstride), skipping those whose state st[r] is idle (pred as for the
// hop epilogues; st may be null)
int launch_clover_apply_batched(nq_ctx* ctx, int mode, int dag, const SpinorView& out, const SpinorView& in,
                                const SpinorView& blk, double kappa, double mu, int nr, const CgState* st, int pred);
int launch_clover_apply(nq_ctx* ctx, int mode, const SpinorView& out, const SpinorView& in, const SpinorView& blk,
                        const SpinorView& aux, double kappa, double mu, int dag, const double a[2][2],
                        const double b[2][2]);
int launch_aos_to_soa(nq_ctx* ctx, const double* aos, int ncomp, int par, int host_lx, const SpinorView& out);
int launch_soa_to_aos(nq_ctx* ctx, const SpinorView& in, int ncomp, int par, int host_lx, double* aos);
// device site order of a lattice (t tile up to tbmax, x/y patch tx x ty)
Geo make_geo(int T, int LX, int LY, int LZ, int tbmax, int tx, int ty);
// choose the site-order tiles of a fresh context (no fields yet) by timing the hops
// on synthetic data, or take them from the tile cache; sets ctx->geo
int tune_site_tiles(nq_ctx* ctx);
int launch_build_links(nq_ctx* ctx, const double* aos_e, const double* aos_o, int lk, void* out,
                       const double det[4][2]);
int launch_lx_split_gauge(nq_ctx* ctx, const double* lx, double* evn, double* odd);

// BLAS-1 (all fused; operands FP64 or sloppy; nr right-hand sides per launch
// sharing the views' rstride; st points at nr consecutive states).
// pred: 0 always run, 1 skip right-hand sides that are done, 2 run only when reliable,
// 3 skip when done or reliable.
int launch_axpby(nq_ctx* ctx, const SpinorView& y, const double a[2], const SpinorView& x,
                 const double b[2], int fin, CgState* st, int pred, int nr = 1);
int launch_lincomb(nq_ctx* ctx, const SpinorView& out, double a, const SpinorView& x, double b,
                   const SpinorView& y, double c, const SpinorView& z, int fin, CgState* st,
                   int nr = 1, int pred = 0);
int launch_dot(nq_ctx* ctx, const SpinorView& a, const SpinorView& b, int fin, CgState* st,
               int nr = 1, int pred = 0);
// p = r + beta p; pred 1: skip done or reliable-pending right-hand sides; 2: only the
// reliable-pending ones (after FIN_RR_TRUE_HOLD)
int launch_cg_pupdate(nq_ctx* ctx, const SpinorView& p, const SpinorView& r, CgState* st, int nr = 1, int pred = 1,
                      const SpinorView& x = SpinorView());
// state[i].reliable = 0 for i < nr
int launch_clear_reliable(nq_ctx* ctx, CgState* st, int nr);
// x += alpha p ; r -= alpha s ; |r|^2 -> fin  (alpha from state, skipped when done or
// while a reliable update is due)
int launch_cg_update(nq_ctx* ctx, const SpinorView& x, const SpinorView& p, const SpinorView& r,
                     const SpinorView& s, CgState* st, int fin, int nr = 1);
int launch_reliable_accum(nq_ctx* ctx, const SpinorView& x64, const SpinorView& xs, CgState* st, int nr = 1,
                          const SpinorView& p = SpinorView(), const SpinorView& p_prev = SpinorView());
// p_out = r + beta p_in (one right-hand side, p_in and p_out distinct; skipped when done or
// reliable-pending); with x, when the state's xstep is set, first
// x += alpha_prev p_prev + alpha p_in, p_prev being p_out's old content
int launch_cg_pupdate_pair(nq_ctx* ctx, const SpinorView& p_in, const SpinorView& p_out, const SpinorView& r,
                           CgState* st, const SpinorView& x = SpinorView());
int launch_convert(nq_ctx* ctx, const SpinorView& out, const SpinorView& in, int nr = 1,
                   CgState* st = nullptr, int pred = 0);
int launch_zero(nq_ctx* ctx, const SpinorView& v, int nr = 1);
// out_h = a_h x + b_h y on the upper/lower spin halves (y optional)
int launch_spin_diag(nq_ctx* ctx, const SpinorView& out, const double a[2][2], const SpinorView& x,
                     const double b[2][2], const SpinorView& y, int nr = 1);
int reduce_norm2(nq_ctx* ctx, const SpinorView& v, double* out);
// sum the per-chunk partials of the last reducing launch into state[0..nr) with fin;
// pred_reliable as in HopEpi (the same condition the reducing launch used)
int launch_finalize(nq_ctx* ctx, int nr, CgState* st, int fin, int pred_reliable);
int reduce_dot(nq_ctx* ctx, const SpinorView& a, const SpinorView& b, double* out);
}  // namespace nq
