/*
 * nissa_amd — MI355X (gfx950) native fermion-solve path for nissa.
 *
 * C ABI: plain pointers, sizes and int status codes.  Host buffers use nissa's
 * CPU ("AoS") layout (src/base/field.hpp:408-416):
 *   spincolor double[4][3][2], color double[3][2], quad_su3 double[4][3][3][2];
 *   lx sites t-major / z-fastest, eo sites = lx>>1 within each parity
 *   (src/geometry/geometry_lx.cpp:51-60, src/geometry/geometry_eo.cpp:67-77).
 * Gauge links are passed with every phase already applied (boundary
 * conditions, staggered eta phases), as nissa's callers do before calling
 * the operators (src/operations/gaugeconf.cpp:125-182, src/hmc/backfield.cpp:230-245).
 *
 * Every entry point returns NQ_OK (0) or an error code; nq_last_error() gives
 * the message of the last failure on the calling thread.  A context owns one
 * HIP stream on one device; calls on one context are serialised, distinct
 * contexts may be used from different threads.
 */
#ifndef NISSA_AMD_H
#define NISSA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (nissa CRASHes on these conditions, src/base/debug.hpp:25) */
#define NQ_OK 0
#define NQ_EINVAL 1   /* bad argument / shape mismatch            */
#define NQ_ENOCONV 2  /* "exit without converging"  (cg_invert_template_threaded.hpp:146-147) */
#define NQ_ENAN 3     /* "nanned" / "invalid norm" (cg_invert_template_threaded.hpp:55,113)   */
#define NQ_EOOM 4     /* device allocation failed                 */
#define NQ_EDEVICE 5  /* HIP runtime error / no device            */
#define NQ_EBUSY 6    /* admission refused (serving)              */

/* field coverage (nissa FieldCoverage, src/base/field.hpp:106) */
#define NQ_EVEN 0
#define NQ_ODD 1
#define NQ_LX 2
/* field kind */
#define NQ_SPINCOLOR 0
#define NQ_COLOR 1
/* solver precision of the inner Krylov iteration */
#define NQ_PREC_F64 0      /* plain FP64 CG                                          */
#define NQ_PREC_F32 1      /* FP32 spinors + FP32 18-real links, FP64 reliable updates */
#define NQ_PREC_F16 2      /* FP16 spinors + per-site FP32 norm, FP32 12-real links  */
#define NQ_PREC_F16H 3     /* FP16 spinors + per-site norm, FP16 12-real links       */

typedef struct nq_ctx nq_ctx;
typedef struct nq_gauge nq_gauge;
typedef struct nq_field nq_field;
typedef struct nq_clover nq_clover;

typedef struct nq_solve_stats {
  int iters;             /* Krylov iterations (sloppy ones for mixed precision)      */
  int reliable_updates;  /* FP64 true-residual recomputations inside the loop       */
  int converged;         /* 1 when the FP64 true residual met the request           */
  int reserved;
  double rel_res_sq;     /* |b - A x|^2 / |b|^2 recomputed in FP64 after the loop   */
  double seconds;        /* device time of the solve (HIP events on the ctx stream) */
} nq_solve_stats;

/* ------------------------------------------------------------------ context */
const char* nq_last_error(void);
int nq_device_count(int* n);
/* replaces nissa's initGrid(T,L) geometry set-up on one rank (src/base/init.cpp:563).
   From 2^18 sites per parity up, the device site order (t tile, x/y patch) is chosen by
   timing the hops on synthetic data at creation, or read from the tile cache that run
   wrote (NQ_TUNE_CACHE, else $HOME/.cache/nissa_amd/site_tiles.txt), as nissa's kernel
   tuner keeps its table on disk (src/threads/benchmarks.hpp:118-200,207-328).
   NQ_TTILE / NQ_XYTILE fix the tiles, NQ_TUNE_TILES=0 keeps the defaults.  The order
   changes global sums at the rounding level only. */
int nq_ctx_create(int T, int LX, int LY, int LZ, int device, nq_ctx** out);
void nq_ctx_destroy(nq_ctx* ctx);
int nq_ctx_sync(nq_ctx* ctx);
/* the context's site-order tiles: t tile, x patch, y patch */
int nq_ctx_site_tiles(const nq_ctx* ctx, int* tb, int* tx, int* ty);
/* per-kernel timing ledger (HIP events around every launch when enabled) */
int nq_ctx_set_profiling(nq_ctx* ctx, int on);
int nq_ctx_kernel_count(nq_ctx* ctx, int* n);
int nq_ctx_kernel_stat(nq_ctx* ctx, int i, char* name, int name_len, int64_t* launches,
                       double* total_ms, double* bytes_per_launch);
int nq_ctx_reset_kernel_stats(nq_ctx* ctx);

/* ------------------------------------------------------------------ gauge */
/* EoField<quad_su3> as two host buffers of Vh quad_su3 (src/base/field.hpp:1586-1599) */
int nq_gauge_create_eo(nq_ctx* ctx, const double* conf_evn, const double* conf_odd, nq_gauge** out);
/* LxField<quad_su3>: split on the device like split_lx_vector_into_eo_parts */
int nq_gauge_create_lx(nq_ctx* ctx, const double* conf_lx, nq_gauge** out);
void nq_gauge_destroy(nq_gauge* g);
/* No counterpart in the reference: for callers that cache gauge handles by host buffer
   (the shims' gauge cache).  nissa changes confs in place (put_boundaries_conditions,
   adapt_theta: operations/gaugeconf.cpp:125-160; add_or_rem_stagphases_to_conf:
   hmc/backfield.cpp:230-245; a new conf read into the same buffer), so a cached
   handle is valid only while the host conf is unchanged.  nq_host_conf_fingerprint
   hashes a sample of the buffer (the first and last site and ~65536 doubles spread
   over every real of a site); nq_gauge_matches_* set *same = 1 when the host conf's
   fingerprint equals the one taken when the handle was made from it.  A change
   confined to a few links can escape the sample: callers that make such changes
   destroy the handle themselves. */
int nq_host_conf_fingerprint(const double* data, int64_t ndoubles, uint64_t* out);
int nq_gauge_matches_lx(const nq_gauge* g, const double* conf_lx, int* same);
int nq_gauge_matches_eo(const nq_gauge* g, const double* conf_evn, const double* conf_odd, int* same);
/* device bytes held by the gauge handle (all precisions built so far) */
int64_t nq_gauge_device_bytes(const nq_gauge* g);
/* global_plaquette_lx_conf (operations/gauge/plaquette.hpp:52-84): the average of
   Re tr(U_mu U_nu U_mu^+ U_nu^+)/3 over the temporal planes and over the spatial
   planes, then the mean of the two; computed on the device from the handle's links */
int nq_gauge_plaquette(nq_ctx* ctx, const nq_gauge* g, double* plaq);
/* read_ildg_gauge_conf (io/reader.hpp:75-98, io/ILDG_File.cpp): the "ildg-binary-data"
   record of a LIME file (big-endian, 64- or 32-bit as its "ildg-format" record says;
   sites t slowest ... x fastest, directions x,y,z,t) into an lx conf in nissa's
   order (mu = 0 is t; quad_su3_ildg_to_nissa_reord, su3_op.hpp:217-225).
   checks (may be NULL) receives the "scidac-checksum" record's suma, sumb (0, 0
   when absent) and the ones computed from the data (ildgChecksum, io/checksum.hpp).
   nissa only warns when they differ, and so does this: compare them. */
int nq_read_ildg_gauge_conf(const char* path, int T, int LX, int LY, int LZ, double* conf_lx, uint32_t checks[4]);
/* read_real_vector(out, path, record_name) (io/reader.hpp:52-71): the named binary
   record (e.g. "scidac-binary-data" of a spincolor source or propagator file) as
   nreals_per_site reals per site in nissa's degree order, same site order and
   checksums as above; the precision (64 or 32 bit) follows from the record length */
/* unitarity_check_lx_conf (operations/gaugeconf.hpp:204-245), which nissa runs after
   reading a conf: per link |U U^+ - 1| (Frobenius, su3_get_non_unitariness,
   new_types/su3_op.hpp:3037-3055); out = {average, maximum}, nbroken (may be NULL)
   counts the links above 1e-13 */
int nq_host_unitarity_check_lx(int T, int LX, int LY, int LZ, const double* conf_lx, double out[2], int64_t* nbroken);
int nq_read_lime_real_vector(const char* path, const char* record, int nreals_per_site, int T, int LX, int LY, int LZ,
                             double* out_lx, uint32_t checks[4]);

/* ------------------------------------------------------------------ fields */
int nq_field_create(nq_ctx* ctx, int kind, int coverage, nq_field** out);
void nq_field_destroy(nq_field* f);
int nq_field_upload(nq_field* f, const double* host_aos);
int nq_field_download(const nq_field* f, double* host_aos);
int nq_field_copy(nq_field* dst, const nq_field* src);
int nq_field_norm2(const nq_field* f, double* out);

/* ------------------------------------------------------------------ twisted mass (device handles) */
/* tmn2Deo_or_tmn2Doe_eos (dirac_operator_tmD_eoprec_portable.hpp:36-169); out parity != in parity */
int nq_tm_hop(nq_ctx* ctx, const nq_gauge* g, const double* anis, nq_field* out, const nq_field* in);
/* tmDee_or_oo_eos / inv_tmDee_or_oo_eos (..._portable.hpp:174-222) */
int nq_tm_Dee(nq_ctx* ctx, double kappa, double mu, int inverse, nq_field* out, const nq_field* in);
/* tmDkern_eoprec_eos (..._portable.cpp:36-51) */
int nq_tm_kern(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
               nq_field* out_odd, const nq_field* in_odd);
/* tmDkern_eoprec_square_eos (..._portable.cpp:54-65) */
int nq_tm_kern_square(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
                      nq_field* out_odd, const nq_field* in_odd);
/* No counterpart in the reference: nq_tm_kern_square for callers that cannot allocate
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Its two FP64 temporaries live in a static 128 MiB device array of the library, so it
   calls no hipMalloc, hipFree or other allocating routine; lattices of more than 349504
   sites per parity (24^3x48 fits, 32^3x64 does not) are refused with NQ_EINVAL.  Every
   call on one GPU shares that array: calls that overlap on one GPU race.  Same kernels
   and results as nq_tm_kern_square; the launch ledger does not record it. */
int nq_tm_kern_square_noalloc(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
                              nq_field* out_odd, const nq_field* in_odd);
/* apply_tmQ on lx (dirac_operator_tmQ_portable.cpp:29-176) */
int nq_tm_apply_Q(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
                  nq_field* out_lx, const nq_field* in_lx);
/* apply_tmQ2 (dirac_operator_tmQ2.hpp:8): Q(-mu) Q(+mu) on lx */
int nq_tm_apply_Q2(nq_ctx* ctx, const nq_gauge* g, double kappa, double mu,
                   nq_field* out_lx, const nq_field* in_lx);
/* inv_tmDkern_eoprec_square_eos_cg (cg_invert_tmD_eoprec.cpp:61-78): solve K K^+ sol = src.
   residue is nissa's squared relative residual. */
int nq_tm_cg_kernsq(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
                    int niter, double residue, int prec, nq_field* sol_odd,
                    const nq_field* guess_odd, const nq_field* src_odd, nq_solve_stats* st);
/* inv_tmD_cg_eoprec (cg_invert_tmD_eoprec.cpp:152-213): full eo-preconditioned solve of
   D sol = src on lx, plus the FP64 check |g5 Q sol - src|^2/|src|^2.
   guess_Koo (odd field or NULL) is nissa's std::optional<OddField<spincolor>> guess_Koo
   (:102-103,152-153): the start of the K K^+ CG of eq. 9; on return it holds that CG's
   solution (the reference's "if a guess was passed, return new one", :137; nissa takes
   the optional by value, so its caller never sees the update: a shim that keeps nissa's
   semantics passes a copy). */
int nq_tm_solve_eoprec(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, double mu,
                       int niter, double residue, int prec, nq_field* sol_lx, nq_field* guess_Koo,
                       const nq_field* src_lx, nq_solve_stats* st, double* check_rel_res_sq);
/* inv_tmQ2_cg (cg_invert_tmQ2.cpp:12-19): CG on Q(-mu)Q(+mu) over lx */
int nq_tmQ2_cg(nq_ctx* ctx, const nq_gauge* g, double kappa, double mu, int niter, double residue,
               nq_field* sol_lx, const nq_field* guess_lx, const nq_field* src_lx, nq_solve_stats* st);
/* inv_tmQ2_cgm (cgm_invert_tmQ2.cpp:46-53): multi-mass CG on Q(-m_k)Q(+m_k) = Q(0)^2 + m_k^2
   over lx (shifts m_k^2 of the mu = 0 operator), sols[k] per mass; true_res_sq[k] gets the
   FP64 check |Q(-m)Q(+m) x_k - src|^2/|src|^2.  nmass <= 16. */
int nq_tmQ2_cgm(nq_ctx* ctx, const nq_gauge* g, double kappa, int nmass, const double* masses, int niter,
                const double* residues, nq_field** sols_lx, const nq_field* src_lx, nq_solve_stats* st,
                double* true_res_sq);
/* inv_tmDQ_cgm (cgm_invert_tmQ2.cpp:55-75): the same for the source g5 src */
int nq_tmDQ_cgm(nq_ctx* ctx, const nq_gauge* g, double kappa, int nmass, const double* masses, int niter,
                const double* residues, nq_field** sols_lx, const nq_field* src_lx, nq_solve_stats* st,
                double* true_res_sq);
/* multi-RHS eo solve: nrhs sources share the gauge handle, kappa and the iteration;
   mu[r] per right-hand side (per-RHS diagonal). Residuals are per RHS.  guess_Koo (NULL,
   or nrhs odd fields of which any may be NULL) as in nq_tm_solve_eoprec, per source. */
int nq_tm_solve_eoprec_multi(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis,
                             int nrhs, const double* mu, int niter, double residue, int prec,
                             nq_field** sol_lx, nq_field** guess_Koo, const nq_field* const* src_lx,
                             nq_solve_stats* st);

/* ------------------------------------------------------------------ twisted clover (device handles)
   Clover term from a gauge handle, kept on the device (FP64, both parities); the
   inverse twisted clover blocks are built on first use for each (kappa, mu). */
/* clover_term(Cl, cSW, conf) (operations/su3_paths/clover_term.hpp:88-96; chromo_operator,
   clover_term.cpp:56-104): four-leaf clover times -cSW/4 */
int nq_clover_create(nq_ctx* ctx, const nq_gauge* g, double csw, nq_clover** out);
/* the clover term as nissa's LxField<clover_term_t> AoS [V][4][3][3][2] (test/debug) */
int nq_clover_download_lx(const nq_clover* cl, double* cl_lx);
void nq_clover_destroy(nq_clover* cl);
/* tmclovDee_or_oo_eos (dirac_operators/tmclovD_eoprec/dirac_operator_tmclovD_eoprec_portable.cpp:19-42):
   (1/(2 kappa) + i mu g5 + Cl) on the field's parity, mu -> -mu when dag */
int nq_tmclov_Dee(nq_ctx* ctx, nq_clover* cl, double kappa, double mu, int dag, nq_field* out, const nq_field* in);
/* invert_twisted_clover_term (clover_term.hpp:160-270) + inv_tmclovDee_or_oo_eos
   (dirac_operator_tmclovD_eoprec.hpp:22-43) */
int nq_tmclov_inv_Dee(nq_ctx* ctx, nq_clover* cl, double kappa, double mu, int dag, nq_field* out,
                      const nq_field* in);
/* tmclovDkern_eoprec_eos (..._portable.cpp:45-70), odd -> odd */
int nq_tmclov_kern(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                   int dag, nq_field* out, const nq_field* in);
/* tmclovDkern_eoprec_square_eos (..._portable.cpp:73-86): K(dag) then K */
int nq_tmclov_kern_square(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                          double mu, nq_field* out, const nq_field* in);
/* apply_tmclovQ (dirac_operators/tmclovQ/dirac_operator_tmclovQ.cpp:14-190), lx fields */
int nq_tmclov_apply_Q(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                      nq_field* out, const nq_field* in);
/* apply_tmclovQ2 (dirac_operators/tmclovQ2/dirac_operator_tmclovQ2.hpp:9-20): Q(-mu) Q(+mu), lx */
int nq_tmclov_apply_Q2(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                       nq_field* out, const nq_field* in);
/* inv_tmclovQ2_cg (inverters/twisted_clover/cg_invert_tmclovQ2.hpp:9-24, cg_64_invert_tmclovQ2_portable.hpp:14-40):
   cg_invert on Q(-mu) Q(+mu) over lx fields, FP64; guess may be null */
int nq_tmclovQ2_cg(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, double mu,
                   int niter, double residue, nq_field* sol, const nq_field* guess, const nq_field* src,
                   nq_solve_stats* st);
/* inv_tmclovQ2_cgm (inverters/twisted_clover/cgm_invert_tmclovQ2.cpp:8-28): multi-shift CG on
   Q(0)^2 + m_k^2 over lx fields, FP64; sols[k] per mass, true_res[k] (optional) the FP64
   |b - (Q(0)^2 + m_k^2) x_k|^2 / |b|^2 */
int nq_tmclovQ2_cgm(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis, int nmass,
                    const double* masses, int niter, const double* residues, nq_field** sols, const nq_field* src,
                    nq_solve_stats* st, double* true_res);
/* inv_tmclovDkern_eoprec_square_eos_cg (inverters/twisted_clover/cg_invert_tmclovD_eoprec.cpp:33-54) */
int nq_tmclov_cg_kernsq(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                        double mu, int niter, double residue, int prec, nq_field* sol, const nq_field* guess,
                        const nq_field* src, nq_solve_stats* st);
/* inv_tmclovD_cg_eoprec (cg_invert_tmclovD_eoprec.cpp:56-232): eqs. 8-11 with the clover
   blocks, CG on K K^+ (prec NQ_PREC_F64, or NQ_PREC_F32 / NQ_PREC_F16 / NQ_PREC_F16H inner
   iterations with FP64 reliable updates; FP16 inner iterations use FP16 copies of the blocks),
   FP64 check |D sol - src|^2/|src|^2 */
int nq_tmclov_solve_eoprec(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                           double mu, int niter, double residue, int prec, nq_field* sol, const nq_field* src,
                           nq_solve_stats* st, double* check);
/* inv_tmclovD_cg_eoprec for nrhs sources at one (kappa, mu) and the handle's cSW, as the
   propagator driver calls it for its 12 spin-colour sources (projects/ib/prop.cpp:93-96):
   the K K^+ CG of up to 16 sources runs in the same launches, with the clover blocks of a
   site read once for all of them; sources stream through the slots; each result as
   nq_tmclov_solve_eoprec's (FP64 check in st[i].rel_res_sq) */
int nq_tmclov_solve_eoprec_multi(nq_ctx* ctx, const nq_gauge* g, nq_clover* cl, double kappa, const double* anis,
                                 double mu, int nrhs, int niter, double residue, int prec, nq_field** sols,
                                 const nq_field* const* srcs, nq_solve_stats* st);

/* ------------------------------------------------------------------ staggered */
/* apply_st2Doe (dirac_operator_stD.hpp:47-77) */
int nq_st_2Doe(nq_ctx* ctx, const nq_gauge* g, nq_field* out_odd, const nq_field* in_evn);
/* apply_stDeo_half (dirac_operator_stD.hpp:94-127) */
int nq_st_Deo_half(nq_ctx* ctx, const nq_gauge* g, nq_field* out_evn, const nq_field* in_odd);
/* apply_stD2ee_m2 (dirac_operator_stD.hpp:130-225): mass2 in - 1/4 Deo Doe in */
int nq_st_DDee(nq_ctx* ctx, const nq_gauge* g, double mass2, nq_field* out_evn, const nq_field* in_evn);
/* evn_apply_stD / odd_apply_stD (dag = 0) and their _dag (dag = 1), apply_stD
   (dirac_operator_stD.hpp:16-41,254-291): D = m + 1/2 hop per parity; a NULL output
   skips that parity */
int nq_st_apply_D(nq_ctx* ctx, const nq_gauge* g, double m, int dag, nq_field* out_evn, nq_field* out_odd,
                  const nq_field* in_evn, const nq_field* in_odd);
/* inv_evn_stD_cg (sol_odd NULL) / inv_stD_cg (cg_invert_evn_stD.hpp:17-31,
   cg_invert_stD.hpp:14-88): CG on m^2 - 1/4 Deo Doe for D^+ src, odd part rebuilt,
   check |D sol - src|^2/|src|^2 (full solve) */
int nq_st_solve_D(nq_ctx* ctx, const nq_gauge* g, double m, int niter, double residue, nq_field* sol_evn,
                  nq_field* sol_odd, const nq_field* guess_evn, const nq_field* src_evn,
                  const nq_field* src_odd, nq_solve_stats* st, double* check_rel_res_sq);
/* inv_stD2ee_m2_cg (cg_invert_stD2ee_m2.hpp:18-43) */
int nq_st_cg(nq_ctx* ctx, const nq_gauge* g, double mass2, int niter, double residue,
             nq_field* sol_evn, const nq_field* guess_evn, const nq_field* src_evn, nq_solve_stats* st);
/* cgm_invert with ApplyStD2eeM2Functor (cgm_invert_template_threaded.hpp:13-295,
   cgm_invert_stD2ee_m2_portable.hpp:10-29), batched over nrhs sources in one Krylov
   loop per source: sols[r*nshift+k]. prec NQ_PREC_F64 or NQ_PREC_F32 (FP32 inner,
   FP64 reliable updates). true_res_sq[r*nshift+k] gets the FP64 check. */
int nq_st_cgm(nq_ctx* ctx, const nq_gauge* g, int nrhs, int nshift, const double* shifts,
              int niter, const double* residues, int prec, nq_field** sols,
              const nq_field* const* srcs, nq_solve_stats* st, double* true_res_sq);

/* summ_src_and_all_inv_stD2ee_m2_cgm_portable (cgm_invert_stD2ee_m2_portable.hpp:31-55):
   chi_r = cons src_r + sum_k weights[k] x_{r,k},  (m2-less D2ee + poles[k]) x_{r,k} = src_r,
   one CGM over the rational approximation's poles per source (rat_approx_t,
   rat_approx.hpp:16-60), residue for every pole; nrhs sources batched. */
int nq_st_cgm_rat_sum(nq_ctx* ctx, const nq_gauge* g, int nrhs, int degree, double cons,
                      const double* poles, const double* weights, int niter, double residue, int prec,
                      nq_field** chis, const nq_field* const* srcs, nq_solve_stats* st);

/* summ_the_rootst_eoimpr_quark_force (hmc/fermions/rootst_eoimpr_quark_force.cpp:17-85): g holds the
   links with background field and staggered phases; u1b_evn/odd the background field ([Vh][4]
   complex, nissa eo order); the multi-shift solve runs on the poles to `residue` (prec
   NQ_PREC_F64 or NQ_PREC_F32 with FP64 reliable updates); F_evn/F_odd (nissa EoField<quad_su3>
   AoS) are summed into, before the TA projection and the product with U, as in the reference */
int nq_st_rootst_force(nq_ctx* ctx, const nq_gauge* g, const double* u1b_evn, const double* u1b_odd, int ndeg,
                       const double* poles, const double* weights, int niter, double residue, int prec,
                       const nq_field* pf, double* F_evn, double* F_odd, nq_solve_stats* st);

/* ------------------------------------------------------------------ host-buffer drop-ins */
/* Same operations over nissa host buffers; the gauge handle stays resident. */
/* tmDkern_eoprec_square_eos (..._portable.cpp:54-65) on OddField<spincolor> host buffers
   (Vh x 24 doubles, nissa eo order) */
int nq_tm_kern_square_host(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis,
                           double mu, double* out_odd, const double* in_odd);
/* inv_tmD_cg_eoprec (cg_invert_tmD_eoprec.cpp:152-213) on LxField<spincolor> host buffers;
   guess_Koo (NULL or Vh x 24 doubles, nissa eo order) as in nq_tm_solve_eoprec */
int nq_tm_solve_eoprec_host(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis,
                            double mu, int niter, double residue, int prec, double* sol_lx,
                            double* guess_Koo, const double* src_lx, nq_solve_stats* st,
                            double* check_rel_res_sq);

/* ------------------------------------------------------------------ host-side harness helpers */
/* Seeded synthetic gauge field U = exp(i eps H) per link, H traceless Hermitian with
   N(0,1/2) components in the Gell-Mann basis (nissa herm_put_to_gauss, random.hpp:311-338);
   counter-based Philox4x32-10 keyed by (seed, site, mu): identical on any host. */
int nq_host_gauge_random_lx(int T, int LX, int LY, int LZ, uint64_t seed, double eps, double* conf_lx);
/* the algebra element H and link for one (site, mu), for tests */
int nq_host_link_algebra(uint64_t seed, int64_t lx, int mu, double eps, double* H18, double* U18);
/* put_boundaries_conditions (gaugeconf.cpp:125-147): U_mu *= exp(i pi theta_mu / L_mu) */
int nq_host_gauge_phase_lx(int T, int LX, int LY, int LZ, const double* theta_in_pi, double* conf_lx);
/* add_or_rem_stagphases_to_conf (backfield.cpp:230-245) on an lx conf */
int nq_host_gauge_stagphases_lx(int T, int LX, int LY, int LZ, double* conf_lx);
/* sources: kind 0 = Z2 (+-1 real), 1 = Gaussian complex (N(0,1/2) per real),
   2 = point source (site0, spin-colour index sc) ; ncomp = 12 (spincolor) or 3 (color);
   parity_mask: 1 = even sites only, 2 = odd only, 3 = all (lx buffer) */
int nq_host_source_lx(int T, int LX, int LY, int LZ, uint64_t seed, int kind, int ncomp,
                      int parity_mask, int64_t site0, int sc, double* out_lx);
int nq_host_split_lx(int T, int LX, int LY, int LZ, int ndoubles, const double* lx, double* evn, double* odd);
int nq_host_paste_lx(int T, int LX, int LY, int LZ, int ndoubles, double* lx, const double* evn, const double* odd);

/* ------------------------------------------------------------------ serving */
/* Propagator-solve server (the caller of this path: the propagator driver solving 12
   spin-colour sources per mass on one configuration, projects/ib/prop.cpp:41-106,925-964).
   Gauge fields stay resident per volume; sources of queued requests join running
   batches of 16 slots as slots free up (continuous batching, per-source mu); a batch
   group is created only when its device footprint fits the HBM budget; volumes run on
   their own streams and threads.  A solution is bit-identical alone and in any mix. */
typedef struct nq_server nq_server;
typedef struct nq_serve_request {
  int volume;               /* id from nq_server_add_volume                             */
  int prec;                 /* NQ_PREC_* of the inner iteration                         */
  double kappa, mu;
  double residue;           /* nissa residue (squared relative), checked in FP64 per source */
  int niter;
  int nsrc;                 /* number of sources                                        */
  int src_kind;             /* 0: host lx spincolor buffers src[k];
                               1: point sources at lx site site0, spin-colour k = 0..nsrc-1 */
  int reserved;
  int64_t site0;
  const double* const* src; /* src_kind 0: nsrc host buffers, read when the source starts */
  double* const* sol;       /* nsrc host lx buffers written when each source finishes, or NULL */
  double csw;               /* 0: twisted mass; > 0: twisted clover with this cSW
                               (inv_tmclovD_cg_eoprec; its sources batch per (kappa, mu, cSW)) */
} nq_serve_request;
typedef struct nq_serve_result {
  int64_t id;
  int status;               /* NQ_OK, or the first error of its sources (NQ_EBUSY: refused) */
  int nsrc;
  int iters_max;            /* largest Krylov iteration count over the sources          */
  int reliable_updates;     /* summed over the sources                                  */
  double max_rel_res_sq;    /* worst FP64 |D sol - src|^2/|src|^2 over the sources        */
  double t_submit, t_admit, t_done;  /* seconds on the server clock                     */
} nq_serve_result;
/* hbm_budget_bytes <= 0: free device memory minus 4 GiB; max_groups_per_volume 1..3 */
int nq_server_create(int device, int64_t hbm_budget_bytes, int max_groups_per_volume, nq_server** out);
/* an lx conf (phases applied) that stays resident; returns the volume id */
int nq_server_add_volume(nq_server* s, int T, int LX, int LY, int LZ, const double* conf_lx, int* volume_id);
/* non-blocking; host source/solution buffers must stay valid until the request is done */
int nq_server_submit(nq_server* s, const nq_serve_request* req, int64_t* id);
/* blocks until request `id` is done (NQ_EBUSY on timeout); releases the request */
int nq_server_wait(nq_server* s, int64_t id, double timeout_s, nq_serve_result* out);
double nq_server_clock(const nq_server* s);
int64_t nq_server_device_bytes(const nq_server* s);  /* resident gauges + admitted groups */
void nq_server_destroy(nq_server* s);

/* ------------------------------------------------------------------ instrumentation */
/* napply back-to-back hop applies on device fields, timed with HIP events on the
   ctx stream; prec selects the storage the kernel reads (NQ_PREC_*). */
int nq_bench_tm_hop(nq_ctx* ctx, const nq_gauge* g, int prec, int napply, nq_field* out,
                    const nq_field* in, double* ms_per_apply, double* bytes_per_apply);
/* the same for nrhs right-hand sides (copies of `in`) per launch, as the batched solves
   run it; bytes count the links once and the spinors per right-hand side */
int nq_bench_tm_hop_multi(nq_ctx* ctx, const nq_gauge* g, int prec, int nrhs, int napply,
                          const nq_field* in, double* ms_per_apply, double* bytes_per_apply);

#ifdef __cplusplus
}
#endif
#endif
