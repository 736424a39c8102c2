/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C (C11 + OpenMP + __float128) restatement of nissa's fermion-solve
 * path, used only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  It is never the product: the library
 * under nissa_amd/ does not link, load or call anything here.
 *
 * Layouts follow nissa's CPU ("AoS") layout exactly
 * (src/base/field.hpp:408-416, src/new_types/su3.hpp:19-55):
 *   spincolor  double[4][3][2]     (Dirac index, colour, re/im)
 *   color      double[3][2]
 *   su3        double[3][3][2]
 *   quad_su3   double[4][3][3][2]  (direction mu: 0=t,1=x,2=y,3=z)
 * lx sites are ordered t-major, z fastest (src/geometry/geometry_lx.cpp:51-60);
 * eo sites are the rank among same-parity sites in lx order
 * (src/geometry/geometry_eo.cpp:67-77), which for even extents is lx>>1.
 * An eo gauge field is two buffers (even sites, odd sites), each Vh quad_su3,
 * as nissa's EoField<quad_su3> (src/base/field.hpp:1586-1599).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * of oracle.c.
 */
#ifndef NISSA_AMD_ORACLE_H
#define NISSA_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_geom or_geom;

/* geometry (geometry_lx.cpp:51-60,403-424; geometry_eo.cpp:41-113) */
/* OpenMP threads the oracle's parallel loops use (omp_get_max_threads) */
int or_omp_max_threads(void);
or_geom* or_geom_create(int T, int LX, int LY, int LZ);
void     or_geom_destroy(or_geom* g);
int64_t  or_geom_vol(const or_geom* g);
int64_t  or_geom_lx_of_eo(const or_geom* g, int par, int64_t ieo);
int64_t  or_geom_eo_of_lx(const or_geom* g, int64_t lx);
int      or_geom_parity(const or_geom* g, int64_t lx);
int64_t  or_geom_neigh_lx(const or_geom* g, int64_t lx, int mu, int up);

/* lx <-> eo (geometry_mix.hpp:21-82); ndoubles = doubles per site */
void or_split_lx(const or_geom* g, int ndoubles, const double* lx, double* evn, double* odd);
void or_paste_lx(const or_geom* g, int ndoubles, double* lx, const double* evn, const double* odd);

/* ---------------- twisted mass (dirac_operator_tmD_eoprec_portable.*) -------- */
/* out[par=out_par] = sum_mu (1-g_mu)U psi(x+mu) + (1+g_mu)U^+ psi(x-mu)   (tmn2D) */
void or_tm_hop(const or_geom* g, int out_par, double* out,
               const double* conf_evn, const double* conf_odd,
               const double* anis /* NULL = isotropic */, const double* in);
void or_tmDee(const or_geom* g, double* out, double kappa, double mu, const double* in);
void or_inv_tmDee(const or_geom* g, double* out, double kappa, double mu, const double* in);
void or_tm_put_together_g5(const or_geom* g, double* out, const double* temp);
void or_tmDkern(const or_geom* g, double* out, double* temp_evn,
                const double* conf_evn, const double* conf_odd,
                double kappa, const double* anis, double mu, const double* in);
void or_tmDkern_square(const or_geom* g, double* out, double* temp_odd, double* temp_evn,
                       const double* conf_evn, const double* conf_odd,
                       double kappa, const double* anis, double mu, const double* in);
/* Q = g5 D on the full lattice (dirac_operator_tmQ_portable.cpp:29-176) */
void or_apply_tmQ(const or_geom* g, double* out, const double* conf_lx,
                  double kappa, const double* anis, double mu, const double* in);
/* legacy Q+Q- on lx (dirac_operator_tmQ2.hpp:8, body as the commented original) */
void or_apply_tmQ2(const or_geom* g, double* out, const double* conf_lx,
                   double kappa, double* temp, double mu, const double* in);

/* ---------------- staggered (dirac_operator_stD.hpp) ------------------------- */
void or_st_2Doe(const or_geom* g, double* out_odd, const double* conf_evn,
                const double* conf_odd, const double* in_evn);
void or_st_Deo_half(const or_geom* g, double* out_evn, const double* conf_evn,
                    const double* conf_odd, const double* in_odd);
void or_stD2ee_m2(const or_geom* g, double* out_evn, const double* conf_evn,
                  const double* conf_odd, double* temp_odd, double mass2,
                  const double* in_evn);
void or_apply_stD(const or_geom* g, double* out_evn, double* out_odd,
                  const double* conf_evn, const double* conf_odd, double m,
                  const double* in_evn, const double* in_odd);

/* ---------------- reductions (field.hpp:316-367,863-905) --------------------- */
double or_scalar_prod(int64_t nsites, int ndoubles, const double* a, const double* b);

/* ---------------- inverters --------------------------------------------------- */
typedef struct {
  int iters;              /* CG iterations (final_iter) */
  double rel_res_sq;      /* true residual |b-Ax|^2/|b|^2 recomputed at the end */
  int converged;          /* 1 if iter<niter, as nissa's CRASH condition */
} or_cg_stats;

/* inv_tmDkern_eoprec_square_eos_cg_64 -> cg_invert (cg_invert_template_threaded.hpp:19-150) */
int or_tm_cg_kernsq(const or_geom* g, double* sol, const double* guess,
                    const double* conf_evn, const double* conf_odd,
                    double kappa, const double* anis, double mu, int niter,
                    double residue, const double* source, or_cg_stats* st);
/* inv_tmD_cg_eoprec_native + residual check (cg_invert_tmD_eoprec.cpp:102-213) */
int or_tm_solve_eoprec(const or_geom* g, double* sol_lx, const double* conf_lx,
                       double kappa, const double* anis, double mu, int niter,
                       double residue, const double* source_lx, or_cg_stats* st,
                       double* check_rel_res_sq);
/* legacy inv_tmQ2_cg: CG on Q(-mu)Q(+mu) over lx (cg_invert_tmQ2.cpp:12-19) */
int or_tmQ2_cg(const or_geom* g, double* sol, const double* guess, const double* conf_lx,
               double kappa, double mu, int niter, double residue,
               const double* source, or_cg_stats* st);
/* inv_stD2ee_m2_cg (cg_invert_stD2ee_m2.hpp:18-43) */
int or_st_cg(const or_geom* g, double* sol, const double* conf_evn, const double* conf_odd,
             double mass2, int niter, double residue, const double* source, or_cg_stats* st);
/* cgm_invert with the staggered functor (cgm_invert_template_threaded.hpp:13-295,
   cgm_invert_stD2ee_m2_portable.hpp:10-29); sols is nshift pointers */
int or_st_cgm(const or_geom* g, double** sols, const double* conf_evn, const double* conf_odd,
              int nshift, const double* shifts, int niter, const double* residues,
              const double* source, or_cg_stats* st, double* true_res_sq /* nshift */);
/* cgm on legacy lx Q0^2 + m^2 (cgm_invert_tmQ2.cpp:46-53): masses m[k] -> shifts m^2 */
int or_tmQ2_cgm(const or_geom* g, double** sols, const double* conf_lx, double kappa,
                int nmass, const double* masses, int niter, const double* residues,
                const double* source, or_cg_stats* st, double* true_res_sq);
/* rational sum chi = c0*src + sum_k w_k x_k (cgm_invert_stD2ee_m2_portable.hpp:31-55) */
int or_st_rat_summ(const or_geom* g, double* chi, const double* conf_evn, const double* conf_odd,
                   int ndeg, double cons, const double* poles, const double* weights,
                   int niter, double residue, const double* source, or_cg_stats* st);

/* Quad-precision defect correction for K K^+ y = phi: residuals and the solution
   accumulated in __float128, inner CG in double
   (cg_128_invert_template_threaded.cpp:56-95, float_128.hpp:28-290).
   Returns the final relative squared residual computed in __float128. */
int or_tm_cg128_kernsq(const or_geom* g, double* sol, const double* conf_evn,
                       const double* conf_odd, double kappa, const double* anis,
                       double mu, double residue, const double* source,
                       double* final_rel_res_sq, int* outer_iters);
/* K K^+ applied with every operation in __float128 (inputs in double, output rounded) */
void or_tmDkern_square_q(const or_geom* g, double* out, const double* conf_evn,
                         const double* conf_odd, double kappa, const double* anis,
                         double mu, const double* in);

/* ---------------- helpers for tests ------------------------------------------ */
/* Average plaquette of an lx conf (nissa global_plaquette_lx_conf normalisation:
   Re tr / 3 averaged over the 6 planes and all sites). */
double or_plaquette_lx(const or_geom* g, const double* conf_lx);

/* twisted clover (clover_term.hpp/.cpp, dirac_operator_tmclovD_eoprec*, dirac_operator_tmclovQ.cpp,
   cg_invert_tmclovD_eoprec.cpp) */
void or_clover_term_lx(const or_geom* g, double* cl, const double* conf_lx, double csw);
void or_tmclovDee(const or_geom* g, double* out, const double* cl, double kappa, double mu, int dag,
                  const double* in);
void or_tmclov_invert(const or_geom* g, double* inv, const double* cl, double kappa, double mass);
void or_inv_tmclovDee(const or_geom* g, double* out, const double* inv, int dag, const double* in);
void or_tmclovDkern(const or_geom* g, double* out, double* temp, const double* ce, const double* co,
                    double kappa, const double* anis, const double* cl_odd, const double* inv_evn, int dag,
                    double mu, const double* in);
void or_tmclovDkern_square(const or_geom* g, double* out, double* t1, double* t2, const double* ce,
                           const double* co, double kappa, const double* anis, const double* cl_odd,
                           const double* inv_evn, double mu, const double* in);
void or_apply_tmclovQ(const or_geom* g, double* out, const double* conf_lx, double kappa, const double* anis,
                      const double* cl_lx, double mu, const double* in);
int or_rootst_force(const or_geom* g, double* Fe, double* Fo, const double* ce, const double* co,
                    const double* u1e, const double* u1o, const double* pf, int ndeg, const double* poles,
                    const double* weights, int niter, double residue);
int or_tmclov_solve_eoprec(const or_geom* g, double* sol_lx, const double* conf_lx, double kappa,
                           const double* anis, double csw, double mu, int niter, double residue,
                           const double* source_lx, or_cg_stats* st, double* check);
/* Random site-local gauge rotation: conf' = O(x) U_mu(x) O(x+mu)^+, psi' = O psi */
void or_gauge_transform_lx(const or_geom* g, double* conf_lx, const double* omega /*V su3*/);
void or_color_rotate_lx(const or_geom* g, int nspin, double* psi, const double* omega);

#ifdef __cplusplus
}
#endif
#endif
