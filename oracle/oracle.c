/*
 * CPU ORACLE — test infrastructure only (see oracle.h).
 *
 * Each function restates the reference algorithm named beside it.  How each
 * one is pinned (tests/test_oracle_*.py):
 *   geometry           closed forms: ieo == lx>>1, parity, neighbour wrap      pinned
 *   or_tm_hop          plane-wave closed form with nissa's gamma matrices
 *                      (dirac.cpp:43-58) at U=1; gauge covariance; gamma5
 *                      hermiticity; tmLQCD propagator fixture on L4T8conf     pinned
 *   tmDee/inv/put_tog  dense Schur complement of apply_tmQ on 4^4             pinned
 *   or_apply_tmQ       plane-wave closed form; gauge covariance; fixture       pinned
 *   kern / kernsq      dense Schur complement; K(-mu) = K(mu)^+ identity       pinned
 *   staggered ops      plane-wave closed form with eta phases; covariance;
 *                      D_eo^+ = -D_oe identity; dense Schur complement         pinned
 *   reductions         exact integer-valued sums; __float128 sums              pinned
 *   CG / CGM           dense direct solve (4^4), true residual recomputation,
 *                      CGM shift == single-shift CG                            pinned
 *   cg128              residual in __float128 below 1e-28                      pinned
 *   or_st_rat_summ     the same sum from dense solves on 4x4x2x2; partial
 *                      fractions of 1/((x+a)(x+b))                             pinned
 *   or_rootst_force    finite-difference gradient of the pseudofermion action  pinned
 *   plaquette          L4T8conf fixture value 0.6350055722288719              pinned
 *   clover term        constant abelian fields in each plane against
 *                      -cSW/4 sum_{mu<nu} gamma_mu gamma_nu F_munu built from
 *                      nissa's gamma matrices; hermitian blocks; gauge covariance pinned
 *   tmclov Dee/inv/K   U=1 reduces to the tm operators; Dee^-1 Dee = 1;
 *                      K(-mu) = K(mu)^+ identity; dense Schur complement of
 *                      apply_tmclovQ on 2^4 x 4; eo solve solves the lx system  pinned
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

int or_omp_max_threads(void) { return omp_get_max_threads(); }

struct or_geom {
  int L[4];
  int64_t V, Vh;
  int32_t* up;      /* [V][4] lx neighbours */
  int32_t* dw;
  int8_t* par;      /* [V] */
  int32_t* eo_of_lx;
  int32_t* lx_of_eo[2];
  int32_t* eoup[2]; /* [Vh][4] eo index of the (opposite parity) neighbour */
  int32_t* eodw[2];
};

/* lx index of coordinates, t-major z fastest (geometry_lx.cpp:51-60) */
static int64_t lx_of_coord(const int L[4], const int x[4])
{
  int64_t i = 0;
  for (int mu = 0; mu < 4; mu++) i = i * L[mu] + x[mu];
  return i;
}
static void coord_of_lx(const int L[4], int64_t i, int x[4])
{
  for (int mu = 3; mu >= 0; mu--) { x[mu] = (int)(i % L[mu]); i /= L[mu]; }
}

or_geom* or_geom_create(int T, int LX, int LY, int LZ)
{
  or_geom* g = (or_geom*)calloc(1, sizeof(or_geom));
  g->L[0] = T; g->L[1] = LX; g->L[2] = LY; g->L[3] = LZ;
  for (int mu = 0; mu < 4; mu++)
    if (g->L[mu] < 2 || g->L[mu] % 2) { free(g); return NULL; }
  g->V = (int64_t)T * LX * LY * LZ;
  g->Vh = g->V / 2;
  g->up = (int32_t*)malloc(sizeof(int32_t) * g->V * 4);
  g->dw = (int32_t*)malloc(sizeof(int32_t) * g->V * 4);
  g->par = (int8_t*)malloc(g->V);
  g->eo_of_lx = (int32_t*)malloc(sizeof(int32_t) * g->V);
  for (int p = 0; p < 2; p++) {
    g->lx_of_eo[p] = (int32_t*)malloc(sizeof(int32_t) * g->Vh);
    g->eoup[p] = (int32_t*)malloc(sizeof(int32_t) * g->Vh * 4);
    g->eodw[p] = (int32_t*)malloc(sizeof(int32_t) * g->Vh * 4);
  }
  /* neighbours with periodic wrap on one rank (geometry_lx.cpp:403-424) */
  for (int64_t i = 0; i < g->V; i++) {
    int x[4];
    coord_of_lx(g->L, i, x);
    g->par[i] = (int8_t)((x[0] + x[1] + x[2] + x[3]) % 2);
    for (int mu = 0; mu < 4; mu++) {
      int y[4] = {x[0], x[1], x[2], x[3]};
      y[mu] = (x[mu] + 1) % g->L[mu];
      g->up[i * 4 + mu] = (int32_t)lx_of_coord(g->L, y);
      y[mu] = (x[mu] - 1 + g->L[mu]) % g->L[mu];
      g->dw[i * 4 + mu] = (int32_t)lx_of_coord(g->L, y);
    }
  }
  /* eo labelling: rank among same-parity sites in lx order (geometry_eo.cpp:67-77) */
  int64_t cnt[2] = {0, 0};
  for (int64_t i = 0; i < g->V; i++) {
    int p = g->par[i];
    g->eo_of_lx[i] = (int32_t)cnt[p];
    g->lx_of_eo[p][cnt[p]] = (int32_t)i;
    cnt[p]++;
  }
  for (int p = 0; p < 2; p++)
    for (int64_t e = 0; e < g->Vh; e++) {
      int64_t i = g->lx_of_eo[p][e];
      for (int mu = 0; mu < 4; mu++) {
        g->eoup[p][e * 4 + mu] = g->eo_of_lx[g->up[i * 4 + mu]];
        g->eodw[p][e * 4 + mu] = g->eo_of_lx[g->dw[i * 4 + mu]];
      }
    }
  return g;
}

void or_geom_destroy(or_geom* g)
{
  if (!g) return;
  free(g->up); free(g->dw); free(g->par); free(g->eo_of_lx);
  for (int p = 0; p < 2; p++) { free(g->lx_of_eo[p]); free(g->eoup[p]); free(g->eodw[p]); }
  free(g);
}

int64_t or_geom_vol(const or_geom* g) { return g->V; }
int64_t or_geom_lx_of_eo(const or_geom* g, int par, int64_t ieo) { return g->lx_of_eo[par][ieo]; }
int64_t or_geom_eo_of_lx(const or_geom* g, int64_t lx) { return g->eo_of_lx[lx]; }
int or_geom_parity(const or_geom* g, int64_t lx) { return g->par[lx]; }
int64_t or_geom_neigh_lx(const or_geom* g, int64_t lx, int mu, int up)
{
  return up ? g->up[lx * 4 + mu] : g->dw[lx * 4 + mu];
}

/* split_lx_vector_into_eo_parts / paste_eo_parts_into_lx_vector (geometry_mix.hpp:21-60) */
void or_split_lx(const or_geom* g, int nd, const double* lx, double* evn, double* odd)
{
  double* dst[2] = {evn, odd};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < g->V; i++)
    memcpy(dst[g->par[i]] + (int64_t)g->eo_of_lx[i] * nd, lx + i * nd, sizeof(double) * nd);
}
void or_paste_lx(const or_geom* g, int nd, double* lx, const double* evn, const double* odd)
{
  const double* src[2] = {evn, odd};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < g->V; i++)
    memcpy(lx + i * nd, src[g->par[i]] + (int64_t)g->eo_of_lx[i] * nd, sizeof(double) * nd);
}

/* ---- the tm operator core in double and in __float128 ---- */
#define R double
#define SFX(n) d_##n
#include "oracle_tm_core.inc"
#undef R
#undef SFX
#define R __float128
#define SFX(n) q_##n
#include "oracle_tm_core.inc"
#undef R
#undef SFX

void or_tm_hop(const or_geom* g, int op, double* out, const double* ce, const double* co,
               const double* anis, const double* in)
{
  d_hop(g, op, out, ce, co, anis, in);
}
void or_tmDee(const or_geom* g, double* out, double kappa, double mu, const double* in) { d_tmDee(g, out, kappa, mu, in); }
void or_inv_tmDee(const or_geom* g, double* out, double kappa, double mu, const double* in) { d_inv_tmDee(g, out, kappa, mu, in); }
void or_tm_put_together_g5(const or_geom* g, double* out, const double* temp) { d_put_tog(g, out, temp); }
void or_tmDkern(const or_geom* g, double* out, double* temp, const double* ce, const double* co,
                double kappa, const double* anis, double mu, const double* in)
{
  d_kern(g, out, temp, ce, co, kappa, anis, mu, in);
}
void or_tmDkern_square(const or_geom* g, double* out, double* t1, double* t2, const double* ce,
                       const double* co, double kappa, const double* anis, double mu, const double* in)
{
  d_kernsq(g, out, t1, t2, ce, co, kappa, anis, mu, in);
}

void or_tmDkern_square_q(const or_geom* g, double* out, const double* ce, const double* co,
                         double kappa, const double* anis, double mu, const double* in)
{
  const int64_t n = g->Vh * 24;
  __float128* qi = (__float128*)malloc(sizeof(__float128) * n);
  __float128* qo = (__float128*)malloc(sizeof(__float128) * n);
  __float128* t1 = (__float128*)malloc(sizeof(__float128) * n);
  __float128* t2 = (__float128*)malloc(sizeof(__float128) * n);
  for (int64_t i = 0; i < n; i++) qi[i] = in[i];
  q_kernsq(g, qo, t1, t2, ce, co, kappa, anis, mu, qi);
  for (int64_t i = 0; i < n; i++) out[i] = (double)qo[i];
  free(qi); free(qo); free(t1); free(t2);
}

/* ---- full-lattice Q = g5 D (dirac_operator_tmQ_portable.cpp:29-176) ---- */
void or_apply_tmQ(const or_geom* g, double* out, const double* conf, double kappa,
                  const double* anis, double mass, const double* in)
{
  const double kcf = 1 / (2 * kappa);
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->V; X++) {
    d_cx o[4][3];
    for (int id = 0; id < 4; id++)
      for (int ic = 0; ic < 3; ic++) { o[id][ic].re = 0; o[id][ic].im = 0; }
    for (int mu = 0; mu < 4; mu++) {
      d_cx t0[3], t1[3], t2[3], t3[3];
      const double* s = in + (int64_t)g->up[X * 4 + mu] * 24;
      for (int ic = 0; ic < 3; ic++) {
        d_cx s0 = d_sc_get(s, 0, ic), s1 = d_sc_get(s, 1, ic), s2 = d_sc_get(s, 2, ic), s3 = d_sc_get(s, 3, ic);
        switch (mu) {
        case 0: t0[ic] = d_c_add(s0, s2); t1[ic] = d_c_add(s1, s3); break;
        case 1: t0[ic] = d_c_iadd(s0, s3); t1[ic] = d_c_iadd(s1, s2); break;
        case 2: t0[ic] = d_c_add(s0, s3); t1[ic] = d_c_sub(s1, s2); break;
        default: t0[ic] = d_c_iadd(s0, s2); t1[ic] = d_c_isub(s1, s3); break;
        }
      }
      const double* Uf = conf + (X * 4 + mu) * 18;
      d_su3_col(t2, Uf, t0);
      d_su3_col(t3, Uf, t1);
      if (anis && mu > 0)
        for (int ic = 0; ic < 3; ic++) { t2[ic] = d_c_scal(t2[ic], *anis); t3[ic] = d_c_scal(t3[ic], *anis); }
      for (int ic = 0; ic < 3; ic++) {
        o[0][ic] = d_c_add(o[0][ic], t2[ic]);
        o[1][ic] = d_c_add(o[1][ic], t3[ic]);
        switch (mu) {
        case 0: o[2][ic] = d_c_add(o[2][ic], t2[ic]); o[3][ic] = d_c_add(o[3][ic], t3[ic]); break;
        case 1: o[2][ic] = d_c_isub(o[2][ic], t3[ic]); o[3][ic] = d_c_isub(o[3][ic], t2[ic]); break;
        case 2: o[2][ic] = d_c_sub(o[2][ic], t3[ic]); o[3][ic] = d_c_add(o[3][ic], t2[ic]); break;
        default: o[2][ic] = d_c_isub(o[2][ic], t2[ic]); o[3][ic] = d_c_iadd(o[3][ic], t3[ic]); break;
        }
      }
      const int64_t Xdw = g->dw[X * 4 + mu];
      s = in + Xdw * 24;
      for (int ic = 0; ic < 3; ic++) {
        d_cx s0 = d_sc_get(s, 0, ic), s1 = d_sc_get(s, 1, ic), s2 = d_sc_get(s, 2, ic), s3 = d_sc_get(s, 3, ic);
        switch (mu) {
        case 0: t0[ic] = d_c_sub(s0, s2); t1[ic] = d_c_sub(s1, s3); break;
        case 1: t0[ic] = d_c_isub(s0, s3); t1[ic] = d_c_isub(s1, s2); break;
        case 2: t0[ic] = d_c_sub(s0, s3); t1[ic] = d_c_add(s1, s2); break;
        default: t0[ic] = d_c_isub(s0, s2); t1[ic] = d_c_iadd(s1, s3); break;
        }
      }
      const double* Ub = conf + (Xdw * 4 + mu) * 18;
      d_su3dag_col(t2, Ub, t0);
      d_su3dag_col(t3, Ub, t1);
      if (anis && mu > 0)
        for (int ic = 0; ic < 3; ic++) { t2[ic] = d_c_scal(t2[ic], *anis); t3[ic] = d_c_scal(t3[ic], *anis); }
      for (int ic = 0; ic < 3; ic++) {
        o[0][ic] = d_c_add(o[0][ic], t2[ic]);
        o[1][ic] = d_c_add(o[1][ic], t3[ic]);
        switch (mu) {
        case 0: o[2][ic] = d_c_sub(o[2][ic], t2[ic]); o[3][ic] = d_c_sub(o[3][ic], t3[ic]); break;
        case 1: o[2][ic] = d_c_iadd(o[2][ic], t3[ic]); o[3][ic] = d_c_iadd(o[3][ic], t2[ic]); break;
        case 2: o[2][ic] = d_c_add(o[2][ic], t3[ic]); o[3][ic] = d_c_sub(o[3][ic], t2[ic]); break;
        default: o[2][ic] = d_c_iadd(o[2][ic], t2[ic]); o[3][ic] = d_c_isub(o[3][ic], t3[ic]); break;
        }
      }
    }
    /* -1/2 on the hop, the diagonal, and gamma5 (:164-174) */
    const double* i0 = in + X * 24;
    double* d = out + X * 24;
    for (int id = 0; id < 4; id++)
      for (int c = 0; c < 3; c++) {
        const double ir = i0[(id * 3 + c) * 2], ii = i0[(id * 3 + c) * 2 + 1];
        if (id < 2) {
          d[(id * 3 + c) * 2] = -0.5 * o[id][c].re + kcf * ir - mass * ii;
          d[(id * 3 + c) * 2 + 1] = -0.5 * o[id][c].im + kcf * ii + mass * ir;
        } else {
          d[(id * 3 + c) * 2] = +0.5 * o[id][c].re - kcf * ir - mass * ii;
          d[(id * 3 + c) * 2 + 1] = +0.5 * o[id][c].im - kcf * ii + mass * ir;
        }
      }
  }
}

/* apply_tmQ2: temp = Q(+mu) in; out = Q(-mu) temp (dirac_operator_tmQ2.cpp:19-32, the
   commented original body with RL=0) */
void or_apply_tmQ2(const or_geom* g, double* out, const double* conf, double kappa, double* temp,
                   double mu, const double* in)
{
  double* t = temp ? temp : (double*)malloc(sizeof(double) * g->V * 24);
  or_apply_tmQ(g, t, conf, kappa, NULL, +mu, in);
  or_apply_tmQ(g, out, conf, kappa, NULL, -mu, t);
  if (!temp) free(t);
}

/* ---- staggered (dirac_operator_stD.hpp) ---- */
/* colour helpers over double[3][2] */
static inline void col_zero(d_cx o[3]) { for (int c = 0; c < 3; c++) { o[c].re = 0; o[c].im = 0; } }
static inline void col_load(d_cx v[3], const double* p) { for (int c = 0; c < 3; c++) { v[c].re = p[2 * c]; v[c].im = p[2 * c + 1]; } }

/* out(x, parity op) = sum_mu U_mu(x) in(x+mu) - U_mu(x-mu)^+ in(x-mu) */
static void st_hop(const or_geom* g, int op, double* out, const double* ce, const double* co, const double* in)
{
  const double* conf[2] = {ce, co};
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->Vh; X++) {
    d_cx o[3], v[3], t[3];
    col_zero(o);
    for (int mu = 0; mu < 4; mu++) {
      const int64_t up = g->eoup[op][X * 4 + mu], dw = g->eodw[op][X * 4 + mu];
      col_load(v, in + up * 6);
      d_su3_col(t, conf[op] + (X * 4 + mu) * 18, v);
      for (int c = 0; c < 3; c++) o[c] = d_c_add(o[c], t[c]);
      col_load(v, in + dw * 6);
      d_su3dag_col(t, conf[!op] + (dw * 4 + mu) * 18, v);
      for (int c = 0; c < 3; c++) o[c] = d_c_sub(o[c], t[c]);
    }
    for (int c = 0; c < 3; c++) { out[X * 6 + 2 * c] = o[c].re; out[X * 6 + 2 * c + 1] = o[c].im; }
  }
}

/* apply_st2Doe (:47-77) */
void or_st_2Doe(const or_geom* g, double* out, const double* ce, const double* co, const double* in)
{
  st_hop(g, 1, out, ce, co, in);
}
/* apply_stDeo_half (:94-127): includes the 0.25 */
void or_st_Deo_half(const or_geom* g, double* out, const double* ce, const double* co, const double* in)
{
  st_hop(g, 0, out, ce, co, in);
  for (int64_t i = 0; i < g->Vh * 6; i++) out[i] *= 0.25;
}
/* apply_stD2ee_m2 (:130-225): m2 in - 1/4 Deo Doe in */
void or_stD2ee_m2(const or_geom* g, double* out, const double* ce, const double* co, double* temp,
                  double mass2, const double* in)
{
  st_hop(g, 1, temp, ce, co, in);
  st_hop(g, 0, out, ce, co, temp);
  const int64_t n = g->Vh * 6;
  if (mass2 != 0) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = mass2 * in[i] - out[i] * 0.25;
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] *= -0.25;
  }
}
/* apply_stD = (evn_apply_stD, odd_apply_stD) (:16-41,254-291) */
void or_apply_stD(const or_geom* g, double* oe, double* oo, const double* ce, const double* co,
                  double m, const double* ie, const double* io)
{
  or_st_Deo_half(g, oe, ce, co, io);
  for (int64_t i = 0; i < g->Vh * 6; i++) oe[i] = ie[i] * m + 2 * oe[i];
  or_st_2Doe(g, oo, ce, co, ie);
  for (int64_t i = 0; i < g->Vh * 6; i++) oo[i] = io[i] * m + 0.5 * oo[i];
}

/* ---- reductions: per-site partial, then stride-halving tree (field.hpp:334-355,863-891) ---- */
double or_scalar_prod(int64_t nsites, int nd, const double* a, const double* b)
{
  if (nsites <= 0) return 0;
  double* buf = (double*)malloc(sizeof(double) * nsites);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < nsites; s++) {
    double r = 0;
    for (int k = 0; k < nd; k++) r += a[s * nd + k] * b[s * nd + k];
    buf[s] = r;
  }
  int64_t n = nsites;
  while (n > 1) {
    const int64_t stride = (n + 1) / 2, nred = n / 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nred; i++) buf[i] += buf[i + stride];
    n = stride;
  }
  double r = buf[0];
  free(buf);
  return r;
}

/* ---- generic CG (cg_invert_template_threaded.hpp:19-150) ---- */
typedef void (*or_op)(void* ctx, double* out, const double* in);

static int cg_generic(int64_t nsites, int nd, double* sol, const double* guess, or_op f, void* ctx,
                      int niter, double residue, const double* source, or_cg_stats* st)
{
  const int64_t n = nsites * nd;
  double* s = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* r = (double*)malloc(sizeof(double) * n);
  if (guess) memcpy(sol, guess, sizeof(double) * n);
  else memset(sol, 0, sizeof(double) * n);
  const double sourceNorm2 = or_scalar_prod(nsites, nd, source, source);
  int rc = 0;
  if (sourceNorm2 == 0 || isnan(sourceNorm2)) { rc = -1; goto done; } /* "invalid norm" */
  f(ctx, s, sol);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) { r[i] = source[i] - s[i]; p[i] = r[i]; }
  double delta = or_scalar_prod(nsites, nd, r, r), lambda;
  int iter = 0, final_iter = 0;
  do {
    final_iter = ++iter;
    f(ctx, s, p);
    const double alpha = or_scalar_prod(nsites, nd, s, p);
    const double omega = delta / alpha;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) { sol[i] += p[i] * omega; r[i] -= s[i] * omega; }
    lambda = or_scalar_prod(nsites, nd, r, r);
    const double gammag = lambda / delta;
    delta = lambda;
    if (isnan(gammag)) { rc = -2; goto done; } /* "nanned" */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = r[i] + p[i] * gammag;
  } while (lambda >= residue * sourceNorm2 && iter < niter);
  /* last calculation of the residual */
  f(ctx, s, sol);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) r[i] = source[i] - s[i];
  lambda = or_scalar_prod(nsites, nd, r, r);
  if (st) { st->iters = final_iter; st->rel_res_sq = lambda / sourceNorm2; st->converged = final_iter != niter; }
  if (final_iter == niter) rc = -3; /* "exit without converging" */
done:
  free(s); free(p); free(r);
  return rc;
}

typedef struct {
  const or_geom* g;
  const double *ce, *co, *conf_lx, *anis;
  double kappa, mu;
  double *t1, *t2;
} tm_ctx;

static void op_kernsq(void* v, double* out, const double* in)
{
  tm_ctx* c = (tm_ctx*)v;
  d_kernsq(c->g, out, c->t1, c->t2, c->ce, c->co, c->kappa, c->anis, c->mu, in);
}

int or_tm_cg_kernsq(const or_geom* g, double* sol, const double* guess, const double* ce,
                    const double* co, double kappa, const double* anis, double mu, int niter,
                    double residue, const double* source, or_cg_stats* st)
{
  tm_ctx c = {g, ce, co, NULL, anis, kappa, mu, NULL, NULL};
  c.t1 = (double*)malloc(sizeof(double) * g->Vh * 24);
  c.t2 = (double*)malloc(sizeof(double) * g->Vh * 24);
  int rc = cg_generic(g->Vh, 24, sol, guess, op_kernsq, &c, niter, residue, source, st);
  free(c.t1); free(c.t2);
  return rc;
}

/* inv_tmD_cg_eoprec_native + the residual check of inv_tmD_cg_eoprec
   (cg_invert_tmD_eoprec.cpp:37-58,81-150,192-212) */
int or_tm_solve_eoprec(const or_geom* g, double* sol_lx, const double* conf_lx, double kappa,
                       const double* anis, double mu, int niter, double residue,
                       const double* source_lx, or_cg_stats* st, double* check)
{
  const int64_t Vh = g->Vh, n = Vh * 24;
  double* ce = (double*)malloc(sizeof(double) * Vh * 72);
  double* co = (double*)malloc(sizeof(double) * Vh * 72);
  double* se = (double*)malloc(sizeof(double) * n);
  double* so = (double*)malloc(sizeof(double) * n);
  double* xe = (double*)malloc(sizeof(double) * n);
  double* xo = (double*)malloc(sizeof(double) * n);
  double* et = (double*)malloc(sizeof(double) * n);
  double* phi = (double*)malloc(sizeof(double) * n);
  double* ot = (double*)malloc(sizeof(double) * n);
  or_split_lx(g, 72, conf_lx, ce, co);
  or_split_lx(g, 24, source_lx, se, so);
  /* eq. 8a */
  d_inv_tmDee(g, et, kappa, mu, se);
  /* eq. 8b: phi = g5 (eta_o + 1/2 H_oe et) */
  d_hop(g, 1, phi, ce, co, anis, et);
  for (int64_t i = 0; i < n; i++)
    phi[i] = ((i % 24) < 12) ? so[i] + phi[i] * 0.5 : -so[i] - phi[i] * 0.5;
  /* eq. 9 */
  int rc = or_tm_cg_kernsq(g, ot, NULL, ce, co, kappa, anis, mu, niter, residue, phi, st);
  /* eq. 10 */
  d_kern(g, xo, xe, ce, co, kappa, anis, -mu, ot);
  /* eq. 11 */
  d_hop(g, 0, et, ce, co, anis, xo);
  for (int64_t i = 0; i < n; i++) et[i] = se[i] + et[i] * 0.5;
  d_inv_tmDee(g, xe, kappa, mu, et);
  or_paste_lx(g, 24, sol_lx, xe, xo);
  if (check) {
    double* rv = (double*)malloc(sizeof(double) * g->V * 24);
    or_apply_tmQ(g, rv, conf_lx, kappa, anis, mu, sol_lx);
    for (int64_t i = 0; i < g->V * 24; i++) {
      const double gq = ((i % 24) < 12) ? rv[i] : -rv[i]; /* gamma5 */
      rv[i] = gq - source_lx[i];
    }
    *check = or_scalar_prod(g->V, 24, rv, rv) / or_scalar_prod(g->V, 24, source_lx, source_lx);
    free(rv);
  }
  free(ce); free(co); free(se); free(so); free(xe); free(xo); free(et); free(phi); free(ot);
  return rc;
}

static void op_tmQ2(void* v, double* out, const double* in)
{
  tm_ctx* c = (tm_ctx*)v;
  or_apply_tmQ2(c->g, out, c->conf_lx, c->kappa, c->t1, c->mu, in);
}

int or_tmQ2_cg(const or_geom* g, double* sol, const double* guess, const double* conf_lx, double kappa,
               double mu, int niter, double residue, const double* source, or_cg_stats* st)
{
  tm_ctx c = {g, NULL, NULL, conf_lx, NULL, kappa, mu, NULL, NULL};
  c.t1 = (double*)malloc(sizeof(double) * g->V * 24);
  int rc = cg_generic(g->V, 24, sol, guess, op_tmQ2, &c, niter, residue, source, st);
  free(c.t1);
  return rc;
}

typedef struct {
  const or_geom* g;
  const double *ce, *co;
  double mass2;
  double* temp;
} st_ctx;

static void op_st(void* v, double* out, const double* in)
{
  st_ctx* c = (st_ctx*)v;
  or_stD2ee_m2(c->g, out, c->ce, c->co, c->temp, c->mass2, in);
}

int or_st_cg(const or_geom* g, double* sol, const double* ce, const double* co, double mass2,
             int niter, double residue, const double* source, or_cg_stats* st)
{
  st_ctx c = {g, ce, co, mass2, NULL};
  c.temp = (double*)malloc(sizeof(double) * g->Vh * 6);
  int rc = cg_generic(g->Vh, 6, sol, NULL, op_st, &c, niter, residue, source, st);
  free(c.temp);
  return rc;
}

/* ---- generic CGM (cgm_invert_template_threaded.hpp:13-295) ---- */
typedef void (*or_shift_op)(void* ctx, double* out, double shift, const double* in);

static void sop_st(void* v, double* out, double shift, const double* in)
{
  st_ctx* c = (st_ctx*)v;
  or_stD2ee_m2(c->g, out, c->ce, c->co, c->temp, shift, in);
}
static void sop_tmQ2(void* v, double* out, double shift, const double* in)
{
  tm_ctx* c = (tm_ctx*)v;
  /* apply_tmQ2_m2: mu = sqrt(m2) (dirac_operator_tmQ2.hpp:10-13) */
  or_apply_tmQ2(c->g, out, c->conf_lx, c->kappa, c->t1, sqrt(shift), in);
}

static int cgm_generic(int64_t nsites, int nd, double** sol, int nShift, const double* shifts,
                       or_shift_op f, void* ctx, int niter_max, const double* residue,
                       const double* source, or_cg_stats* st, double* true_res)
{
  const int64_t n = nsites * nd;
  double* s = (double*)malloc(sizeof(double) * n);
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double** ps = (double**)malloc(sizeof(double*) * nShift);
  double *zps = malloc(sizeof(double) * nShift), *zas = malloc(sizeof(double) * nShift),
         *alphas = malloc(sizeof(double) * nShift), *zfs = malloc(sizeof(double) * nShift),
         *betas = malloc(sizeof(double) * nShift), *final_res = malloc(sizeof(double) * nShift);
  int* run_flag = malloc(sizeof(int) * nShift);
  int rc = 0;
  for (int k = 0; k < nShift; k++) {
    ps[k] = (double*)malloc(sizeof(double) * n);
    memcpy(ps[k], source, sizeof(double) * n);
    memset(sol[k], 0, sizeof(double) * n);
  }
  memcpy(p, source, sizeof(double) * n);
  memcpy(r, source, sizeof(double) * n);
  const double source_norm2 = or_scalar_prod(nsites, nd, r, r);
  if (source_norm2 == 0 || isnan(source_norm2)) { rc = -1; goto done; }
  int nrun = nShift, iter = 0, final_iter = 0;
  double betaa = 1, alpha = 0, rr = source_norm2, rfrf, betap;
  for (int k = 0; k < nShift; k++) { zps[k] = zas[k] = 1; alphas[k] = 0; run_flag[k] = 1; final_res[k] = 1; }
  do {
    final_iter = ++iter;
    f(ctx, s, 0.0, p);
    const double pap = or_scalar_prod(nsites, nd, p, s);
    betap = betaa;
    betaa = -rr / pap;
    for (int k = 0; k < nShift; k++)
      if (run_flag[k]) {
        const double ratio = betap / (betaa * alpha * (1 - zas[k] / zps[k]) + betap * (1 - shifts[k] * betaa));
        if (isnan(ratio)) { rc = -2; goto done; }
        zfs[k] = zas[k] * ratio;
        betas[k] = betaa * ratio;
        double* x = sol[k];
        const double* q = ps[k];
        const double b = betas[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) x[i] -= b * q[i];
      }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) r[i] += s[i] * betaa;
    rfrf = or_scalar_prod(nsites, nd, r, r);
    alpha = rfrf / rr;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = r[i] + p[i] * alpha;
    for (int k = 0; k < nShift; k++)
      if (run_flag[k]) {
        alphas[k] = alpha * zfs[k] * betas[k] / (zas[k] * betaa);
        double* q = ps[k];
        const double a = alphas[k], z = zfs[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) q[i] = r[i] * z + q[i] * a;
        zps[k] = zas[k];
        zas[k] = zfs[k];
      }
    rr = rfrf;
    for (int k = 0; k < nShift; k++)
      if (run_flag[k]) {
        final_res[k] = rr * zfs[k] * zfs[k] / source_norm2;
        if (final_res[k] < residue[k]) { run_flag[k] = 0; nrun--; }
      }
  } while (nrun > 0 && iter < niter_max);
  for (int k = 0; k < nShift; k++) {
    f(ctx, s, shifts[k], sol[k]);
    for (int64_t i = 0; i < n; i++) s[i] -= source[i];
    if (true_res) true_res[k] = or_scalar_prod(nsites, nd, s, s) / source_norm2;
  }
  if (st) { st->iters = final_iter; st->converged = final_iter != niter_max; st->rel_res_sq = true_res ? true_res[0] : 0; }
  if (final_iter == niter_max) rc = -3;
done:
  for (int k = 0; k < nShift; k++) free(ps[k]);
  free(ps); free(s); free(r); free(p);
  free(zps); free(zas); free(alphas); free(zfs); free(betas); free(final_res); free(run_flag);
  return rc;
}

int or_st_cgm(const or_geom* g, double** sols, const double* ce, const double* co, int nshift,
              const double* shifts, int niter, const double* residues, const double* source,
              or_cg_stats* st, double* true_res)
{
  st_ctx c = {g, ce, co, 0, NULL};
  c.temp = (double*)malloc(sizeof(double) * g->Vh * 6);
  int rc = cgm_generic(g->Vh, 6, sols, nshift, shifts, sop_st, &c, niter, residues, source, st, true_res);
  free(c.temp);
  return rc;
}

int or_tmQ2_cgm(const or_geom* g, double** sols, const double* conf_lx, double kappa, int nmass,
                const double* masses, int niter, const double* residues, const double* source,
                or_cg_stats* st, double* true_res)
{
  tm_ctx c = {g, NULL, NULL, conf_lx, NULL, kappa, 0, NULL, NULL};
  c.t1 = (double*)malloc(sizeof(double) * g->V * 24);
  double* m2 = (double*)malloc(sizeof(double) * nmass);
  for (int k = 0; k < nmass; k++) m2[k] = masses[k] * masses[k];
  int rc = cgm_generic(g->V, 24, sols, nmass, m2, sop_tmQ2, &c, niter, residues, source, st, true_res);
  free(m2); free(c.t1);
  return rc;
}

int or_st_rat_summ(const or_geom* g, double* chi, const double* ce, const double* co, int ndeg,
                   double cons, const double* poles, const double* weights, int niter,
                   double residue, const double* source, or_cg_stats* st)
{
  const int64_t n = g->Vh * 6;
  double** t = (double**)malloc(sizeof(double*) * ndeg);
  double* res = (double*)malloc(sizeof(double) * ndeg);
  for (int k = 0; k < ndeg; k++) { t[k] = (double*)malloc(sizeof(double) * n); res[k] = residue; }
  int rc = or_st_cgm(g, t, ce, co, ndeg, poles, niter, res, source, st, NULL);
  for (int64_t i = 0; i < n; i++) chi[i] = source[i] * cons;
  for (int k = 0; k < ndeg; k++)
    for (int64_t i = 0; i < n; i++) chi[i] += weights[k] * t[k][i];
  for (int k = 0; k < ndeg; k++) free(t[k]);
  free(t); free(res);
  return rc;
}

/* ---- quad-precision defect correction (cg_128_invert_template_threaded.cpp:56-95) ---- */
int or_tm_cg128_kernsq(const or_geom* g, double* sol, const double* ce, const double* co,
                       double kappa, const double* anis, double mu, double target,
                       const double* source, double* final_res, int* outer)
{
  const int64_t n = g->Vh * 24;
  __float128* solq = (__float128*)calloc(n, sizeof(__float128));
  __float128* resq = (__float128*)malloc(sizeof(__float128) * n);
  __float128* t1 = (__float128*)malloc(sizeof(__float128) * n);
  __float128* t2 = (__float128*)malloc(sizeof(__float128) * n);
  double* isrc = (double*)malloc(sizeof(double) * n);
  double* d = (double*)malloc(sizeof(double) * n);
  const double source_norm = or_scalar_prod(g->Vh, 24, source, source);
  double prev = 1, cur = 1;
  int quit = 0, ext = 0, rc = 0;
  do {
    q_kernsq(g, resq, t1, t2, ce, co, kappa, anis, mu, solq);
    __float128 acc = 0;
    for (int64_t i = 0; i < n; i++) {
      resq[i] = (__float128)source[i] - resq[i];
      isrc[i] = (double)resq[i];
      acc += resq[i] * resq[i];
    }
    cur = (double)(acc / source_norm);
    const double inner = fmax(1e-16, target / cur);
    if (cur >= target) {
      or_cg_stats st;
      int r = or_tm_cg_kernsq(g, d, NULL, ce, co, kappa, anis, mu, 100000, inner, isrc, &st);
      if (r == -1 || r == -2) { rc = r; break; }
      for (int64_t i = 0; i < n; i++) solq[i] += d[i];
    }
    if (ext != 0 && !(cur < prev)) quit = 1; /* not converging */
    prev = cur;
    ext++;
  } while (cur >= target && !quit);
  for (int64_t i = 0; i < n; i++) sol[i] = (double)solq[i];
  if (final_res) *final_res = cur;
  if (outer) *outer = ext;
  free(solq); free(resq); free(t1); free(t2); free(isrc); free(d);
  return rc;
}

/* ---- helpers ---- */
static void su3_mul(double* c, const double* a, const double* b) /* c = a b */
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        const double ar = a[(i * 3 + k) * 2], ai = a[(i * 3 + k) * 2 + 1];
        const double br = b[(k * 3 + j) * 2], bi = b[(k * 3 + j) * 2 + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      c[(i * 3 + j) * 2] = re;
      c[(i * 3 + j) * 2 + 1] = im;
    }
}
static void su3_mul_dag(double* c, const double* a, const double* b) /* c = a b^+ */
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        const double ar = a[(i * 3 + k) * 2], ai = a[(i * 3 + k) * 2 + 1];
        const double br = b[(j * 3 + k) * 2], bi = -b[(j * 3 + k) * 2 + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      c[(i * 3 + j) * 2] = re;
      c[(i * 3 + j) * 2 + 1] = im;
    }
}

/* plaquette.hpp:52-84: (temporal + spatial)/2, each Re tr summed / (V * 9) */
double or_plaquette_lx(const or_geom* g, const double* conf)
{
  double ts[2] = {0, 0};
  for (int64_t A = 0; A < g->V; A++)
    for (int mu = 0; mu < 4; mu++) {
      const int64_t B = g->up[A * 4 + mu];
      for (int nu = mu + 1; nu < 4; nu++) {
        const int64_t C = g->up[A * 4 + nu];
        double ABD[18], ACD[18], P[18];
        su3_mul(ABD, conf + (A * 4 + mu) * 18, conf + (B * 4 + nu) * 18);
        su3_mul(ACD, conf + (A * 4 + nu) * 18, conf + (C * 4 + mu) * 18);
        su3_mul_dag(P, ABD, ACD);
        ts[mu != 0 && nu != 0] += P[0] + P[8] + P[16];
      }
    }
  return (ts[0] / (g->V * 9.0) + ts[1] / (g->V * 9.0)) / 2;
}

void or_gauge_transform_lx(const or_geom* g, double* conf, const double* om)
{
  double tmp[18];
  for (int64_t x = 0; x < g->V; x++)
    for (int mu = 0; mu < 4; mu++) {
      double* U = conf + (x * 4 + mu) * 18;
      su3_mul(tmp, om + x * 18, U);
      su3_mul_dag(U, tmp, om + (int64_t)g->up[x * 4 + mu] * 18);
    }
}

void or_color_rotate_lx(const or_geom* g, int nspin, double* psi, const double* om)
{
  for (int64_t x = 0; x < g->V; x++)
    for (int s = 0; s < nspin; s++) {
      double* v = psi + (x * nspin + s) * 6;
      double o[6];
      for (int i = 0; i < 3; i++) {
        double re = 0, im = 0;
        for (int k = 0; k < 3; k++) {
          const double ar = om[x * 18 + (i * 3 + k) * 2], ai = om[x * 18 + (i * 3 + k) * 2 + 1];
          re += ar * v[2 * k] - ai * v[2 * k + 1];
          im += ar * v[2 * k + 1] + ai * v[2 * k];
        }
        o[2 * i] = re;
        o[2 * i + 1] = im;
      }
      memcpy(v, o, sizeof o);
    }
}

/* ===================================================================== twisted clover
 * clover_term (clover_term.hpp:88-96): chromo_operator (clover_term.cpp:58-104) from the
 * four leaves (topological_charge.hpp:139-190), times -(cSW + 1e-16)/4
 * (chromo_operator_include_cSW, clover_term.hpp:70-83).  Per lx site: 4 su3
 * [G, H, I, J] = [iA, B + iC, iD, E + iF] with A = P3 - P2, B = P4 + P1, C = P5 - P0,
 * D = P3 + P2, E = P4 - P1, F = P5 + P0 and P_i = (Q_i - Q_i^+)/4, Q the sum of the four
 * leaves of plane i (edge order 01,02,03,12,13,23, init.cpp:705-713).  Output [V][4][3][3][2].
 */
static void su3_dag_mul(double* c, const double* a, const double* b) /* c = a^+ b */
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        const double ar = a[(k * 3 + i) * 2], ai = -a[(k * 3 + i) * 2 + 1];
        const double br = b[(k * 3 + j) * 2], bi = b[(k * 3 + j) * 2 + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      c[(i * 3 + j) * 2] = re;
      c[(i * 3 + j) * 2 + 1] = im;
    }
}

void or_clover_term_lx(const or_geom* g, double* cl, const double* conf, double csw)
{
  const double f = -(csw + 1e-16) / 4;
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->V; X++) {
    double P[6][18];
    int iedge = 0;
    for (int mu = 0; mu < 4; mu++) {
      const int64_t A = g->up[X * 4 + mu], D = g->dw[X * 4 + mu];
      for (int nu = mu + 1; nu < 4; nu++, iedge++) {
        const int64_t B = g->up[X * 4 + nu], F = g->dw[X * 4 + nu];
        const int64_t Cs = g->up[D * 4 + nu], E = g->dw[D * 4 + nu], G = g->dw[A * 4 + nu];
#define LK(s, d) (conf + ((s) * 4 + (d)) * 18)
        double t1[18], t2[18], Q[18], L[18];
        /* leaf 1: U_mu(X) U_nu(A) U_mu(B)^+ U_nu(X)^+ */
        su3_mul(t1, LK(X, mu), LK(A, nu));
        su3_mul_dag(t2, t1, LK(B, mu));
        su3_mul_dag(Q, t2, LK(X, nu));
        /* leaf 2: U_nu(X) U_mu(C)^+ U_nu(D)^+ U_mu(D) */
        su3_mul_dag(t1, LK(X, nu), LK(Cs, mu));
        su3_mul_dag(t2, t1, LK(D, nu));
        su3_mul(L, t2, LK(D, mu));
        for (int k = 0; k < 18; k++) Q[k] += L[k];
        /* leaf 3: U_mu(D)^+ U_nu(E)^+ U_mu(E) U_nu(F) */
        /* U_mu(D)^+ U_nu(E)^+ = (U_nu(E) U_mu(D))^+ */
        double t0[18];
        su3_mul(t0, LK(E, nu), LK(D, mu));
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            t1[(i * 3 + j) * 2] = t0[(j * 3 + i) * 2];
            t1[(i * 3 + j) * 2 + 1] = -t0[(j * 3 + i) * 2 + 1];
          }
        su3_mul(t2, t1, LK(E, mu));
        su3_mul(L, t2, LK(F, nu));
        for (int k = 0; k < 18; k++) Q[k] += L[k];
        /* leaf 4: U_nu(F)^+ U_mu(F) U_nu(G) U_mu(X)^+ */
        su3_dag_mul(t1, LK(F, nu), LK(F, mu));
        su3_mul(t2, t1, LK(G, nu));
        su3_mul_dag(L, t2, LK(X, mu));
        for (int k = 0; k < 18; k++) Q[k] += L[k];
#undef LK
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            P[iedge][(i * 3 + j) * 2] = (Q[(i * 3 + j) * 2] - Q[(j * 3 + i) * 2]) / 4;
            P[iedge][(i * 3 + j) * 2 + 1] = (Q[(i * 3 + j) * 2 + 1] + Q[(j * 3 + i) * 2 + 1]) / 4;
          }
      }
    }
    double* o = cl + X * 72;
    for (int k = 0; k < 9; k++) {
      const int r = 2 * k, i = 2 * k + 1;
      const double Ar = P[3][r] - P[2][r], Ai = P[3][i] - P[2][i];
      const double Br = P[4][r] + P[1][r], Bi = P[4][i] + P[1][i];
      const double Cr = P[5][r] - P[0][r], Ci = P[5][i] - P[0][i];
      const double Dr = P[3][r] + P[2][r], Di = P[3][i] + P[2][i];
      const double Er = P[4][r] - P[1][r], Ei = P[4][i] - P[1][i];
      const double Fr = P[5][r] + P[0][r], Fi = P[5][i] + P[0][i];
      /* G = iA, H = B + iC, I = iD, J = E + iF */
      o[0 * 18 + r] = -Ai * f; o[0 * 18 + i] = Ar * f;
      o[1 * 18 + r] = (Br - Ci) * f; o[1 * 18 + i] = (Bi + Cr) * f;
      o[2 * 18 + r] = -Di * f; o[2 * 18 + i] = Dr * f;
      o[3 * 18 + r] = (Er - Fi) * f; o[3 * 18 + i] = (Ei + Fr) * f;
    }
  }
}

/* 3x3 complex matrix (row-major [3][3][2]) times colour, optionally daggered, accumulated
   with sign s into o */
static inline void cl_mv(d_cx o[3], const double* M, int dag, double s, const d_cx v[3])
{
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double mr = dag ? M[(b * 3 + a) * 2] : M[(a * 3 + b) * 2];
      const double mi = dag ? -M[(b * 3 + a) * 2 + 1] : M[(a * 3 + b) * 2 + 1];
      o[a].re += s * (mr * v[b].re - mi * v[b].im);
      o[a].im += s * (mr * v[b].im + mi * v[b].re);
    }
}

/* apply_point_twisted_clover_term_to_halfspincolor (clover_term.hpp:114-137) on the
   spin pair (off, off+1) with the blocks c0 = Cl[off], c1 = Cl[off+1]:
   o0 = z i0 + c0 i0 + c1^+ i1 ; o1 = z i1 + c1 i0 - c0 i1, z = 1/(2 kappa) + i mass */
static void cl_block(double* out, const double* cl, double kappa, double mass, int off, const double* in)
{
  d_cx i0[3], i1[3], o0[3], o1[3];
  col_load(i0, in + off * 6);
  col_load(i1, in + (off + 1) * 6);
  const double zr = 1 / (2 * kappa), zi = mass;
  for (int c = 0; c < 3; c++) {
    o0[c].re = zr * i0[c].re - zi * i0[c].im; o0[c].im = zr * i0[c].im + zi * i0[c].re;
    o1[c].re = zr * i1[c].re - zi * i1[c].im; o1[c].im = zr * i1[c].im + zi * i1[c].re;
  }
  cl_mv(o0, cl + off * 18, 0, +1, i0);
  cl_mv(o0, cl + (off + 1) * 18, 1, +1, i1);
  cl_mv(o1, cl + (off + 1) * 18, 0, +1, i0);
  cl_mv(o1, cl + off * 18, 0, -1, i1);
  for (int c = 0; c < 3; c++) {
    out[off * 6 + 2 * c] = o0[c].re; out[off * 6 + 2 * c + 1] = o0[c].im;
    out[(off + 1) * 6 + 2 * c] = o1[c].re; out[(off + 1) * 6 + 2 * c + 1] = o1[c].im;
  }
}

/* tmclovDee_or_oo_eos (dirac_operator_tmclovD_eoprec_portable.cpp:19-42): upper spins with
   +mu, lower with -mu, mu -> -mu for the dagger; cl: [Vh][4][3][3][2] of that parity */
void or_tmclovDee(const or_geom* g, double* out, const double* cl, double kappa, double mu, int dag,
                  const double* in)
{
  if (dag) mu = -mu;
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->Vh; X++) {
    cl_block(out + X * 24, cl + X * 72, kappa, +mu, 0, in + X * 24);
    cl_block(out + X * 24, cl + X * 72, kappa, -mu, 2, in + X * 24);
  }
}

/* invert_point_twisted_clover_term (clover_term.hpp:160-247): for each chirality block
   hl the inverse of z + [[C(2hl), C(2hl+1)^+], [C(2hl+1), -C(2hl)]], z = 1/(2 kappa) + i mass
   (the same +mass for both blocks), laid out as inv[hl][id][ic][x_id][x_ic] (row (id,ic),
   column (x_id,x_ic)).  The reference solves each column by CG on the squared block to a
   relative residue of 1e-32; here Gauss-Jordan elimination with partial pivoting gives the
   same matrix to rounding.  inv: [Vh][2][2][3][2][3][2]. */
void or_tmclov_invert(const or_geom* g, double* inv, const double* cl, double kappa, double mass)
{
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->Vh; X++)
    for (int hl = 0; hl < 2; hl++) {
      /* build the 6x6 block column by column from unit vectors */
      double M[6][6][2], Iv[6][6][2];
      for (int col = 0; col < 6; col++) {
        double e[24] = {0}, o[24];
        e[(2 * hl + col / 3) * 6 + (col % 3) * 2] = 1;
        cl_block(o, cl + X * 72, kappa, mass, 2 * hl, e);
        for (int row = 0; row < 6; row++) {
          M[row][col][0] = o[(2 * hl + row / 3) * 6 + (row % 3) * 2];
          M[row][col][1] = o[(2 * hl + row / 3) * 6 + (row % 3) * 2 + 1];
          Iv[row][col][0] = row == col;
          Iv[row][col][1] = 0;
        }
      }
      for (int c = 0; c < 6; c++) {
        int p = c;
        double best = -1;
        for (int r = c; r < 6; r++) {
          const double a = M[r][c][0] * M[r][c][0] + M[r][c][1] * M[r][c][1];
          if (a > best) { best = a; p = r; }
        }
        if (p != c)
          for (int k = 0; k < 6; k++)
            for (int ri = 0; ri < 2; ri++) {
              double t = M[c][k][ri]; M[c][k][ri] = M[p][k][ri]; M[p][k][ri] = t;
              t = Iv[c][k][ri]; Iv[c][k][ri] = Iv[p][k][ri]; Iv[p][k][ri] = t;
            }
        /* scale row c by 1/M[c][c] */
        const double dr = M[c][c][0], di = M[c][c][1], dd = dr * dr + di * di;
        const double qr = dr / dd, qi = -di / dd;
        for (int k = 0; k < 6; k++) {
          double a = M[c][k][0], b = M[c][k][1];
          M[c][k][0] = a * qr - b * qi; M[c][k][1] = a * qi + b * qr;
          a = Iv[c][k][0]; b = Iv[c][k][1];
          Iv[c][k][0] = a * qr - b * qi; Iv[c][k][1] = a * qi + b * qr;
        }
        for (int r = 0; r < 6; r++) {
          if (r == c) continue;
          const double fr = M[r][c][0], fi = M[r][c][1];
          for (int k = 0; k < 6; k++) {
            M[r][k][0] -= fr * M[c][k][0] - fi * M[c][k][1];
            M[r][k][1] -= fr * M[c][k][1] + fi * M[c][k][0];
            Iv[r][k][0] -= fr * Iv[c][k][0] - fi * Iv[c][k][1];
            Iv[r][k][1] -= fr * Iv[c][k][1] + fi * Iv[c][k][0];
          }
        }
      }
      double* o = inv + (X * 2 + hl) * 72;
      for (int row = 0; row < 6; row++)
        for (int col = 0; col < 6; col++) {
          o[(row * 6 + col) * 2] = Iv[row][col][0];
          o[(row * 6 + col) * 2 + 1] = Iv[row][col][1];
        }
    }
}

/* inv_tmclovDee_or_oo_eos (dirac_operator_tmclovD_eoprec.hpp:22-43): upper spins times
   inv[0], lower times inv[1]^+; for the dagger lower times inv[1], upper times inv[0]^+ */
void or_inv_tmclovDee(const or_geom* g, double* out, const double* inv, int dag, const double* in)
{
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->Vh; X++)
    for (int hl = 0; hl < 2; hl++) {
      const int use_dag = dag ? (hl == 0) : (hl == 1);
      const double* M = inv + (X * 2 + hl) * 72;
      const double* v = in + X * 24 + hl * 12;
      double* o = out + X * 24 + hl * 12;
      for (int r = 0; r < 6; r++) {
        double re = 0, im = 0;
        for (int c = 0; c < 6; c++) {
          const double mr = use_dag ? M[(c * 6 + r) * 2] : M[(r * 6 + c) * 2];
          const double mi = use_dag ? -M[(c * 6 + r) * 2 + 1] : M[(r * 6 + c) * 2 + 1];
          re += mr * v[2 * c] - mi * v[2 * c + 1];
          im += mr * v[2 * c + 1] + mi * v[2 * c];
        }
        o[2 * r] = re;
        o[2 * r + 1] = im;
      }
    }
}

/* tmclovDkern_eoprec_eos (dirac_operator_tmclovD_eoprec_portable.cpp:45-70):
   out = g5 (Doo(dag) in - 1/4 H_oe Dee^-1(dag) H_eo in) */
void or_tmclovDkern(const or_geom* g, double* out, double* temp, const double* ce, const double* co,
                    double kappa, const double* anis, const double* cl_odd, const double* inv_evn, int dag,
                    double mu, const double* in)
{
  d_hop(g, 0, out, ce, co, anis, in);
  or_inv_tmclovDee(g, temp, inv_evn, dag, out);
  d_hop(g, 1, out, ce, co, anis, temp);
  or_tmclovDee(g, temp, cl_odd, kappa, mu, dag, in);
  d_put_tog(g, out, temp);
}

/* tmclovDkern_eoprec_square_eos (..._portable.cpp:73-86): K(dag) then K */
void or_tmclovDkern_square(const or_geom* g, double* out, double* t1, double* t2, const double* ce,
                           const double* co, double kappa, const double* anis, const double* cl_odd,
                           const double* inv_evn, double mu, const double* in)
{
  or_tmclovDkern(g, t1, t2, ce, co, kappa, anis, cl_odd, inv_evn, 1, mu, in);
  or_tmclovDkern(g, out, t2, ce, co, kappa, anis, cl_odd, inv_evn, 0, mu, t1);
}

/* apply_tmclovQ (dirac_operator_tmclovQ.cpp:14-190) = apply_tmQ + g5 (chromo(Cl) in),
   the chromo part as unsafe_apply_point_chromo_operator_to_spincolor
   (clover_term.hpp:14-31): the twisted clover blocks without the diagonal */
void or_apply_tmclovQ(const or_geom* g, double* out, const double* conf, double kappa, const double* anis,
                      const double* cl_lx, double mu, const double* in)
{
  or_apply_tmQ(g, out, conf, kappa, anis, mu, in);
#pragma omp parallel for schedule(static)
  for (int64_t X = 0; X < g->V; X++) {
    double c[24] = {0};
    const double* v = in + X * 24;
    for (int off = 0; off < 4; off += 2) {
      d_cx i0[3], i1[3], o0[3] = {{0, 0}}, o1[3] = {{0, 0}};
      col_load(i0, v + off * 6);
      col_load(i1, v + (off + 1) * 6);
      cl_mv(o0, cl_lx + X * 72 + off * 18, 0, +1, i0);
      cl_mv(o0, cl_lx + X * 72 + (off + 1) * 18, 1, +1, i1);
      cl_mv(o1, cl_lx + X * 72 + (off + 1) * 18, 0, +1, i0);
      cl_mv(o1, cl_lx + X * 72 + off * 18, 0, -1, i1);
      for (int k = 0; k < 3; k++) {
        c[off * 6 + 2 * k] = o0[k].re; c[off * 6 + 2 * k + 1] = o0[k].im;
        c[(off + 1) * 6 + 2 * k] = o1[k].re; c[(off + 1) * 6 + 2 * k + 1] = o1[k].im;
      }
    }
    for (int i = 0; i < 24; i++) out[X * 24 + i] += i < 12 ? c[i] : -c[i];
  }
}

typedef struct {
  const or_geom* g;
  const double *ce, *co, *anis, *cl_odd, *inv_evn;
  double kappa, mu;
  double *t1, *t2;
} clov_ctx;

static void op_clov_kernsq(void* v, double* out, const double* in)
{
  clov_ctx* c = (clov_ctx*)v;
  or_tmclovDkern_square(c->g, out, c->t1, c->t2, c->ce, c->co, c->kappa, c->anis, c->cl_odd, c->inv_evn, c->mu, in);
}

/* inv_tmclovD_cg_eoprec_native + the residual check of inv_tmclovD_cg_eoprec
   (cg_invert_tmclovD_eoprec.cpp:56-128,179-232) */
int or_tmclov_solve_eoprec(const or_geom* g, double* sol_lx, const double* conf_lx, double kappa,
                           const double* anis, double csw, double mu, int niter, double residue,
                           const double* source_lx, or_cg_stats* st, double* check)
{
  const int64_t V = g->V, Vh = g->Vh, n = Vh * 24;
  double* cl = (double*)malloc(sizeof(double) * V * 72);
  double* cle = (double*)malloc(sizeof(double) * Vh * 72);
  double* clo = (double*)malloc(sizeof(double) * Vh * 72);
  double* inv = (double*)malloc(sizeof(double) * Vh * 144);
  double* ce = (double*)malloc(sizeof(double) * Vh * 72);
  double* co = (double*)malloc(sizeof(double) * Vh * 72);
  double* se = (double*)malloc(sizeof(double) * n);
  double* so = (double*)malloc(sizeof(double) * n);
  double* xe = (double*)malloc(sizeof(double) * n);
  double* xo = (double*)malloc(sizeof(double) * n);
  double* et = (double*)malloc(sizeof(double) * n);
  double* phi = (double*)malloc(sizeof(double) * n);
  double* ot = (double*)malloc(sizeof(double) * n);
  or_clover_term_lx(g, cl, conf_lx, csw);
  or_split_lx(g, 72, cl, cle, clo);
  or_tmclov_invert(g, inv, cle, kappa, mu);
  or_split_lx(g, 72, conf_lx, ce, co);
  or_split_lx(g, 24, source_lx, se, so);
  /* eq. 8a */
  or_inv_tmclovDee(g, et, inv, 0, se);
  /* eq. 8b: phi = g5 (eta_o + 1/2 H_oe et) */
  d_hop(g, 1, phi, ce, co, anis, et);
  for (int64_t i = 0; i < n; i++)
    phi[i] = ((i % 24) < 12) ? so[i] + phi[i] * 0.5 : -so[i] - phi[i] * 0.5;
  /* eq. 9 */
  clov_ctx c = {g, ce, co, anis, clo, inv, kappa, mu, NULL, NULL};
  c.t1 = (double*)malloc(sizeof(double) * n);
  c.t2 = (double*)malloc(sizeof(double) * n);
  int rc = cg_generic(Vh, 24, ot, NULL, op_clov_kernsq, &c, niter, residue, phi, st);
  /* eq. 10 */
  or_tmclovDkern(g, xo, xe, ce, co, kappa, anis, clo, inv, 1, mu, ot);
  /* eq. 11 */
  d_hop(g, 0, et, ce, co, anis, xo);
  for (int64_t i = 0; i < n; i++) et[i] = se[i] + et[i] * 0.5;
  or_inv_tmclovDee(g, xe, inv, 0, et);
  or_paste_lx(g, 24, sol_lx, xe, xo);
  if (check) {
    double* rv = (double*)malloc(sizeof(double) * V * 24);
    or_apply_tmclovQ(g, rv, conf_lx, kappa, anis, cl, mu, sol_lx);
    for (int64_t i = 0; i < V * 24; i++) {
      const double gq = ((i % 24) < 12) ? rv[i] : -rv[i];
      rv[i] = gq - source_lx[i];
    }
    *check = or_scalar_prod(V, 24, rv, rv) / or_scalar_prod(V, 24, source_lx, source_lx);
    free(rv);
  }
  free(c.t1); free(c.t2);
  free(cl); free(cle); free(clo); free(inv); free(ce); free(co); free(se); free(so); free(xe); free(xo);
  free(et); free(phi); free(ot);
  return rc;
}

/* ===================================================================== rooted staggered force
 * summ_the_rootst_eoimpr_quark_force (hmc/fermions/rootst_eoimpr_quark_force.cpp:17-85):
 * the links get the background field and the staggered phases (add_backfield_with_stagphases_to_conf,
 * backfield.cpp:248-284; eta_mu(x) = prod_{nu<mu} (-1)^{x_nu}, geometry_lx.cpp:703-724), the
 * multi-shift CG gives chi_k = (p_k - 1/4 Deo Doe)^-1 pf for the poles p_k, v_k = Doe chi_k
 * (apply_stDoe = 1/2 apply_st2Doe), and per link
 *   F_e[x][mu] += w_k eta u1b_e v_k(x+mu) chi_k(x)^+ ,  F_o[x][mu] -= w_k eta u1b_o chi_k(x+mu) v_k(x)^+ .
 * ce/co: links without phases; u1b: [Vh][4][2] per parity; F: [Vh][4][3][3][2] per parity, summed into. */
static int stag_eta(const int x[4], int mu)
{
  int ph = 1;
  for (int nu = 1; nu <= mu; nu++) ph *= 1 - 2 * (x[nu - 1] % 2);
  return ph;
}

int or_rootst_force(const or_geom* g, double* Fe, double* Fo, const double* ce, const double* co,
                    const double* u1e, const double* u1o, const double* pf, int ndeg, const double* poles,
                    const double* weights, int niter, double residue)
{
  const int64_t Vh = g->Vh;
  const double* C[2] = {ce, co};
  const double* U1[2] = {u1e, u1o};
  double* P[2];
  for (int p = 0; p < 2; p++) {
    P[p] = (double*)malloc(sizeof(double) * Vh * 72);
    for (int64_t i = 0; i < Vh; i++) {
      int x[4];
      coord_of_lx(g->L, g->lx_of_eo[p][i], x);
      for (int mu = 0; mu < 4; mu++) {
        const double fr = U1[p][(i * 4 + mu) * 2] * stag_eta(x, mu), fi = U1[p][(i * 4 + mu) * 2 + 1] * stag_eta(x, mu);
        const double* u = C[p] + (i * 4 + mu) * 18;
        double* o = P[p] + (i * 4 + mu) * 18;
        for (int k = 0; k < 9; k++) {
          o[2 * k] = u[2 * k] * fr - u[2 * k + 1] * fi;
          o[2 * k + 1] = u[2 * k] * fi + u[2 * k + 1] * fr;
        }
      }
    }
  }
  double** chi = (double**)malloc(sizeof(double*) * ndeg);
  double* res = (double*)malloc(sizeof(double) * ndeg);
  for (int k = 0; k < ndeg; k++) {
    chi[k] = (double*)malloc(sizeof(double) * Vh * 6);
    res[k] = residue;
  }
  or_cg_stats st;
  int rc = or_st_cgm(g, chi, P[0], P[1], ndeg, poles, niter, res, pf, &st, NULL);
  double* v = (double*)malloc(sizeof(double) * Vh * 6);
  double* F[2] = {Fe, Fo};
  for (int k = 0; k < ndeg; k++) {
    or_st_2Doe(g, v, P[0], P[1], chi[k]);
    for (int64_t i = 0; i < Vh * 6; i++) v[i] *= 0.5;
    for (int p = 0; p < 2; p++)
      for (int64_t i = 0; i < Vh; i++) {
        int x[4];
        coord_of_lx(g->L, g->lx_of_eo[p][i], x);
        for (int mu = 0; mu < 4; mu++) {
          const int64_t n = g->eoup[p][i * 4 + mu];
          const double* a = p == 0 ? v + n * 6 : chi[k] + n * 6;
          const double* b = p == 0 ? chi[k] + i * 6 : v + i * 6;
          const double s = (p == 0 ? 1.0 : -1.0) * weights[k] * stag_eta(x, mu);
          const double ur = U1[p][(i * 4 + mu) * 2], ui = U1[p][(i * 4 + mu) * 2 + 1];
          double* f = F[p] + (i * 4 + mu) * 18;
          for (int c1 = 0; c1 < 3; c1++)
            for (int c2 = 0; c2 < 3; c2++) {
              /* a_c1 conj(b_c2) u1b */
              const double tr = a[2 * c1] * b[2 * c2] + a[2 * c1 + 1] * b[2 * c2 + 1];
              const double ti = a[2 * c1 + 1] * b[2 * c2] - a[2 * c1] * b[2 * c2 + 1];
              f[(c1 * 3 + c2) * 2] += s * (tr * ur - ti * ui);
              f[(c1 * 3 + c2) * 2 + 1] += s * (tr * ui + ti * ur);
            }
        }
      }
  }
  for (int k = 0; k < ndeg; k++) free(chi[k]);
  free(chi); free(res); free(v); free(P[0]); free(P[1]);
  return rc;
}
