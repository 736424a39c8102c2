// Twisted-clover site-local pieces: the clover term built from the gauge field,
// the inverse of the twisted clover blocks on one parity, and their application.
//
// Clover field of one parity: 36 complex FP64 components per site in the SoA
// record layout of the spinor fields (component k of device site e at
// ((cx<double>*)data)[k*stride + e]): the four 3x3 blocks G, H, I, J of nissa's
// clover_term_t (clover_term.cpp:22-55) row-major, k = 9*b + 3*row + col.
// The twisted clover operator on a site acts on the chirality halves h = 0
// (spins 0,1) and h = 1 (spins 2,3) as the 6x6 block
//   z_h + [[C(2h), C(2h+1)^+], [C(2h+1), -C(2h)]],  z_h = 1/(2 kappa) + i mu_h
// (apply_point_twisted_clover_term_to_halfspincolor, clover_term.hpp:114-137).
// Inverse field: 72 complex per site, block h row-major (k = 36 h + 6 row + col,
// row/col = 3 spin + colour), nissa's inv_clover_term_t[h][id][ic][x_id][x_ic].

#include "clover_block.hpp"

namespace nq {

namespace {

typedef cx<double> C;

__device__ __forceinline__ void m3_mul(C c[9], const C a[9], const C b[9])  // c = a b
{
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C s = a[3 * i] * b[j];
      s = s + a[3 * i + 1] * b[3 + j];
      s = s + a[3 * i + 2] * b[6 + j];
      c[3 * i + j] = s;
    }
}
__device__ __forceinline__ void m3_mul_dag(C c[9], const C a[9], const C b[9])  // c = a b^+
{
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C s = a[3 * i] * conjg(b[3 * j]);
      s = s + a[3 * i + 1] * conjg(b[3 * j + 1]);
      s = s + a[3 * i + 2] * conjg(b[3 * j + 2]);
      c[3 * i + j] = s;
    }
}
__device__ __forceinline__ void m3_dag_mul(C c[9], const C a[9], const C b[9])  // c = a^+ b
{
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C s = conjg(a[i]) * b[j];
      s = s + conjg(a[3 + i]) * b[3 + j];
      s = s + conjg(a[6 + i]) * b[6 + j];
      c[3 * i + j] = s;
    }
}

// link U_mu at lattice coordinates (periodic wrap) from the nissa eo AoS copies
struct Links {
  const double* aos[2];
  int L[4];
  __device__ __forceinline__ void get(int t, int x, int y, int z, int mu, C u[9]) const
  {
    t = (t + L[0]) % L[0];
    x = (x + L[1]) % L[1];
    y = (y + L[2]) % L[2];
    z = (z + L[3]) % L[3];
    const int64_t lx = (((int64_t)t * L[1] + x) * L[2] + y) * L[3] + z;
    const double* p = aos[(t + x + y + z) & 1] + (lx >> 1) * 72 + mu * 18;
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = mk<double>(p[2 * k], p[2 * k + 1]);
  }
};

// clover_term (clover_term.hpp:88-96): the four leaves of each plane
// (four_leaves_point, topological_charge.hpp:139-190, planes 01,02,03,12,13,23),
// P = (Q - Q^+)/4, then G = i(P3 - P2), H = (P4 + P1) + i(P5 - P0),
// I = i(P3 + P2), J = (P4 - P1) + i(P5 + P0) (point_chromo_operator,
// clover_term.cpp:56-98), times f = -(cSW + 1e-16)/4.
__global__ void __launch_bounds__(128) k_clover_build(Geo g, Links lk, int par, double f, SpinorView out)
{
  const uint32_t e = blockIdx.x * 128 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const Site s = site_of(g, par, e);
  const int X[4] = {s.t, s.x, s.y, s.z};
  C G[9], H[9], I[9], J[9];
#pragma unroll
  for (int k = 0; k < 9; k++) G[k] = H[k] = I[k] = J[k] = mk<double>(0, 0);
  int plane = 0;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = mu + 1; nu < 4; nu++, plane++) {
      int A[4] = {X[0], X[1], X[2], X[3]}, B[4] = {X[0], X[1], X[2], X[3]};
      int D[4] = {X[0], X[1], X[2], X[3]}, F[4] = {X[0], X[1], X[2], X[3]};
      A[mu] += 1;
      B[nu] += 1;
      D[mu] -= 1;
      F[nu] -= 1;
      int Cc[4] = {D[0], D[1], D[2], D[3]}, E[4] = {D[0], D[1], D[2], D[3]}, Gs[4] = {A[0], A[1], A[2], A[3]};
      Cc[nu] += 1;
      E[nu] -= 1;
      Gs[nu] -= 1;
      C u1[9], u2[9], t1[9], t2[9], Q[9], L[9];
      // leaf 1: U_mu(X) U_nu(A) U_mu(B)^+ U_nu(X)^+
      lk.get(X[0], X[1], X[2], X[3], mu, u1);
      lk.get(A[0], A[1], A[2], A[3], nu, u2);
      m3_mul(t1, u1, u2);
      lk.get(B[0], B[1], B[2], B[3], mu, u1);
      m3_mul_dag(t2, t1, u1);
      lk.get(X[0], X[1], X[2], X[3], nu, u1);
      m3_mul_dag(Q, t2, u1);
      // leaf 2: U_nu(X) U_mu(C)^+ U_nu(D)^+ U_mu(D)
      lk.get(X[0], X[1], X[2], X[3], nu, u1);
      lk.get(Cc[0], Cc[1], Cc[2], Cc[3], mu, u2);
      m3_mul_dag(t1, u1, u2);
      lk.get(D[0], D[1], D[2], D[3], nu, u1);
      m3_mul_dag(t2, t1, u1);
      lk.get(D[0], D[1], D[2], D[3], mu, u1);
      m3_mul(L, t2, u1);
#pragma unroll
      for (int k = 0; k < 9; k++) Q[k] = Q[k] + L[k];
      // leaf 3: U_mu(D)^+ U_nu(E)^+ U_mu(E) U_nu(F) = (U_nu(E) U_mu(D))^+ U_mu(E) U_nu(F)
      lk.get(E[0], E[1], E[2], E[3], nu, u1);
      lk.get(D[0], D[1], D[2], D[3], mu, u2);
      m3_mul(t1, u1, u2);
      lk.get(E[0], E[1], E[2], E[3], mu, u1);
      m3_dag_mul(t2, t1, u1);
      lk.get(F[0], F[1], F[2], F[3], nu, u1);
      m3_mul(L, t2, u1);
#pragma unroll
      for (int k = 0; k < 9; k++) Q[k] = Q[k] + L[k];
      // leaf 4: U_nu(F)^+ U_mu(F) U_nu(G) U_mu(X)^+
      lk.get(F[0], F[1], F[2], F[3], nu, u1);
      lk.get(F[0], F[1], F[2], F[3], mu, u2);
      m3_dag_mul(t1, u1, u2);
      lk.get(Gs[0], Gs[1], Gs[2], Gs[3], nu, u1);
      m3_mul(t2, t1, u1);
      lk.get(X[0], X[1], X[2], X[3], mu, u1);
      m3_mul_dag(L, t2, u1);
#pragma unroll
      for (int k = 0; k < 9; k++) Q[k] = Q[k] + L[k];
      // P = (Q - Q^+)/4 and its contribution to G, H, I, J
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          const C q = Q[3 * a + b], qt = Q[3 * b + a];
          const C P = mk<double>((q.re - qt.re) / 4, (q.im + qt.im) / 4);
          const C iP = mk<double>(-P.im, P.re);
          const int k = 3 * a + b;
          switch (plane) {
          case 0: H[k] = H[k] - iP; J[k] = J[k] + iP; break;
          case 1: H[k] = H[k] + P; J[k] = J[k] - P; break;
          case 2: G[k] = G[k] - iP; I[k] = I[k] + iP; break;
          case 3: G[k] = G[k] + iP; I[k] = I[k] + iP; break;
          case 4: H[k] = H[k] + P; J[k] = J[k] + P; break;
          default: H[k] = H[k] + iP; J[k] = J[k] + iP; break;
          }
        }
    }
  C* o = (C*)out.data;
  const int64_t st = g.stride;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    o[(0 + k) * st + e] = scal(G[k], f);
    o[(9 + k) * st + e] = scal(H[k], f);
    o[(18 + k) * st + e] = scal(I[k], f);
    o[(27 + k) * st + e] = scal(J[k], f);
  }
}

// invert_point_twisted_clover_term (clover_term.hpp:160-247): the inverse of
// each block z + [[C0, C1^+], [C1, -C0]], z = 1/(2 kappa) + i mass (the same
// mass for both blocks), by Gauss-Jordan elimination with partial pivoting (the
// reference solves the columns by CG on the squared block; the matrix is the same
// to rounding).
__global__ void __launch_bounds__(64) k_clover_invert(Geo g, SpinorView cl, SpinorView inv, double kappa, double mass)
{
  const uint32_t e = blockIdx.x * 64 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const C* c = (const C*)cl.data;
  C* o = (C*)inv.data;
  const int64_t st = g.stride;
  const C z = mk<double>(1 / (2 * kappa), mass);
  for (int h = 0; h < 2; h++) {
    C M[6][6], V[6][6];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        const C c0 = c[(18 * h + 3 * a + b) * st + e], c1 = c[(18 * h + 9 + 3 * a + b) * st + e];
        const C c1t = conjg(c[(18 * h + 9 + 3 * b + a) * st + e]);
        M[a][b] = c0;
        M[a][3 + b] = c1t;
        M[3 + a][b] = c1;
        M[3 + a][3 + b] = mk<double>(-c0.re, -c0.im);
      }
    for (int r = 0; r < 6; r++) {
      M[r][r] = M[r][r] + z;
      for (int k = 0; k < 6; k++) V[r][k] = mk<double>(r == k ? 1.0 : 0.0, 0.0);
    }
    for (int col = 0; col < 6; col++) {
      int p = col;
      double best = -1;
      for (int r = col; r < 6; r++) {
        const double a = M[r][col].re * M[r][col].re + M[r][col].im * M[r][col].im;
        if (a > best) { best = a; p = r; }
      }
      if (p != col)
        for (int k = 0; k < 6; k++) {
          C t = M[col][k]; M[col][k] = M[p][k]; M[p][k] = t;
          t = V[col][k]; V[col][k] = V[p][k]; V[p][k] = t;
        }
      const double dd = M[col][col].re * M[col][col].re + M[col][col].im * M[col][col].im;
      const C q = mk<double>(M[col][col].re / dd, -M[col][col].im / dd);
      for (int k = 0; k < 6; k++) {
        M[col][k] = M[col][k] * q;
        V[col][k] = V[col][k] * q;
      }
      for (int r = 0; r < 6; r++) {
        if (r == col) continue;
        const C fct = M[r][col];
        for (int k = 0; k < 6; k++) {
          M[r][k] = M[r][k] - fct * M[col][k];
          V[r][k] = V[r][k] - fct * V[col][k];
        }
      }
    }
    for (int r = 0; r < 6; r++)
      for (int k = 0; k < 6; k++) o[(36 * h + 6 * r + k) * st + e] = V[r][k];
  }
}

// mode 0 (tmclovDee_or_oo_eos, dirac_operator_tmclovD_eoprec_portable.cpp:19-42):
//   out_h = (z_h + B_h) in_h, mu_0 = +mu, mu_1 = -mu, mu -> -mu for the dagger;
// mode 1 (inv_tmclovDee_or_oo_eos, dirac_operator_tmclovD_eoprec.hpp:22-43):
//   out_h = inv_h in_h, with inv_h^+ for h = 1 (h = 0 for the dagger).
// out = a_h * (that) + b_h * aux (per chirality coefficients, aux optional).
// All block loads of a chirality are issued before its arithmetic; the two
// chiralities are kept apart so that their blocks are not live together.
struct ClovArgs {
  double a[2][2], b[2][2];
};

// FP16 spinors (FP16 blocks): one thread per site, both chiralities (the per-site
// norm of the output covers all twelve components)
template <int MODE, int DAG, int SP>
__global__ void __launch_bounds__(256) k_clover_apply(Geo g, SpinorView out, SpinorView in, SpinorView blk,
                                                      SpinorView aux, double kappa, double mu, ClovArgs k)
{
  using R = typename std::conditional<SP == SP_F64, double, float>::type;
  typedef cx<R> CR;
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  CR v[12], o[12];
  SpAcc<SP, R, 12>::load(in, g.stride, e, v);
  constexpr int BEB = SP == SP_F64 ? 16 : (SP == SP_F32 ? 8 : 4);  // block element bytes
#pragma unroll
  for (int h = 0; h < 2; h++) {
    clover_half<MODE, DAG, R, BEB>(blk.data, g.stride, e, h, kappa, mu, v + 6 * h, o + 6 * h);
#pragma unroll
    for (int q = 0; q < 6; q++) asm volatile("" : "+v"(o[6 * h + q].re), "+v"(o[6 * h + q].im));
  }
  if (aux.data) {
    CR w[12];
    SpAcc<SP, R, 12>::load(aux, g.stride, e, w);
#pragma unroll
    for (int q = 0; q < 12; q++) {
      const int h = q < 6 ? 0 : 1;
      o[q] = mk<R>((R)k.a[h][0], (R)k.a[h][1]) * o[q] + mk<R>((R)k.b[h][0], (R)k.b[h][1]) * w[q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 12; q++) {
      const int h = q < 6 ? 0 : 1;
      o[q] = mk<R>((R)k.a[h][0], (R)k.a[h][1]) * o[q];
    }
  }
  SpAcc<SP, R, 12>::store(out, g.stride, e, o);
}

// FP64 and FP32 spinors: two threads per site, one per chirality.  Wavefronts 2j
// and 2j+1 of a workgroup take chiralities 0 and 1 of the same 64 sites; the six
// components of a chirality are whole records (arrays 6h.. of FP64, 3h.. of FP32
// records), so each wavefront reads and writes only its half.  Fewer registers
// than the one-thread form (FP64 inverse: 168 vs 194 VGPRs) and more waves per
// SIMD: at 32^3x64 the FP64 inverse went from 368 to 333 us, D_ee from 224 to 214 us.
template <int MODE, int DAG, int SP>
__global__ void __launch_bounds__(256) k_clover_apply_h(Geo g, SpinorView out, SpinorView in, SpinorView blk,
                                                        SpinorView aux, double kappa, double mu, ClovArgs k)
{
  static_assert(SP == SP_F64 || SP == SP_F32, "chirality split needs whole records per half");
  using R = typename std::conditional<SP == SP_F64, double, float>::type;
  typedef cx<R> CR;
  constexpr int EB = SP == SP_F64 ? 16 : 8;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int h = wave & 1;
  const uint32_t e = (blockIdx.x * 2u + (uint32_t)(wave >> 1)) * 64u + (threadIdx.x & 63u);
  if (e >= (uint32_t)g.Vh) return;
  const int64_t hoff = (int64_t)h * 6 * g.stride * EB;
  SpinorView ih = in, oh = out;
  ih.data = (char*)in.data + hoff;
  oh.data = (char*)out.data + hoff;
  CR v[6], o[6];
  SpAcc<SP, R, 6>::load(ih, g.stride, e, v);
  clover_half<MODE, DAG, R, EB>(blk.data, g.stride, e, h, kappa, mu, v, o);
  const CR ka = mk<R>((R)k.a[h][0], (R)k.a[h][1]);
  if (aux.data) {
    SpinorView ah = aux;
    ah.data = (char*)aux.data + hoff;
    CR w[6];
    SpAcc<SP, R, 6>::load(ah, g.stride, e, w);
    const CR kb = mk<R>((R)k.b[h][0], (R)k.b[h][1]);
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = ka * o[q] + kb * w[q];
  } else {
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = ka * o[q];
  }
  SpAcc<SP, R, 6>::store(oh, g.stride, e, o);
}

// Batched forms for nr right-hand sides (views with rstride), out_r = block in_r, for
// the batched twisted clover solve.  A right-hand side whose solver state says idle is
// skipped (CgState done; with pred 1 no reliable update due, with pred 2 one due).  FP64 and
// FP32 spinors: a wavefront per chirality keeps the block in registers and applies
// it to every right-hand side; FP16: one thread per site, the block read per
// right-hand side through the cache (the per-site norm spans both chiralities).
__device__ __forceinline__ bool clov_idle(const CgState* st, int r, int pred)
{
  return st && (st[r].done || (pred == 1 && !st[r].reliable) || (pred == 2 && st[r].reliable));
}

template <int MODE, int DAG, int SP>
__global__ void __launch_bounds__(256) k_clover_apply_hb(Geo g, SpinorView out, SpinorView in, SpinorView blk,
                                                         double kappa, double mu, int nr, const CgState* st, int pred)
{
  static_assert(SP == SP_F64 || SP == SP_F32, "chirality split needs whole records per half");
  using R = typename std::conditional<SP == SP_F64, double, float>::type;
  typedef cx<R> CR;
  constexpr int EB = SP == SP_F64 ? 16 : 8;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int h = wave & 1;
  const uint32_t e = (blockIdx.x * 2u + (uint32_t)(wave >> 1)) * 64u + (threadIdx.x & 63u);
  if (e >= (uint32_t)g.Vh) return;
  const int64_t hoff = (int64_t)h * 6 * g.stride * EB;
  SpinorView ih = in, oh = out;
  ih.data = (char*)in.data + hoff;
  oh.data = (char*)out.data + hoff;
  ClovBlk<MODE, R> k;
  clover_load<MODE, R, EB>(blk.data, g.stride, e, h, k);
  for (int r = 0; r < nr; r++) {
    if (clov_idle(st, r, pred)) continue;  // uniform over the grid
    CR v[6], o[6];
    SpAcc<SP, R, 6>::load(ih, g.stride, e, v, r);
    clover_apply_blk<MODE, DAG, R>(k, h, kappa, mu, v, o);
    SpAcc<SP, R, 6>::store(oh, g.stride, e, o, r);
  }
}

template <int MODE, int DAG>
__global__ void __launch_bounds__(256) k_clover_apply_b16(Geo g, SpinorView out, SpinorView in, SpinorView blk,
                                                          double kappa, double mu, int nr, const CgState* st, int pred)
{
  typedef cx<float> CR;
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  for (int r = 0; r < nr; r++) {
    if (clov_idle(st, r, pred)) continue;
    CR v[12], o[12];
    SpAcc<SP_F16, float, 12>::load(in, g.stride, e, v, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      clover_half<MODE, DAG, float, 4>(blk.data, g.stride, e, h, kappa, mu, v + 6 * h, o + 6 * h);
#pragma unroll
      for (int q = 0; q < 6; q++) asm volatile("" : "+v"(o[6 * h + q].re), "+v"(o[6 * h + q].im));
    }
    SpAcc<SP_F16, float, 12>::store(out, g.stride, e, o, r);
  }
}

// FP32 or FP16 packed copy of an FP64 block field (sloppy solves): MODE 0 the clover
// blocks (36 complex per site), 1 the inverse blocks (72)
template <int MODE, int EB>
__global__ void __launch_bounds__(256) k_blocks_pack(Geo g, const C* in, void* out, int* out_of_range)
{
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const bool ok0 = clover_pack<MODE, EB>(in, out, g.stride, e, 0);
  const bool ok1 = clover_pack<MODE, EB>(in, out, g.stride, e, 1);
  if (!(ok0 && ok1)) *out_of_range = 1;  // every writer stores the same value
}

// Re tr of the six plaquettes U_mu(X) U_nu(X+mu) U_mu(X+nu)^+ U_nu(X)^+ of each
// site, summed over the temporal planes (mu = 0) and the spatial ones
// (global_plaquette_lx_conf, operations/gauge/plaquette.hpp:52-84); one partial
// per (parity, temporal/spatial, 64-site chunk)
__global__ void __launch_bounds__(64) k_plaquette(Geo g, Links lk, int par, int64_t nchunks, double* partial)
{
  const uint32_t e = blockIdx.x * 64 + threadIdx.x;
  double ts[2] = {0, 0};
  if (e < (uint32_t)g.Vh) {
    const Site s = site_of(g, par, e);
    const int X[4] = {s.t, s.x, s.y, s.z};
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        int A[4] = {X[0], X[1], X[2], X[3]}, B[4] = {X[0], X[1], X[2], X[3]};
        A[mu] += 1;
        B[nu] += 1;
        C u1[9], u2[9], t1[9], t2[9];
        lk.get(X[0], X[1], X[2], X[3], mu, u1);
        lk.get(A[0], A[1], A[2], A[3], nu, u2);
        m3_mul(t1, u1, u2);
        lk.get(X[0], X[1], X[2], X[3], nu, u1);
        lk.get(B[0], B[1], B[2], B[3], mu, u2);
        m3_mul(t2, u1, u2);
        double re = 0;  // Re tr(t1 t2^+) = sum_ij Re(t1_ij conj(t2_ij))
#pragma unroll
        for (int k = 0; k < 9; k++) re += t1[k].re * t2[k].re + t1[k].im * t2[k].im;
        ts[mu != 0] += re;
      }
  }
  store_partial(ts[0], partial, nchunks, 2 * par, blockIdx.x);
  store_partial(ts[1], partial, nchunks, 2 * par + 1, blockIdx.x);
}

}  // namespace

int launch_plaquette(nq_ctx* ctx, const double* aos_e, const double* aos_o, double ts[2])
{
  const Geo& g = ctx->geo;
  const int64_t nchunks = (g.Vh + 63) / 64;
  if (4 * nchunks > ctx->max_partials) return fail(NQ_EINVAL, "plaquette: reduction scratch too small");
  Links lk;
  lk.aos[0] = aos_e;
  lk.aos[1] = aos_o;
  lk.L[0] = g.T; lk.L[1] = g.LX; lk.L[2] = g.LY; lk.L[3] = g.LZ;
  for (int par = 0; par < 2; par++)
    hipLaunchKernelGGL(k_plaquette, dim3((unsigned)nchunks), dim3(64), 0, ctx->stream, g, lk, par, nchunks, ctx->partial);
  NQ_HIP(hipGetLastError());
  std::vector<double> h(4 * nchunks);
  NQ_HIP(hipMemcpyAsync(h.data(), ctx->partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  // fixed order: parity, then chunk
  ts[0] = ts[1] = 0;
  for (int par = 0; par < 2; par++)
    for (int k = 0; k < 2; k++)
      for (int64_t c = 0; c < nchunks; c++) ts[k] += h[(2 * par + k) * nchunks + c];
  return NQ_OK;
}

int launch_clover_build(nq_ctx* ctx, const double* aos_e, const double* aos_o, int par, double csw, const SpinorView& out)
{
  const Geo& g = ctx->geo;
  Links lk;
  lk.aos[0] = aos_e;
  lk.aos[1] = aos_o;
  lk.L[0] = g.T; lk.L[1] = g.LX; lk.L[2] = g.LY; lk.L[3] = g.LZ;
  const double f = -(csw + 1e-16) / 4;  // chromo_operator_include_cSW (clover_term.hpp:70-83)
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "clover_build", 0.0, &ev);
  hipLaunchKernelGGL(k_clover_build, dim3((unsigned)((g.Vh + 127) / 128)), dim3(128), 0, ctx->stream, g, lk, par, f, out);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_clover_invert(nq_ctx* ctx, const SpinorView& cl, const SpinorView& inv, double kappa, double mass)
{
  const Geo& g = ctx->geo;
  hipEvent_t ev = nullptr;
  prof_begin(ctx, "clover_invert", 0.0, &ev);
  hipLaunchKernelGGL(k_clover_invert, dim3((unsigned)((g.Vh + 63) / 64)), dim3(64), 0, ctx->stream, g, cl, inv, kappa, mass);
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_blocks_sloppy(nq_ctx* ctx, int ncomp, const SpinorView& in, const SpinorView& out)
{
  const Geo& g = ctx->geo;
  if (in.prec != SP_F64 || (out.prec != SP_F32 && out.prec != SP_F16) || (ncomp != 36 && ncomp != 72))
    return fail(NQ_EINVAL, "block conversion: FP64 -> FP32 or FP16, 36 or 72 components");
  if (ctx->capturing) return fail(NQ_EINVAL, "block conversion inside a graph capture");
  const dim3 grid((unsigned)((g.Vh + 255) / 256)), b256(256);
  const C* src = (const C*)in.data;
  int* flag = nullptr;
  NQ_TRY(ctx_temp(ctx, "blk.range", sizeof(int), (void**)&flag));
  NQ_HIP(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
  if (out.prec == SP_F32) {
    if (ncomp == 36) hipLaunchKernelGGL((k_blocks_pack<0, 8>), grid, b256, 0, ctx->stream, g, src, out.data, flag);
    else hipLaunchKernelGGL((k_blocks_pack<1, 8>), grid, b256, 0, ctx->stream, g, src, out.data, flag);
  } else {
    if (ncomp == 36) hipLaunchKernelGGL((k_blocks_pack<0, 4>), grid, b256, 0, ctx->stream, g, src, out.data, flag);
    else hipLaunchKernelGGL((k_blocks_pack<1, 4>), grid, b256, 0, ctx->stream, g, src, out.data, flag);
  }
  NQ_HIP(hipGetLastError());
  if (out.prec == SP_F16) {
    // an FP16 copy of blocks beyond its range (near-singular D_ee at a tiny twisted
    // mass) would turn the sloppy operator into infinities: refuse it instead
    int bad = 0;
    NQ_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    NQ_HIP(hipStreamSynchronize(ctx->stream));
    if (bad) return fail(NQ_EINVAL, "clover blocks beyond the FP16 range (|entry| > 65000): use FP32 inner iterations");
  }
  return NQ_OK;
}

int launch_clover_apply(nq_ctx* ctx, int mode, const SpinorView& out, const SpinorView& in, const SpinorView& blk,
                        const SpinorView& aux, double kappa, double mu, int dag, const double a[2][2],
                        const double b[2][2])
{
  const Geo& g = ctx->geo;
  const int sp = in.prec;
  if (out.prec != sp || blk.prec != sp || (aux.data && aux.prec != sp))
    return fail(NQ_EINVAL, "twisted clover: fields and blocks of one precision (FP64, FP32 or FP16)");
  ClovArgs k;
  for (int h = 0; h < 2; h++)
    for (int c = 0; c < 2; c++) {
      k.a[h][c] = a[h][c];
      k.b[h][c] = b ? b[h][c] : 0.0;
    }
  hipEvent_t ev = nullptr;
  const double cb = sp == SP_F64 ? 16.0 : (sp == SP_F32 ? 8.0 : 4.0);    // bytes per block complex
  const double sb = sp == SP_F64 ? 192.0 : (sp == SP_F32 ? 96.0 : 52.0);  // bytes per spinor site
  const double bytes = (double)g.Vh * ((mode == 0 ? 36 : 72) * cb + (2 + (aux.data ? 1 : 0)) * sb);
  const char* names[3][2] = {{"clover_dee", "clover_inv"}, {"clover_dee_f32", "clover_inv_f32"},
                             {"clover_dee_f16", "clover_inv_f16"}};
  prof_begin(ctx, names[sp == SP_F64 ? 0 : (sp == SP_F32 ? 1 : 2)][mode == 0 ? 0 : 1], bytes, &ev);
  const dim3 grid((unsigned)((g.Vh + 255) / 256)), grid_h((unsigned)((g.Vh + 127) / 128)), blkd(256);
#define NQ_CLOV(M, D, S)                                                                                        \
  do {                                                                                                          \
    if constexpr (S == SP_F16)                                                                                  \
      hipLaunchKernelGGL((k_clover_apply<M, D, S>), grid, blkd, 0, ctx->stream, g, out, in, blk, aux, kappa, mu, k); \
    else                                                                                                        \
      hipLaunchKernelGGL((k_clover_apply_h<M, D, S>), grid_h, blkd, 0, ctx->stream, g, out, in, blk, aux, kappa, mu, \
                         k);                                                                                    \
  } while (0)
  if (sp == SP_F64) {
    if (mode == 0) { if (dag) NQ_CLOV(0, 1, SP_F64); else NQ_CLOV(0, 0, SP_F64); }
    else { if (dag) NQ_CLOV(1, 1, SP_F64); else NQ_CLOV(1, 0, SP_F64); }
  } else if (sp == SP_F32) {
    if (mode == 0) { if (dag) NQ_CLOV(0, 1, SP_F32); else NQ_CLOV(0, 0, SP_F32); }
    else { if (dag) NQ_CLOV(1, 1, SP_F32); else NQ_CLOV(1, 0, SP_F32); }
  } else if (sp == SP_F16) {
    if (mode == 0) { if (dag) NQ_CLOV(0, 1, SP_F16); else NQ_CLOV(0, 0, SP_F16); }
    else { if (dag) NQ_CLOV(1, 1, SP_F16); else NQ_CLOV(1, 0, SP_F16); }
  } else {
    return fail(NQ_EINVAL, "twisted clover: unknown spinor precision");
  }
#undef NQ_CLOV
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

int launch_clover_apply_batched(nq_ctx* ctx, int mode, int dag, const SpinorView& out, const SpinorView& in,
                                const SpinorView& blk, double kappa, double mu, int nr, const CgState* st, int pred)
{
  const Geo& g = ctx->geo;
  const int sp = in.prec;
  if (out.prec != sp || blk.prec != sp || nr < 1 || nr > 16)
    return fail(NQ_EINVAL, "batched twisted clover: fields and blocks of one precision, 1..16 RHS");
  hipEvent_t ev = nullptr;
  const double cb = sp == SP_F64 ? 16.0 : (sp == SP_F32 ? 8.0 : 4.0);
  const double sb = sp == SP_F64 ? 192.0 : (sp == SP_F32 ? 96.0 : 52.0);
  prof_begin(ctx, mode == 0 ? "clover_dee_batch" : "clover_inv_batch",
             (double)g.Vh * ((mode == 0 ? 36 : 72) * cb + 2 * sb * nr), &ev);
  const dim3 grid((unsigned)((g.Vh + 255) / 256)), grid_h((unsigned)((g.Vh + 127) / 128)), b256(256);
#define NQ_CLB(M, D)                                                                                             \
  do {                                                                                                           \
    if (sp == SP_F64)                                                                                            \
      hipLaunchKernelGGL((k_clover_apply_hb<M, D, SP_F64>), grid_h, b256, 0, ctx->stream, g, out, in, blk, kappa, mu, \
                         nr, st, pred);                                                                          \
    else if (sp == SP_F32)                                                                                       \
      hipLaunchKernelGGL((k_clover_apply_hb<M, D, SP_F32>), grid_h, b256, 0, ctx->stream, g, out, in, blk, kappa, mu, \
                         nr, st, pred);                                                                          \
    else                                                                                                         \
      hipLaunchKernelGGL((k_clover_apply_b16<M, D>), grid, b256, 0, ctx->stream, g, out, in, blk, kappa, mu, nr, st, \
                         pred);                                                                                  \
  } while (0)
  if (mode == 0) { if (dag) NQ_CLB(0, 1); else NQ_CLB(0, 0); }
  else { if (dag) NQ_CLB(1, 1); else NQ_CLB(1, 0); }
#undef NQ_CLB
  NQ_HIP(hipGetLastError());
  prof_end(ctx, ev);
  return NQ_OK;
}

}  // namespace nq
