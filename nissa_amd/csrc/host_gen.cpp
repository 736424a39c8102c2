// Host-side harness helpers: seeded synthetic gauge fields, boundary and
// staggered phases, sources, lx<->eo reordering.  Pure host code (no HIP calls),
// so it is callable on machines without a GPU.
//
// Gauge construction follows nissa's "hot" random link recipe in distribution:
// H traceless Hermitian built from four complex Gaussians of unit complex variance
// (herm_put_to_gauss, src/base/random.hpp:311-338), then U = exp(i eps H)
// (safe_hermitian_exact_i_exponentiate, src/new_types/su3_op.hpp:2880-3031).  The
// random stream is our own: Philox4x32-10 (Salmon et al., SC'11) keyed by the seed
// and counted by (site, direction), so every host generates identical bytes.
// The exponential is computed by scaling and squaring of a Taylor series and then
// projected onto SU(3) (Gram-Schmidt of rows 0,1; row 2 = conj(row0 x row1)).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "nq_internal.hpp"

namespace nq {

// static-partition parallel loop over [0, n) on the host's hardware threads
template <class F>
static void parallel_for(int64_t n, F&& f)
{
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), (n + 4095) / 4096));
  if (nt == 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      const int64_t a = n * t / nt, b = n * (t + 1) / nt;
      for (int64_t i = a; i < b; i++) f(i);
    });
  for (auto& x : th) x.join();
}

struct Philox {
  static inline void round(uint32_t c[4], const uint32_t k[2])
  {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
  // 10-round Philox4x32 on counter (c0..c3) with key (k0,k1)
  static inline void gen(uint32_t out[4], uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t key)
  {
    uint32_t c[4] = {c0, c1, c2, c3};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    for (int r = 0; r < 10; r++) {
      round(c, k);
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; i++) out[i] = c[i];
  }
};

// uniform in (0,1) with 53 random bits
static inline double u53(uint32_t a, uint32_t b)
{
  const uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  return ((double)(m & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}

// n Gaussian reals N(0, sigma^2) for stream (seed, a, b)
static void gauss_stream(uint64_t seed, uint64_t a, uint32_t b, double sigma, int n, double* out)
{
  for (int i = 0; i < n; i += 2) {
    uint32_t r[4];
    Philox::gen(r, (uint32_t)a, (uint32_t)(a >> 32), b, (uint32_t)(i / 2), seed);
    const double u1 = u53(r[0], r[1]), u2 = u53(r[2], r[3]);
    const double rad = std::sqrt(-2.0 * std::log(u1)) * sigma;
    const double th = 2.0 * M_PI * u2;
    out[i] = rad * std::cos(th);
    if (i + 1 < n) out[i + 1] = rad * std::sin(th);
  }
}

typedef double m3[3][3][2];

static void mat_mul(m3 c, const m3 a, const m3 b)
{
  m3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += a[i][k][0] * b[k][j][0] - a[i][k][1] * b[k][j][1];
        im += a[i][k][0] * b[k][j][1] + a[i][k][1] * b[k][j][0];
      }
      t[i][j][0] = re;
      t[i][j][1] = im;
    }
  std::memcpy(c, t, sizeof(m3));
}

// exp(A) for a 3x3 complex A with ||A|| small-to-moderate: scaling and squaring
static void mat_exp(m3 out, const m3 A)
{
  double nrm = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) nrm += A[i][j][0] * A[i][j][0] + A[i][j][1] * A[i][j][1];
  nrm = std::sqrt(nrm);
  int s = 0;
  while (nrm > 0.25) { nrm *= 0.5; s++; }
  const double scale = std::ldexp(1.0, -s);
  m3 B, term, acc;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int r = 0; r < 2; r++) {
        B[i][j][r] = A[i][j][r] * scale;
        acc[i][j][r] = (i == j && r == 0) ? 1.0 : 0.0;
        term[i][j][r] = acc[i][j][r];
      }
  for (int k = 1; k <= 18; k++) {
    mat_mul(term, term, B);
    const double inv = 1.0 / k;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int r = 0; r < 2; r++) {
          term[i][j][r] *= inv;
          acc[i][j][r] += term[i][j][r];
        }
  }
  for (int k = 0; k < s; k++) mat_mul(acc, acc, acc);
  std::memcpy(out, acc, sizeof(m3));
}

// project onto SU(3): orthonormalise rows 0,1, row 2 = conj(row0 x row1)
static void su3_project(m3 U)
{
  double n0 = 0;
  for (int j = 0; j < 3; j++) n0 += U[0][j][0] * U[0][j][0] + U[0][j][1] * U[0][j][1];
  n0 = 1.0 / std::sqrt(n0);
  for (int j = 0; j < 3; j++) { U[0][j][0] *= n0; U[0][j][1] *= n0; }
  // c = <row0, row1> = sum conj(u0) u1
  double cr = 0, ci = 0;
  for (int j = 0; j < 3; j++) {
    cr += U[0][j][0] * U[1][j][0] + U[0][j][1] * U[1][j][1];
    ci += U[0][j][0] * U[1][j][1] - U[0][j][1] * U[1][j][0];
  }
  for (int j = 0; j < 3; j++) {
    U[1][j][0] -= cr * U[0][j][0] - ci * U[0][j][1];
    U[1][j][1] -= cr * U[0][j][1] + ci * U[0][j][0];
  }
  double n1 = 0;
  for (int j = 0; j < 3; j++) n1 += U[1][j][0] * U[1][j][0] + U[1][j][1] * U[1][j][1];
  n1 = 1.0 / std::sqrt(n1);
  for (int j = 0; j < 3; j++) { U[1][j][0] *= n1; U[1][j][1] *= n1; }
  for (int j = 0; j < 3; j++) {
    const int a = (j + 1) % 3, b = (j + 2) % 3;
    // (u0 x u1)_j = u0_a u1_b - u0_b u1_a ; row2 = conj of it
    const double re = (U[0][a][0] * U[1][b][0] - U[0][a][1] * U[1][b][1]) -
                      (U[0][b][0] * U[1][a][0] - U[0][b][1] * U[1][a][1]);
    const double im = (U[0][a][0] * U[1][b][1] + U[0][a][1] * U[1][b][0]) -
                      (U[0][b][0] * U[1][a][1] + U[0][b][1] * U[1][a][0]);
    U[2][j][0] = re;
    U[2][j][1] = -im;
  }
}

// H from 8 reals with N(0,1/2) each, same placement as herm_put_to_gauss
static void algebra(uint64_t seed, int64_t lx, int mu, m3 H)
{
  double r[8];
  gauss_stream(seed, (uint64_t)lx, (uint32_t)(mu + 1), std::sqrt(0.5), 8, r);
  const double s3 = 1.0 / std::sqrt(3.0);
  std::memset(H, 0, sizeof(m3));
  H[0][0][0] = r[2] + s3 * r[7];
  H[1][1][0] = -r[2] + s3 * r[7];
  H[2][2][0] = -2.0 * s3 * r[7];
  H[0][1][0] = H[1][0][0] = r[0];
  H[1][0][1] = r[1];
  H[0][1][1] = -r[1];
  H[0][2][0] = H[2][0][0] = r[3];
  H[2][0][1] = r[4];
  H[0][2][1] = -r[4];
  H[1][2][0] = H[2][1][0] = r[5];
  H[2][1][1] = r[6];
  H[1][2][1] = -r[6];
}

static void link_of(uint64_t seed, int64_t lx, int mu, double eps, m3 H, m3 U)
{
  algebra(seed, lx, mu, H);
  m3 A;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      // A = i eps H
      A[i][j][0] = -eps * H[i][j][1];
      A[i][j][1] = eps * H[i][j][0];
    }
  mat_exp(U, A);
  su3_project(U);
}

static bool dims_ok(int T, int LX, int LY, int LZ)
{
  const int L[4] = {T, LX, LY, LZ};
  for (int i = 0; i < 4; i++)
    if (L[i] < 2 || (L[i] & 1)) return false;
  return true;
}

}  // namespace nq

using namespace nq;

extern "C" int nq_host_link_algebra(uint64_t seed, int64_t lx, int mu, double eps, double* H18, double* U18)
{
  if (!H18 || !U18 || mu < 0 || mu > 3) return nq::fail(NQ_EINVAL, "nq_host_link_algebra: bad argument");
  m3 H, U;
  link_of(seed, lx, mu, eps, H, U);
  std::memcpy(H18, H, sizeof(m3));
  std::memcpy(U18, U, sizeof(m3));
  return NQ_OK;
}

extern "C" int nq_host_gauge_random_lx(int T, int LX, int LY, int LZ, uint64_t seed, double eps, double* conf)
{
  if (!conf || !dims_ok(T, LX, LY, LZ)) return nq::fail(NQ_EINVAL, "nq_host_gauge_random_lx: bad dims/pointer");
  const int64_t V = (int64_t)T * LX * LY * LZ;
  parallel_for(V, [&](int64_t x) {
    for (int mu = 0; mu < 4; mu++) {
      m3 H, U;
      link_of(seed, x, mu, eps, H, U);
      std::memcpy(conf + (x * 4 + mu) * 18, U, sizeof(m3));
    }
  });
  return NQ_OK;
}

extern "C" int nq_host_gauge_phase_lx(int T, int LX, int LY, int LZ, const double* theta, double* conf)
{
  if (!conf || !theta || !dims_ok(T, LX, LY, LZ)) return nq::fail(NQ_EINVAL, "nq_host_gauge_phase_lx: bad argument");
  const int L[4] = {T, LX, LY, LZ};
  double c[4], s[4];
  for (int mu = 0; mu < 4; mu++) {
    c[mu] = std::cos(theta[mu] * M_PI / L[mu]);
    s[mu] = std::sin(theta[mu] * M_PI / L[mu]);
  }
  const int64_t V = (int64_t)T * LX * LY * LZ;
  parallel_for(V, [&](int64_t x) {
    for (int mu = 0; mu < 4; mu++) {
      double* U = conf + (x * 4 + mu) * 18;
      for (int k = 0; k < 9; k++) {
        const double re = U[2 * k], im = U[2 * k + 1];
        U[2 * k] = re * c[mu] - im * s[mu];
        U[2 * k + 1] = re * s[mu] + im * c[mu];
      }
    }
  });
  return NQ_OK;
}

extern "C" int nq_host_gauge_stagphases_lx(int T, int LX, int LY, int LZ, double* conf)
{
  if (!conf || !dims_ok(T, LX, LY, LZ)) return nq::fail(NQ_EINVAL, "nq_host_gauge_stagphases_lx: bad argument");
  const int L[4] = {T, LX, LY, LZ};
  const int64_t V = (int64_t)T * LX * LY * LZ;
  parallel_for(V, [&](int64_t x) {
    int c[4];
    int64_t r = x;
    for (int mu = 3; mu >= 0; mu--) { c[mu] = (int)(r % L[mu]); r /= L[mu]; }
    // eta_0 = 1, eta_mu = prod_{nu<mu} (-1)^x_nu  (geometry_lx.cpp:703-713)
    int eta = 1;
    for (int mu = 0; mu < 4; mu++) {
      if (mu > 0) eta *= (c[mu - 1] & 1) ? -1 : 1;
      if (eta < 0) {
        double* U = conf + (x * 4 + mu) * 18;
        for (int k = 0; k < 18; k++) U[k] = -U[k];
      }
    }
  });
  return NQ_OK;
}

extern "C" int nq_host_source_lx(int T, int LX, int LY, int LZ, uint64_t seed, int kind, int ncomp,
                                 int parity_mask, int64_t site0, int sc, double* out)
{
  if (!out || !dims_ok(T, LX, LY, LZ) || (ncomp != 12 && ncomp != 3) || parity_mask < 1 || parity_mask > 3 ||
      kind < 0 || kind > 2)
    return nq::fail(NQ_EINVAL, "nq_host_source_lx: bad argument");
  const int L[4] = {T, LX, LY, LZ};
  const int64_t V = (int64_t)T * LX * LY * LZ;
  const int nd = 2 * ncomp;
  std::memset(out, 0, sizeof(double) * V * nd);
  if (kind == 2) {
    if (site0 < 0 || site0 >= V || sc < 0 || sc >= ncomp) return nq::fail(NQ_EINVAL, "nq_host_source_lx: bad point");
    out[site0 * nd + 2 * sc] = 1.0;
    return NQ_OK;
  }
  parallel_for(V, [&](int64_t x) {
    int c[4];
    int64_t r = x;
    for (int mu = 3; mu >= 0; mu--) { c[mu] = (int)(r % L[mu]); r /= L[mu]; }
    const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
    if (!((parity_mask >> par) & 1)) return;
    if (kind == 0) {
      uint32_t rr[4];
      for (int k = 0; k < ncomp; k += 4) {
        Philox::gen(rr, (uint32_t)x, (uint32_t)(x >> 32), 0x5A5A0000u, (uint32_t)k, seed);
        for (int j = 0; j < 4 && k + j < ncomp; j++) out[x * nd + 2 * (k + j)] = (rr[j] & 1) ? 1.0 : -1.0;
      }
    } else {
      gauss_stream(seed, (uint64_t)x, 0xA5A50000u, std::sqrt(0.5), nd, out + x * nd);
    }
  });
  return NQ_OK;
}

extern "C" int nq_host_split_lx(int T, int LX, int LY, int LZ, int nd, const double* lx, double* evn, double* odd)
{
  if (!lx || !evn || !odd || nd <= 0 || !dims_ok(T, LX, LY, LZ)) return nq::fail(NQ_EINVAL, "nq_host_split_lx: bad argument");
  const int L[4] = {T, LX, LY, LZ};
  const int64_t V = (int64_t)T * LX * LY * LZ;
  parallel_for(V, [&](int64_t x) {
    int c[4];
    int64_t r = x;
    for (int mu = 3; mu >= 0; mu--) { c[mu] = (int)(r % L[mu]); r /= L[mu]; }
    double* d = ((c[0] + c[1] + c[2] + c[3]) & 1) ? odd : evn;
    std::memcpy(d + (x >> 1) * nd, lx + x * nd, sizeof(double) * nd);
  });
  return NQ_OK;
}

extern "C" int nq_host_paste_lx(int T, int LX, int LY, int LZ, int nd, double* lx, const double* evn, const double* odd)
{
  if (!lx || !evn || !odd || nd <= 0 || !dims_ok(T, LX, LY, LZ)) return nq::fail(NQ_EINVAL, "nq_host_paste_lx: bad argument");
  const int L[4] = {T, LX, LY, LZ};
  const int64_t V = (int64_t)T * LX * LY * LZ;
  parallel_for(V, [&](int64_t x) {
    int c[4];
    int64_t r = x;
    for (int mu = 3; mu >= 0; mu--) { c[mu] = (int)(r % L[mu]); r /= L[mu]; }
    const double* s = ((c[0] + c[1] + c[2] + c[3]) & 1) ? odd : evn;
    std::memcpy(lx + x * nd, s + (x >> 1) * nd, sizeof(double) * nd);
  });
  return NQ_OK;
}
