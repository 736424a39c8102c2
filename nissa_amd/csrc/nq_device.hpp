// Device-side building blocks for gfx950: complex arithmetic, eo geometry,
// SoA field accessors in FP64/FP32/FP16(+norm), link accessors (18/12 reals),
// and the deterministic block reduction with last-block finalisation.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "nq_internal.hpp"

namespace nq {

template <class R>
struct cx {
  R re, im;
};

template <class R>
__device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <class R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b)
{
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ cx<R> scal(cx<R> a, R s) { return {a.re * s, a.im * s}; }
// a + i b , a - i b
template <class R>
__device__ __forceinline__ cx<R> iadd(cx<R> a, cx<R> b) { return {a.re - b.im, a.im + b.re}; }
template <class R>
__device__ __forceinline__ cx<R> isub(cx<R> a, cx<R> b) { return {a.re + b.im, a.im - b.re}; }
template <class R>
__device__ __forceinline__ cx<R> conjg(cx<R> a) { return {a.re, -a.im}; }
// acc += a*b
template <class R>
__device__ __forceinline__ void fma_to(cx<R>& acc, cx<R> a, cx<R> b)
{
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
}

// ---------------------------------------------------------------- geometry
struct Site {
  int t, x, y, z, th, tl, xt, xi, yt, yi;
  uint32_t dlx;  // rank in the device site order
};

__device__ __forceinline__ Site site_of(const Geo& g, int par, uint32_t e)
{
  Site s;
  const uint32_t q = 2u * e;
  uint32_t r = q / (uint32_t)g.TB;
  const int tl0 = (int)(q - r * (uint32_t)g.TB);
  const uint32_t r1 = r / (uint32_t)g.LZ;
  s.z = (int)(r - r1 * (uint32_t)g.LZ);
  s.yi = (int)(r1 & (uint32_t)(g.TY - 1));
  const uint32_t r2 = r1 >> g.logTY;
  s.xi = (int)(r2 & (uint32_t)(g.TX - 1));
  const uint32_t r3 = r2 >> g.logTX;
  const uint32_t r4 = r3 / (uint32_t)g.NYT;
  s.yt = (int)(r3 - r4 * (uint32_t)g.NYT);
  const uint32_t r5 = r4 / (uint32_t)g.NXT;
  s.xt = (int)(r4 - r5 * (uint32_t)g.NXT);
  s.th = (int)r5;
  s.x = s.xt * g.TX + s.xi;
  s.y = s.yt * g.TY + s.yi;
  const int b = (par + tl0 + s.x + s.y + s.z) & 1;
  s.tl = tl0 + b;
  s.t = s.th * g.TB + s.tl;
  s.dlx = q + (uint32_t)b;
  return s;
}

// nissa lx index (t-major, z fastest) of a site
__device__ __forceinline__ int64_t nissa_lx(const Geo& g, int t, int x, int y, int z)
{
  return (((int64_t)t * g.LX + x) * g.LY + y) * g.LZ + z;
}

// device eo index of the neighbour in direction mu (0=t..3=z), up (fwd) or down
__device__ __forceinline__ uint32_t neigh_eo(const Geo& g, const Site& s, int mu, bool up)
{
  int64_t d = s.dlx;
  switch (mu) {
  case 0:
    if (up) d += s.tl == g.TB - 1 ? (1 - g.TB) + (s.th == g.NTH - 1 ? -(int64_t)(g.NTH - 1) * g.Sth : g.Sth) : 1;
    else d += s.tl == 0 ? (g.TB - 1) + (s.th == 0 ? (int64_t)(g.NTH - 1) * g.Sth : -g.Sth) : -1;
    break;
  case 1:
    if (up) d += s.xi < g.TX - 1 ? g.Sxi : -(int64_t)(g.TX - 1) * g.Sxi + (s.xt == g.NXT - 1 ? -(int64_t)(g.NXT - 1) * g.Sxt : g.Sxt);
    else d += s.xi > 0 ? -g.Sxi : (int64_t)(g.TX - 1) * g.Sxi + (s.xt == 0 ? (int64_t)(g.NXT - 1) * g.Sxt : -g.Sxt);
    break;
  case 2:
    if (up) d += s.yi < g.TY - 1 ? g.Syi : -(int64_t)(g.TY - 1) * g.Syi + (s.yt == g.NYT - 1 ? -(int64_t)(g.NYT - 1) * g.Syt : g.Syt);
    else d += s.yi > 0 ? -g.Syi : (int64_t)(g.TY - 1) * g.Syi + (s.yt == 0 ? (int64_t)(g.NYT - 1) * g.Syt : -g.Syt);
    break;
  default:
    d += up ? (s.z == g.LZ - 1 ? -(int64_t)(g.LZ - 1) * g.Sz : g.Sz) : (s.z == 0 ? (int64_t)(g.LZ - 1) * g.Sz : -g.Sz);
    break;
  }
  return (uint32_t)(d >> 1);
}

// nissa eo index (lx>>1) of the neighbour (layout conversion only)
__device__ __forceinline__ int64_t nissa_neigh_eo(const Geo& g, const Site& s, int mu, bool up)
{
  int c[4] = {s.t, s.x, s.y, s.z};
  const int L[4] = {g.T, g.LX, g.LY, g.LZ};
  c[mu] = up ? (c[mu] + 1 == L[mu] ? 0 : c[mu] + 1) : (c[mu] == 0 ? L[mu] - 1 : c[mu] - 1);
  return nissa_lx(g, c[0], c[1], c[2], c[3]) >> 1;
}

// ---------------------------------------------------------------- buffer loads
// Field and link reads go through raw buffer loads: the base is a uniform
// (SGPR) descriptor per array, the per-lane site offset is one 32-bit VGPR and
// the component offset k*stride is a uniform SGPR, so a 12-component load
// costs one address register instead of twelve 64-bit pointers.
typedef unsigned int nq_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int nq_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(uint32_t)bytes, 0x00020000);
}
// (the loaded vectors are bit-cast whole: extracting lanes of the integer vector
// first lets the compiler narrow a 64-bit load to a single dword)
typedef double nq_dv2 __attribute__((ext_vector_type(2)));
typedef float nq_fv2 __attribute__((ext_vector_type(2)));
// CP = cache policy bits of the load: 0 default, 2 non-temporal (streamed data
// that should not displace reusable lines, used for the links)
template <int CP = 0>
__device__ __forceinline__ double2 bload_d2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const nq_dv2 v = __builtin_bit_cast(nq_dv2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, CP));
  return make_double2(v.x, v.y);
}
template <int CP = 0>
__device__ __forceinline__ float2 bload_f2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const nq_fv2 v = __builtin_bit_cast(nq_fv2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CP));
  return make_float2(v.x, v.y);
}
template <int CP = 0>
__device__ __forceinline__ float2 bload_h2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, CP);
  return __half22float2(__builtin_bit_cast(__half2, v));
}
constexpr int kLinkCP = 2;
__device__ __forceinline__ float bload_f1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// ---------------------------------------------------------------- packed records
// Fields and link slots store NC complex elements of EB bytes per site in 16-byte
// records: element k belongs to group q = k / G (G = 16/EB elements per record).
// Group q < NG = NC/G is an array of 16-byte records over the sites; the NL = NC%G
// leftover elements form one array of NL*EB-byte records after them:
//   group q of site e:   base + q*stride*16 + e*16 + (k%G)*EB
//   leftover of site e:  base + NG*stride*16 + e*NL*EB + (k-NG*G)*EB
// A field of NC elements occupies NC*EB*stride bytes in any precision.  FP64
// (G = 1) is plain structure-of-arrays; FP32 packs 2 and FP16 4 complex numbers
// per record, so every lane issues 16-byte loads in all precisions.
template <int EB, int NC>
struct Pack {
  static constexpr int G = 16 / EB, NG = NC / G, NL = NC % G;
  __host__ __device__ static constexpr int64_t offset(int k, int64_t e, int64_t stride)
  {
    return k < NG * G ? (int64_t)(k / G) * stride * 16 + e * 16 + (k % G) * EB
                      : (int64_t)NG * stride * 16 + e * NL * EB + (k - NG * G) * EB;
  }
};

typedef float nq_fv4 __attribute__((ext_vector_type(4)));
typedef _Float16 nq_hv2 __attribute__((ext_vector_type(2)));
typedef _Float16 nq_hv4 __attribute__((ext_vector_type(4)));
typedef _Float16 nq_hv8 __attribute__((ext_vector_type(8)));
typedef unsigned int nq_u3 __attribute__((ext_vector_type(3)));

// the G = 16/EB complex numbers of one 16-byte record
template <int EB, class R>
__device__ __forceinline__ void decode_record(const nq_u4 u, cx<R>* s)
{
  if constexpr (EB == 16) {
    const nq_dv2 d = __builtin_bit_cast(nq_dv2, u);
    s[0] = mk<R>((R)d.x, (R)d.y);
  } else if constexpr (EB == 8) {
    const nq_fv4 f = __builtin_bit_cast(nq_fv4, u);
    s[0] = mk<R>((R)f.x, (R)f.y);
    s[1] = mk<R>((R)f.z, (R)f.w);
  } else {
    const nq_hv8 h = __builtin_bit_cast(nq_hv8, u);
#pragma unroll
    for (int j = 0; j < 4; j++) s[j] = mk<R>((R)(float)h[2 * j], (R)(float)h[2 * j + 1]);
  }
}

// load NC complex elements of site e from a packed array at `base` (num_records
// `bytes`); element type selected by EB: 16 double, 8 float, 4 half
template <int EB, int NC, int CP, class R>
__device__ __forceinline__ void load_packed(const void* base, int64_t bytes, int64_t stride, uint32_t e, cx<R>* s)
{
  using P = Pack<EB, NC>;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, bytes);
#pragma unroll
  for (int q = 0; q < P::NG; q++)
    decode_record<EB, R>(__builtin_amdgcn_raw_buffer_load_b128(rs, e * 16u, (uint32_t)(q * stride * 16), CP),
                         s + q * P::G);
  if constexpr (P::NL > 0) {
    const uint32_t so = (uint32_t)(P::NG * stride * 16), vo = e * (uint32_t)(P::NL * EB);
    constexpr int k0 = P::NG * P::G;
    if constexpr (EB == 8) {  // one float complex
      const nq_fv2 f = __builtin_bit_cast(nq_fv2, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, CP));
      s[k0] = mk<R>((R)f.x, (R)f.y);
    } else if constexpr (P::NL == 1) {
      const nq_hv2 h = __builtin_bit_cast(nq_hv2, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, CP));
      s[k0] = mk<R>((R)(float)h.x, (R)(float)h.y);
    } else if constexpr (P::NL == 2) {
      const nq_hv4 h = __builtin_bit_cast(nq_hv4, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, CP));
      s[k0] = mk<R>((R)(float)h[0], (R)(float)h[1]);
      s[k0 + 1] = mk<R>((R)(float)h[2], (R)(float)h[3]);
    } else {
      const nq_u3 u = __builtin_amdgcn_raw_buffer_load_b96(rs, vo, so, CP);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const nq_hv2 h = __builtin_bit_cast(nq_hv2, (unsigned)u[j]);
        s[k0 + j] = mk<R>((R)(float)h.x, (R)(float)h.y);
      }
    }
  }
}

template <class T, int CP>
__device__ __forceinline__ void put(T* p, const T& v)
{
  if constexpr (CP == 2) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// store NC complex elements of site e (FP16 values already scaled)
template <int EB, int NC, int CP, class R>
__device__ __forceinline__ void store_packed(void* base, int64_t stride, uint32_t e, const cx<R>* s, R scale = (R)1)
{
  using P = Pack<EB, NC>;
  char* b = (char*)base;
#pragma unroll
  for (int q = 0; q < P::NG; q++) {
    char* p = b + (int64_t)q * stride * 16 + (int64_t)e * 16;
    if constexpr (EB == 16) {
      put<nq_dv2, CP>((nq_dv2*)p, nq_dv2{(double)s[q].re, (double)s[q].im});
    } else if constexpr (EB == 8) {
      put<nq_fv4, CP>((nq_fv4*)p, nq_fv4{(float)s[2 * q].re, (float)s[2 * q].im, (float)s[2 * q + 1].re,
                                         (float)s[2 * q + 1].im});
    } else {
      nq_hv8 h;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        h[2 * j] = (_Float16)((float)s[4 * q + j].re * (float)scale);
        h[2 * j + 1] = (_Float16)((float)s[4 * q + j].im * (float)scale);
      }
      put<nq_u4, CP>((nq_u4*)p, __builtin_bit_cast(nq_u4, h));
    }
  }
  if constexpr (P::NL > 0) {
    constexpr int k0 = P::NG * P::G;
    char* p = b + (int64_t)P::NG * stride * 16 + (int64_t)e * (P::NL * EB);
    if constexpr (EB == 8) {
      put<nq_fv2, CP>((nq_fv2*)p, nq_fv2{(float)s[k0].re, (float)s[k0].im});
    } else {
#pragma unroll
      for (int j = 0; j < P::NL; j++) {
        const nq_hv2 h = {(_Float16)((float)s[k0 + j].re * (float)scale), (_Float16)((float)s[k0 + j].im * (float)scale)};
        put<unsigned, CP>((unsigned*)p + j, __builtin_bit_cast(unsigned, h));
      }
    }
  }
}

// ---------------------------------------------------------------- spinor access
// Load NC complex components of site e of right-hand side r; compute type R.
// CP: cache policy of the accesses (0 default; kNT for data used once by the
// kernel, which should not displace the neighbour spinors from L2).
// FP16 fields carry one FP32 norm per site and right-hand side (the largest
// |component|), stored at norm[r*stride + e]; the halves hold component/norm.
constexpr int kNT = 2;
template <int SP, class R, int NC, int CP = 0>
struct SpAcc {
  static constexpr int EB = SP == SP_F64 ? 16 : (SP == SP_F32 ? 8 : 4);  // bytes per complex
  __device__ __forceinline__ static void load(const SpinorView& v, int64_t stride, uint32_t e, cx<R> s[NC], int r = 0)
  {
    const char* base = (const char*)v.data + (int64_t)r * v.rstride * EB;
    load_packed<EB, NC, CP, R>(base, (int64_t)NC * stride * EB, stride, e, s);
    if constexpr (SP == SP_F16) {
      const __amdgpu_buffer_rsrc_t rn = make_rsrc(v.norm + (int64_t)r * stride, stride * 4);
      const R nrm = (R)bload_f1(rn, e * 4, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) s[k] = mk<R>(s[k].re * nrm, s[k].im * nrm);
    }
  }
  __device__ __forceinline__ static void store(const SpinorView& v, int64_t stride, uint32_t e, const cx<R> s[NC],
                                               int r = 0)
  {
    char* base = (char*)v.data + (int64_t)r * v.rstride * EB;
    if constexpr (SP == SP_F16) {
      float m = 0.f;
#pragma unroll
      for (int k = 0; k < NC; k++) m = fmaxf(m, fmaxf(fabsf((float)s[k].re), fabsf((float)s[k].im)));
      // sites whose largest component is below 1e-30 are stored as zero: 1/m would
      // overflow for denormal m (fields far from a point source decay that far)
      if (!(m >= 1e-30f)) m = 0.f;
      const float inv = m > 0.f ? 1.0f / m : 0.f;
      store_packed<EB, NC, CP, R>(base, stride, e, s, (R)inv);
      v.norm[(int64_t)r * stride + e] = m;
    } else {
      store_packed<EB, NC, CP, R>(base, stride, e, s);
    }
  }
};

// ---------------------------------------------------------------- LDS staging
// A workgroup of kStageSites = 256 threads covers 256 consecutive sites of one
// parity; it stages the records of the opposite-parity field at the same 256
// positions in LDS ([record][site], 16-byte entries, plus the FP16 norms), and its
// hops read every neighbour that falls inside that range from LDS instead of
// memory: the +-z neighbours (also across the z boundary, the rows are whole), the
// +-t neighbours inside a time tile and the +-y neighbours between the block's rows.
// Out-of-range neighbours are loaded from memory as usual.  The decoded values are
// the same either way (same records, same decoding).
constexpr int kStageSites = 256;

template <int SP, int NC>
struct Stage {
  static constexpr int EB = SP == SP_F64 ? 16 : (SP == SP_F32 ? 8 : 4);
  static constexpr int NG = Pack<EB, NC>::NG;
  static_assert(Pack<EB, NC>::NL == 0, "LDS staging needs whole 16-byte records");
  static constexpr int BYTES = NG * kStageSites * 16 + (SP == SP_F16 ? kStageSites * 4 : 0);
  // thread t stages site e (valid: e < Vh) of right-hand side r
  __device__ __forceinline__ static void put(char* lds, const SpinorView& v, int64_t stride, uint32_t e, bool valid,
                                             int t, int r)
  {
    const char* base = (const char*)v.data + (int64_t)r * v.rstride * EB;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, (int64_t)NC * stride * EB);
#pragma unroll
    for (int q = 0; q < NG; q++) {
      nq_u4 u = {0u, 0u, 0u, 0u};
      if (valid) u = __builtin_amdgcn_raw_buffer_load_b128(rs, e * 16u, (uint32_t)(q * stride * 16), 0);
      *(nq_u4*)(lds + (q * kStageSites + t) * 16) = u;
    }
    if constexpr (SP == SP_F16) {
      float n = 0.f;
      if (valid) n = v.norm[(int64_t)r * stride + e];
      ((float*)(lds + NG * kStageSites * 16))[t] = n;
    }
  }
  // site at offset d (< kStageSites) of the staged range
  template <class R>
  __device__ __forceinline__ static void get(const char* lds, int d, cx<R> s[NC])
  {
#pragma unroll
    for (int q = 0; q < NG; q++) decode_record<EB, R>(*(const nq_u4*)(lds + (q * kStageSites + d) * 16), s + q * (16 / EB));
    if constexpr (SP == SP_F16) {
      const R nrm = (R)((const float*)(lds + NG * kStageSites * 16))[d];
#pragma unroll
      for (int k = 0; k < NC; k++) s[k] = mk<R>(s[k].re * nrm, s[k].im * nrm);
    }
  }
};

// ---------------------------------------------------------------- link access
// Link loads of the register-blocked (multi-RHS) hops: the RB groups of one site
// range run in adjacent workgroups of one XCD and read the same links, which a
// default-policy load keeps in that XCD's L2 for the next group (measured against
// the non-temporal policy at 32^3x64: 12-source FP64 hop 3023 -> 2853 us, batched
// multi-shift solve of 12 sources FP32 2.37 -> 2.24 s, FP64 2.72 -> 2.45 s).
constexpr int kRbLinkCP = 0;
template <int LK, class R, int CP = kLinkCP>
struct LkAcc {
  static constexpr int NC = lk_nc(LK);
  static constexpr int EB = lk_eb(LK);
  // U as u[3][3]; slot s is a packed array of NC elements at base + s*NC*stride*EB bytes
  __device__ __forceinline__ static void load(const void* base, int64_t stride, int slot, uint32_t e,
                                              const double* det, cx<R> u[9])
  {
    const char* sb = (const char*)base + (int64_t)slot * NC * stride * EB;
    load_packed<EB, NC, CP, R>(sb, (int64_t)NC * stride * EB, stride, e, u);
    if constexpr (NC == 6) {
      // row 2 = det * conj(row0 x row1)
      const cx<R> dt = mk<R>((R)det[0], (R)det[1]);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        const cx<R> c = u[a] * u[3 + b] - u[b] * u[3 + a];
        u[6 + j] = dt * conjg(c);
      }
    }
  }
};

// v_out = U v
template <class R>
__device__ __forceinline__ void mat_vec(const cx<R> u[9], const cx<R> v[3], cx<R> o[3])
{
#pragma unroll
  for (int i = 0; i < 3; i++) {
    cx<R> a = u[i * 3] * v[0];
    fma_to(a, u[i * 3 + 1], v[1]);
    fma_to(a, u[i * 3 + 2], v[2]);
    o[i] = a;
  }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline void finalize(CgState* st, int fin, double total)
{
  switch (fin) {
  case FIN_STORE0: st->extra[0] = total; break;
  case FIN_STORE1: st->extra[1] = total; break;
  case FIN_PAP:
    st->pAp = total;
    st->alpha_prev = st->alpha;
    st->alpha = st->rr / total;
    if (!(total > 0.0) || isnan(total)) { st->nan = 1; st->done = 1; }
    break;
  case FIN_RR: {
    const double rr_new = total;
    st->iter += 1;
    st->beta = rr_new / st->rr;
    st->rr = rr_new;
    if (isnan(st->beta)) { st->nan = 1; st->done = 1; }
    if (!(rr_new >= st->target) || st->iter >= st->niter) st->done = 1;
    break;
  }
  case FIN_RR_SLOPPY: {
    // mixed CG: when |r| has dropped by delta since the last FP64 residual (or the
    // sloppy residual reached the target, or the iteration budget is spent) raise
    // `reliable`: the sloppy kernels then idle until the host has run the FP64
    // reliable update, whose finalisation (FIN_RR_TRUE) clears it again.
    const double rr_new = total;
    st->iter += 1;
    st->rr_new = rr_new;
    if (isnan(rr_new)) { st->nan = 1; st->done = 1; break; }
    if (sqrt(rr_new) < st->delta * st->r0 || rr_new < st->target || st->iter >= st->niter) {
      st->reliable = 1;
    } else {
      st->beta = rr_new / st->rr;
      st->rr = rr_new;
      // paired x updates: this p update applies two terms when the last one left its own
      st->xstep = st->xlag;
      st->xlag = !st->xlag;
    }
    break;
  }
  case FIN_RR_TRUE:
  case FIN_RR_TRUE_HOLD: {
    const double rr_true = total;
    st->nreliable += 1;
    st->true_rr = rr_true;
    st->beta = rr_true / st->rr;
    st->rr = rr_true;
    st->r0 = sqrt(rr_true);
    st->xlag = 0;  // the reliable update has taken every x term
    if (fin == FIN_RR_TRUE) st->reliable = 0;
    if (isnan(rr_true)) { st->nan = 1; st->done = 1; }
    if (rr_true < st->target || st->iter >= st->niter) st->done = 1;
    break;
  }
  case FIN_INIT_B2:
    st->b2 = total;
    st->target *= total;
    st->rr = total;
    st->r0 = sqrt(total);
    st->true_rr = total;
    if (!(total > 0.0) || isnan(total)) { st->nan = 1; st->done = 1; }
    break;
  case FIN_RR_INIT:
    st->rr = total;
    st->r0 = sqrt(total);
    st->true_rr = total;
    if (isnan(total)) { st->nan = 1; st->done = 1; }
    if (!(total >= st->target)) st->done = 1;
    break;
  default: break;
  }
}

// Thread mapping shared by all field kernels: a block covers W chunks of 64
// consecutive sites for NRB right-hand sides of group g (Shape); wavefront w
// handles right-hand side g*NRB + w % NRB of chunk w / NRB.  Launch: 64*W*NRB
// threads, ceil(nchunks/W)*NRG blocks.
struct Map {
  int r;
  int64_t chunk;
  uint32_t e;
  bool active;
};
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2);
// the logical block order gives every XCD one contiguous range of chunks, so
// neighbouring sites are processed by the same XCD.
__device__ __forceinline__ int64_t xcd_block()
{
  const uint32_t b = blockIdx.x, nb = gridDim.x;
  const uint32_t xcd = b & 7u, i = b >> 3, q8 = nb >> 3, r8 = nb & 7u;
  return (int64_t)xcd * q8 + (xcd < r8 ? xcd : r8) + i;
}

// lb: the logical block (xcd_block() for a one-shot launch; a persistent kernel
// passes the block it is working on)
__device__ __forceinline__ Map map_thread_lb(const Shape& S, int64_t Vh, int64_t lb)
{
  Map m;
  // the wave index is uniform across the wavefront: keep r and chunk in SGPRs so
  // that buffer descriptors built from them stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int rg = (int)(lb % S.NRG);
  m.r = rg * S.NRB + wave % S.NRB;
  m.chunk = (lb / S.NRG) * S.W + wave / S.NRB;
  if (m.r >= S.NR) {
    // padding wavefront of the last right-hand-side group: no site, no partial
    m.r = S.NR - 1;
    m.chunk = INT64_MAX / 128;
    m.e = 0;
    m.active = false;
    return m;
  }
  const int64_t e = m.chunk * 64 + lane;
  m.active = e < Vh;
  m.e = (uint32_t)e;
  return m;
}

__device__ __forceinline__ Map map_thread(const Shape& S, int64_t Vh) { return map_thread_lb(S, Vh, xcd_block()); }

// Thread mapping of the register-blocked (multi-RHS) kernels (Shape W = 4, NRB = 1):
// with g.rb_bpc > 0 the 256-site blocks are visited z slab by z slab, logical block
// l -> slab l / rb_ncol and column l % rb_ncol -> stored block column * rb_bpc + slab.
// A batched hop moves several times the bytes per site of a one-RHS hop, so in the
// storage order the neighbour rows a y tile away (thousands of sites) have left the
// XCD's L2 when they are read again; inside a slab the +-x and +-y neighbour blocks
// of a block are a few blocks apart in the visit order.  Each (right-hand side,
// chunk) pair is computed as in the storage order, so results do not change.
__device__ __forceinline__ Map map_thread_rb(const Shape& S, const Geo& g, int nr)
{
  Map m;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t lb = xcd_block();
  const int rg = (int)(lb % S.NRG);
  int64_t blk = lb / S.NRG;
  if (g.rb_bpc > 0 && S.W == 4 && nr > 1) {  // one right-hand side: storage order
    const int64_t zb = blk / g.rb_ncol, col = blk - zb * g.rb_ncol;
    blk = col * g.rb_bpc + zb;
  }
  m.r = rg * S.NRB + wave % S.NRB;
  m.chunk = blk * S.W + wave / S.NRB;
  if (m.r >= S.NR) {
    m.r = S.NR - 1;
    m.chunk = INT64_MAX / 128;
    m.e = 0;
    m.active = false;
    return m;
  }
  const int64_t e = m.chunk * 64 + lane;
  m.active = e < g.Vh;
  m.e = (uint32_t)e;
  return m;
}

// Deterministic reductions over 64-site chunks.  Every wavefront holds the
// partial of one (right-hand side r, chunk c) pair and stores it at
// partial[r*nchunks + c] with a plain store; a one-block finalisation kernel
// launched right after (k_finalize) sums the chunk partials of each r in a fixed
// pattern and updates state[r].  No atomics or device-scope fences are needed:
// the kernel boundary publishes the partials.  The summation order depends only
// on the lattice, never on how many right-hand sides share the launch or on the
// block shape, so a result is bit-identical alone or in a batch.
__device__ __forceinline__ void store_partial(double v, double* partial, int64_t nchunks, int r, int64_t chunk)
{
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && chunk < nchunks) partial[(int64_t)r * nchunks + chunk] = v;
}

// One workgroup per right-hand side: 1024 lanes each sum a fixed strided subset of
// the chunk partials (4 interleaved chains), then a fixed LDS tree down to one
// wavefront and a fixed shuffle tree.  The order depends only on nchunks.
__device__ __forceinline__ double block_sum_partials(const double* pr, int64_t nchunks, double* vl)
{
  const int vt = threadIdx.x;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int64_t c = vt;
  for (; c + 3072 < nchunks; c += 4096) {
    a0 += pr[c];
    a1 += pr[c + 1024];
    a2 += pr[c + 2048];
    a3 += pr[c + 3072];
  }
  for (; c < nchunks; c += 1024) a0 += pr[c];
  vl[vt] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int h = 512; h >= 64; h >>= 1) {
    if (vt < h) vl[vt] += vl[vt + h];
    __syncthreads();
  }
  return vt < 64 ? wave_sum(vl[vt]) : 0.0;
}

// block_sum_partials with 256 threads: each thread takes the chains of the virtual
// lanes t, t+256, t+512, t+768 and the LDS tree the same pairs, so the sum has the
// bits of the 1024-thread version (vl: 1024 doubles)
__device__ __forceinline__ double block_sum_partials_256(const double* pr, int64_t nchunks, double* vl)
{
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int vt = t + 256 * q;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int64_t c = vt;
    for (; c + 3072 < nchunks; c += 4096) {
      a0 += pr[c];
      a1 += pr[c + 1024];
      a2 += pr[c + 2048];
      a3 += pr[c + 3072];
    }
    for (; c < nchunks; c += 1024) a0 += pr[c];
    vl[vt] = (a0 + a1) + (a2 + a3);
  }
  __syncthreads();
  for (int h = 512; h >= 64; h >>= 1) {
    for (int vt = t; vt < h; vt += 256) vl[vt] += vl[vt + h];
    __syncthreads();
  }
  return t < 64 ? wave_sum(vl[t]) : 0.0;
}

}  // namespace nq
