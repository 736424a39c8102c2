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
  int t, x, y, z;
  uint32_t lx;
};

__device__ __forceinline__ Site site_of(const Geo& g, int par, uint32_t e)
{
  Site s;
  const uint32_t q = 2u * e;
  uint32_t r = q / (uint32_t)g.LZ;
  const int z0 = (int)(q - r * (uint32_t)g.LZ);
  const uint32_t r2 = r / (uint32_t)g.LY;
  s.y = (int)(r - r2 * (uint32_t)g.LY);
  const uint32_t r3 = r2 / (uint32_t)g.LX;
  s.x = (int)(r2 - r3 * (uint32_t)g.LX);
  s.t = (int)r3;
  const int b = (par + s.t + s.x + s.y) & 1;
  s.z = z0 + b;
  s.lx = q + (uint32_t)b;
  return s;
}

// eo index of the neighbour in direction mu (0=t..3=z), up (fwd) or down
__device__ __forceinline__ uint32_t neigh_eo(const Geo& g, const Site& s, int mu, bool up)
{
  int64_t lx = s.lx;
  switch (mu) {
  case 0:
    lx += up ? (s.t == g.T - 1 ? -(int64_t)(g.T - 1) * g.VS : g.VS) : (s.t == 0 ? (int64_t)(g.T - 1) * g.VS : -g.VS);
    break;
  case 1:
    lx += up ? (s.x == g.LX - 1 ? -(int64_t)(g.LX - 1) * g.S2 : g.S2) : (s.x == 0 ? (int64_t)(g.LX - 1) * g.S2 : -g.S2);
    break;
  case 2:
    lx += up ? (s.y == g.LY - 1 ? -(int64_t)(g.LY - 1) * g.LZ : g.LZ) : (s.y == 0 ? (int64_t)(g.LY - 1) * g.LZ : -(int64_t)g.LZ);
    break;
  default:
    lx += up ? (s.z == g.LZ - 1 ? -(g.LZ - 1) : 1) : (s.z == 0 ? (g.LZ - 1) : -1);
    break;
  }
  return (uint32_t)(lx >> 1);
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
__device__ __forceinline__ double2 bload_d2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const nq_dv2 v = __builtin_bit_cast(nq_dv2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 bload_f2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const nq_fv2 v = __builtin_bit_cast(nq_fv2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 bload_h2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  return __half22float2(__builtin_bit_cast(__half2, v));
}
__device__ __forceinline__ float bload_f1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// ---------------------------------------------------------------- spinor access
// Load NC complex components of site e; compute type R.
template <int SP, class R, int NC>
struct SpAcc {
  static constexpr int EB = SP == SP_F64 ? 16 : (SP == SP_F32 ? 8 : 4);  // bytes per complex
  // component k of site e of right-hand side r lives at data[r*rstride + k*stride + e]
  __device__ __forceinline__ static void load(const SpinorView& v, int64_t stride, uint32_t e, cx<R> s[NC], int r = 0)
  {
    const char* base = (const char*)v.data + (int64_t)r * v.rstride * EB;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, (int64_t)NC * stride * EB);
    const uint32_t vo = e * EB;
    if constexpr (SP == SP_F64) {
#pragma unroll
      for (int k = 0; k < NC; k++) {
        const double2 d = bload_d2(rs, vo, (uint32_t)(k * stride * EB));
        s[k] = mk<R>((R)d.x, (R)d.y);
      }
    } else if constexpr (SP == SP_F32) {
#pragma unroll
      for (int k = 0; k < NC; k++) {
        const float2 d = bload_f2(rs, vo, (uint32_t)(k * stride * EB));
        s[k] = mk<R>((R)d.x, (R)d.y);
      }
    } else {
      const __amdgpu_buffer_rsrc_t rn = make_rsrc(v.norm + (int64_t)r * stride, stride * 4);
      const R nrm = (R)bload_f1(rn, e * 4, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) {
        const float2 d = bload_h2(rs, vo, (uint32_t)(k * stride * EB));
        s[k] = mk<R>((R)d.x * nrm, (R)d.y * nrm);
      }
    }
  }
  __device__ __forceinline__ static void store(const SpinorView& v, int64_t stride, uint32_t e, const cx<R> s[NC],
                                               int r = 0)
  {
    const int64_t o = (int64_t)r * v.rstride + e;
    if constexpr (SP == SP_F64) {
      double2* p = (double2*)v.data + o;
#pragma unroll
      for (int k = 0; k < NC; k++) p[k * stride] = make_double2((double)s[k].re, (double)s[k].im);
    } else if constexpr (SP == SP_F32) {
      float2* p = (float2*)v.data + o;
#pragma unroll
      for (int k = 0; k < NC; k++) p[k * stride] = make_float2((float)s[k].re, (float)s[k].im);
    } else {
      float m = 0.f;
#pragma unroll
      for (int k = 0; k < NC; k++) m = fmaxf(m, fmaxf(fabsf((float)s[k].re), fabsf((float)s[k].im)));
      const float inv = m > 0.f ? 1.0f / m : 0.f;
      __half2* p = (__half2*)v.data + o;
#pragma unroll
      for (int k = 0; k < NC; k++)
        p[k * stride] = __floats2half2_rn((float)s[k].re * inv, (float)s[k].im * inv);
      v.norm[(int64_t)r * stride + e] = m;
    }
  }
};

// ---------------------------------------------------------------- link access
template <int LK, class R>
struct LkAcc {
  static constexpr int NC = (LK == LK_F64_18 || LK == LK_F32_18) ? 9 : 6;
  static constexpr int EB = LK == LK_F64_18 ? 16 : (LK == LK_F16_12 ? 4 : 8);
  // U as u[3][3]; the slot array starts at base + slot*NC*stride elements
  __device__ __forceinline__ static void load(const void* base, int64_t stride, int slot, uint32_t e,
                                              const double* det, cx<R> u[9])
  {
    const char* sb = (const char*)base + (int64_t)slot * NC * stride * EB;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(sb, (int64_t)NC * stride * EB);
    const uint32_t vo = e * EB;
    if constexpr (LK == LK_F64_18) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const double2 d = bload_d2(rs, vo, (uint32_t)(k * stride * EB));
        u[k] = mk<R>((R)d.x, (R)d.y);
      }
    } else if constexpr (LK == LK_F32_18) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const float2 d = bload_f2(rs, vo, (uint32_t)(k * stride * EB));
        u[k] = mk<R>((R)d.x, (R)d.y);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const float2 d = LK == LK_F32_12 ? bload_f2(rs, vo, (uint32_t)(k * stride * EB))
                                         : bload_h2(rs, vo, (uint32_t)(k * stride * EB));
        u[k] = mk<R>((R)d.x, (R)d.y);
      }
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

// Block sum in fixed order (wavefront butterflies then waves in order).
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds)
{
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; i++) t += lds[i];
  }
  return t;  // valid in thread 0
}

__device__ inline void finalize(CgState* st, int fin, double total)
{
  switch (fin) {
  case FIN_STORE0: st->extra[0] = total; break;
  case FIN_STORE1: st->extra[1] = total; break;
  case FIN_PAP:
    st->pAp = total;
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
    const double rr_new = total;
    st->iter += 1;
    st->rr_new = rr_new;
    if (isnan(rr_new)) { st->nan = 1; st->done = 1; break; }
    if (sqrt(rr_new) < st->delta * st->r0 || rr_new < st->target || st->iter >= st->niter) {
      st->reliable = 1;
    } else {
      st->reliable = 0;
      st->beta = rr_new / st->rr;
      st->rr = rr_new;
    }
    break;
  }
  case FIN_RR_TRUE: {
    const double rr_true = total;
    st->nreliable += 1;
    st->true_rr = rr_true;
    st->beta = rr_true / st->rr;
    st->rr = rr_true;
    st->r0 = sqrt(rr_true);
    st->reliable = 0;
    if (isnan(rr_true)) { st->nan = 1; st->done = 1; }
    if (rr_true < st->target || st->iter >= st->niter) st->done = 1;
    break;
  }
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
// consecutive sites for each of NR right-hand sides; wavefront w handles
// right-hand side w % NR of chunk w / NR.  Launch: 64*W*NR threads,
// ceil(nchunks/W) blocks.
struct Map {
  int r;
  int64_t chunk;
  uint32_t e;
  bool active;
};
__device__ __forceinline__ Map map_thread(int W, int NR, int64_t Vh)
{
  Map m;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  m.r = wave % NR;
  m.chunk = (int64_t)blockIdx.x * W + wave / NR;
  const int64_t e = m.chunk * 64 + lane;
  m.active = e < Vh;
  m.e = (uint32_t)e;
  return m;
}

// Deterministic grid reduction over 64-site chunks.  Every wavefront holds the
// partial of one (right-hand side r, chunk c) pair, stores it at
// partial[r*nchunks + c], and the last block to arrive sums, for each r, the
// chunk partials with one wavefront in a fixed pattern (lane i takes chunks
// i, i+64, ...; then butterflies) and finalises state[r].  The summation order
// therefore depends only on the lattice, never on how many right-hand sides
// share the launch or on the block shape.
__device__ __forceinline__ void grid_reduce_rhs(double v, double* partial, int64_t nchunks, int r, int64_t chunk,
                                                int nr, CgState* st, unsigned int* counter, int fin)
{
  __shared__ int am_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0 && r >= 0 && chunk < nchunks) {
    __hip_atomic_store(&partial[(int64_t)r * nchunks + chunk], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (am_last) {
    __threadfence();
    for (int rr = wave; rr < nr; rr += nwaves) {
      CgState* s = st + rr;
      if (s->done) continue;
      double acc = 0;
      for (int64_t c = lane; c < nchunks; c += 64)
        acc += __hip_atomic_load(&partial[(int64_t)rr * nchunks + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      acc = wave_sum(acc);
      if (lane == 0) finalize(s, fin, acc);
    }
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace nq
