// Layout conversions between nissa's host AoS buffers and the device SoA
// layouts, and construction of the per-output-parity link slot arrays.
//
// nissa AoS (src/base/field.hpp:408-416): site-major, internal degrees
// [id][ic][ri] for spincolor, [mu][ic1][ic2][ri] for quad_su3, sites in nissa's
// lx / eo order.  On the device a field of one parity is SoA over complex
// components (stride = padded Vh) in the device site order (t tiled, see Geo), so
// a wavefront reading one component of 64 consecutive sites issues one 1 KiB
// (FP64) coalesced transaction.  Links are re-laid per output parity p as 8
// slots (U_mu(x), U_mu(x-mu)^+) so the hop streams them without gathers.

#include "nq_device.hpp"

namespace nq {

// host index of device site e of parity par: nissa lx (host_lx) or nissa eo = lx>>1
__device__ __forceinline__ int64_t host_index(const Geo& g, int par, uint32_t e, int host_lx)
{
  const Site s = site_of(g, par, e);
  const int64_t lx = nissa_lx(g, s.t, s.x, s.y, s.z);
  return host_lx ? lx : lx >> 1;
}

template <int NC>
__global__ void __launch_bounds__(kBlock) k_aos_to_soa(Geo g, const double* aos, int par, int host_lx, SpinorView out)
{
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const int64_t src = host_index(g, par, e, host_lx);
  const double* p = aos + (int64_t)src * 2 * NC;
  cx<double> v[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) v[k] = mk<double>(p[2 * k], p[2 * k + 1]);
  SpAcc<SP_F64, double, NC>::store(out, g.stride, e, v);
}

template <int NC>
__global__ void __launch_bounds__(kBlock) k_soa_to_aos(Geo g, SpinorView in, int par, int host_lx, double* aos)
{
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= (uint32_t)g.Vh) return;
  const int64_t dst = host_index(g, par, e, host_lx);
  cx<double> v[NC];
  SpAcc<SP_F64, double, NC>::load(in, g.stride, e, v);
  double* p = aos + (int64_t)dst * 2 * NC;
#pragma unroll
  for (int k = 0; k < NC; k++) { p[2 * k] = v[k].re; p[2 * k + 1] = v[k].im; }
}

int launch_aos_to_soa(nq_ctx* ctx, const double* aos, int ncomp, int par, int host_lx, const SpinorView& out)
{
  if (out.prec != SP_F64) return fail(NQ_EINVAL, "upload target must be FP64");
  const int64_t nb = blocks_for(ctx->geo.Vh);
  if (ncomp == 12)
    hipLaunchKernelGGL(k_aos_to_soa<12>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, ctx->geo, aos, par, host_lx, out);
  else
    hipLaunchKernelGGL(k_aos_to_soa<3>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, ctx->geo, aos, par, host_lx, out);
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

int launch_soa_to_aos(nq_ctx* ctx, const SpinorView& in, int ncomp, int par, int host_lx, double* aos)
{
  if (in.prec != SP_F64) return fail(NQ_EINVAL, "download source must be FP64");
  const int64_t nb = blocks_for(ctx->geo.Vh);
  if (ncomp == 36)
    hipLaunchKernelGGL(k_soa_to_aos<36>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, ctx->geo, in, par, host_lx, aos);
  else if (ncomp == 12)
    hipLaunchKernelGGL(k_soa_to_aos<12>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, ctx->geo, in, par, host_lx, aos);
  else
    hipLaunchKernelGGL(k_soa_to_aos<3>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, ctx->geo, in, par, host_lx, aos);
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

// store element k of site e of one link slot array (packed records, see Pack)
template <int LK>
__device__ __forceinline__ void put_link(void* slot_base, int k, int64_t e, int64_t stride, double re, double im)
{
  constexpr int NC = lk_nc(LK);
  constexpr int EB = lk_eb(LK);
  char* p = (char*)slot_base + Pack<EB, NC>::offset(k, e, stride);
  if constexpr (EB == 16) *(double2*)p = make_double2(re, im);
  else if constexpr (EB == 8) *(float2*)p = make_float2((float)re, (float)im);
  else *(__half2*)p = __floats2half2_rn((float)re, (float)im);
}

// out layout: [op][slot][nc][stride]
template <int LK>
__global__ void __launch_bounds__(kBlock) k_build_links(Geo g, const double* ae, const double* ao, void* out)
{
  constexpr int NC = lk_nc(LK);
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  const int op = blockIdx.y;
  if (e >= (uint32_t)g.Vh) return;
  const double* conf[2] = {ae, ao};
  const Site s = site_of(g, op, e);
  const int64_t ne = nissa_lx(g, s.t, s.x, s.y, s.z) >> 1;
  constexpr int EB = lk_eb(LK);
  const int64_t slot_bytes = (int64_t)NC * g.stride * EB;
  char* par_base = (char*)out + (int64_t)op * 8 * slot_bytes;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    // forward: U_mu(x) as stored
    const double* Uf = conf[op] + (ne * 4 + mu) * 18;
    char* bf = par_base + (int64_t)(2 * mu) * slot_bytes;
#pragma unroll
    for (int k = 0; k < NC; k++) put_link<LK>(bf, k, e, g.stride, Uf[2 * k], Uf[2 * k + 1]);
    // backward: U_mu(x-mu)^+ , element (i,j) = conj(U(j,i))
    const int64_t n = nissa_neigh_eo(g, s, mu, false);
    const double* Ub = conf[!op] + (n * 4 + mu) * 18;
    char* bb = par_base + (int64_t)(2 * mu + 1) * slot_bytes;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const int i = k / 3, j = k % 3;
      put_link<LK>(bb, k, e, g.stride, Ub[2 * (j * 3 + i)], -Ub[2 * (j * 3 + i) + 1]);
    }
  }
}

int launch_build_links(nq_ctx* ctx, const double* ae, const double* ao, int lk, void* out, const double (*)[2])
{
  const int64_t nb = blocks_for(ctx->geo.Vh);
  const dim3 grid((unsigned)nb, 2);
  switch (lk) {
  case LK_F64_18: hipLaunchKernelGGL(k_build_links<LK_F64_18>, grid, dim3(kBlock), 0, ctx->stream, ctx->geo, ae, ao, out); break;
  case LK_F32_18: hipLaunchKernelGGL(k_build_links<LK_F32_18>, grid, dim3(kBlock), 0, ctx->stream, ctx->geo, ae, ao, out); break;
  case LK_F32_12: hipLaunchKernelGGL(k_build_links<LK_F32_12>, grid, dim3(kBlock), 0, ctx->stream, ctx->geo, ae, ao, out); break;
  case LK_F64_12: hipLaunchKernelGGL(k_build_links<LK_F64_12>, grid, dim3(kBlock), 0, ctx->stream, ctx->geo, ae, ao, out); break;
  default: hipLaunchKernelGGL(k_build_links<LK_F16_12>, grid, dim3(kBlock), 0, ctx->stream, ctx->geo, ae, ao, out); break;
  }
  NQ_HIP(hipGetLastError());
  return NQ_OK;
}

}  // namespace nq
