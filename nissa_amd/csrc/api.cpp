// C ABI: contexts, device memory, fields, gauge handles, profiling ledger.
// Replaces nissa's field containers and device plumbing on the hot path
// (src/base/field.hpp:659-1805, src/base/memory_manager.hpp:221-434,
// src/threads/threads.hpp:86-265) with explicit device handles.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nq_internal.hpp"

namespace nq {

static thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int dev_alloc(DevBuf& b, size_t bytes)
{
  if (b.p && b.bytes >= bytes) return NQ_OK;
  dev_free(b);
  if (bytes == 0) return NQ_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(NQ_EOOM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return NQ_OK;
}

void dev_free(DevBuf& b)
{
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int ctx_temp(nq_ctx* ctx, const std::string& role, size_t bytes, void** out)
{
  DevBuf& b = ctx->temps[role];
  NQ_TRY(dev_alloc(b, bytes));
  *out = b.p;
  return NQ_OK;
}

// free the temporaries whose role starts with `prefix` (after the stream has drained:
// kernels already enqueued may still read them)
int ctx_temp_release(nq_ctx* ctx, const std::string& prefix)
{
  NQ_HIP(hipStreamSynchronize(ctx->stream));
  for (auto it = ctx->temps.begin(); it != ctx->temps.end();) {
    if (it->first.compare(0, prefix.size(), prefix) == 0) {
      dev_free(it->second);
      it = ctx->temps.erase(it);
    } else {
      ++it;
    }
  }
  return NQ_OK;
}

// Fingerprint of a host conf: 64-bit words of a sample mixed with the splitmix64
// finaliser.  The sample is the first and last 72 doubles (one quad_su3 site) and about
// 65536 doubles spread with a stride coprime to 72, so that every one of the 72 reals
// of a site (4 directions x 3 x 3 complex) is sampled about 900 times: an in-place
// change of every link of one direction (boundary phases, staggered phases, a new
// conf read into the same buffer) changes it.  A change confined to a few links can
// escape the sample.
uint64_t conf_fingerprint(const double* p, int64_t n)
{
  uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
  auto mix = [&h](double d) {
    uint64_t w;
    std::memcpy(&w, &d, 8);
    uint64_t z = h ^ w;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    h = z ^ (z >> 31);
  };
  int64_t stride = std::max<int64_t>(1, n / 65536) | 1;
  while (stride % 3 == 0) stride += 2;
  for (int64_t i = 0; i < std::min<int64_t>(n, 72); i++) mix(p[i]);
  for (int64_t i = 0; i < n; i += stride) mix(p[i]);
  for (int64_t i = std::max<int64_t>(0, n - 72); i < n; i++) mix(p[i]);
  return h;
}

int spinor_bytes(int prec, int ncomp, const Geo& g, size_t* data_bytes, size_t* norm_bytes)
{
  const size_t n = (size_t)ncomp * (size_t)g.stride;
  switch (prec) {
  case SP_F64: *data_bytes = n * 16; *norm_bytes = 0; break;
  case SP_F32: *data_bytes = n * 8; *norm_bytes = 0; break;
  case SP_F16: *data_bytes = n * 4; *norm_bytes = (size_t)g.stride * 4; break;
  default: return fail(NQ_EINVAL, "bad spinor precision %d", prec);
  }
  return NQ_OK;
}

// temporary field of nr right-hand sides (rstride = ncomp*stride elements)
int alloc_spinor_nr(nq_ctx* ctx, const std::string& role, int prec, int ncomp, int nr, SpinorView* v)
{
  size_t db = 0, nb = 0;
  NQ_TRY(spinor_bytes(prec, ncomp, ctx->geo, &db, &nb));
  void* p = nullptr;
  NQ_TRY(ctx_temp(ctx, role, (db + nb) * nr + 256, &p));
  v->data = p;
  v->prec = prec;
  v->ncomp = ncomp;
  v->rstride = (int64_t)ncomp * ctx->geo.stride;
  v->norm = nb ? (float*)((char*)p + db * nr) : nullptr;
  return NQ_OK;
}

int alloc_spinor(nq_ctx* ctx, const std::string& role, int prec, int ncomp, SpinorView* v)
{
  return alloc_spinor_nr(ctx, role, prec, ncomp, 1, v);
}

SpinorView view_of(const nq_field* f, int parity)
{
  SpinorView v;
  v.prec = SP_F64;
  v.ncomp = f->ncomp();
  v.rstride = 0;
  const int slot = f->coverage == NQ_LX ? parity : 0;
  v.data = (char*)f->buf.p + (size_t)slot * f->ncomp() * f->ctx->geo.stride * 16;
  return v;
}

// lx field as one field of 2*ncomp components over Vh sites (site-local ops only)
SpinorView view_lx_flat(const nq_field* f)
{
  SpinorView v = view_of(f, 0);
  v.ncomp = 2 * f->ncomp();
  return v;
}

int ensure_state(nq_ctx* ctx, int n)
{
  if (ctx->nstate >= n) return NQ_OK;
  return fail(NQ_EINVAL, "solver state pool too small (%d < %d)", ctx->nstate, n);
}

// ---------------------------------------------------------------- profiling
void prof_begin(nq_ctx* ctx, const char* name, double bytes, hipEvent_t* a)
{
  *a = nullptr;
  if (!ctx->profiling) return;
  hipEvent_t e0, e1;
  if (ctx->event_pool.size() >= 2) {
    e0 = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    e1 = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  } else {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return;
  }
  (void)hipEventRecord(e0, ctx->stream);
  ctx->pending.push_back({name, e0, e1, bytes});
  *a = e1;
}

void prof_begin_ext(nq_ctx* ctx, const char* name, double bytes, hipEvent_t* a, hipEvent_t* b)
{
  *a = *b = nullptr;
  if (!ctx->profiling) return;
  hipEvent_t e0, e1;
  if (ctx->event_pool.size() >= 2) {
    e0 = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    e1 = ctx->event_pool.back();
    ctx->event_pool.pop_back();
  } else {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return;
  }
  ctx->pending.push_back({name, e0, e1, bytes});
  *a = e0;
  *b = e1;
}

void prof_end_ext(nq_ctx* ctx)
{
  ctx->hop_ev[0] = ctx->hop_ev[1] = nullptr;
  if (ctx->pending.size() > 8192) prof_flush(ctx);
}

void prof_end(nq_ctx* ctx, hipEvent_t b)
{
  if (!b) return;
  (void)hipEventRecord(b, ctx->stream);
  if (ctx->pending.size() > 8192) prof_flush(ctx);
}

void prof_flush(nq_ctx* ctx)
{
  if (ctx->pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KernelStat& s = ctx->stats[p.name];
      s.launches++;
      s.total_ms += ms;
      s.total_bytes += p.bytes;
    }
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
}

// Device site order of a T x LX x LY x LZ lattice: t tiles of TB slices (the largest
// even divisor of T up to tbmax), x/y patches of TX x TY sites (powers of two up to
// 16, halved until they divide LX and LY).
Geo make_geo(int T, int LX, int LY, int LZ, int tbmax, int tx, int ty)
{
  Geo g{};
  g.T = T; g.LX = LX; g.LY = LY; g.LZ = LZ;
  g.Vh = (int64_t)T * LX * LY * LZ / 2;
  g.stride = (g.Vh + 63) / 64 * 64;
  int tb = 2;
  for (int d = 2; d <= std::min(T, tbmax); d += 2)
    if (T % d == 0) tb = d;
  g.TB = tb;
  g.NTH = T / g.TB;
  auto pow2 = [](int v) { return v >= 1 && v <= 16 && (v & (v - 1)) == 0; };
  if (!pow2(tx)) tx = 1;
  if (!pow2(ty)) ty = 1;
  while (LX % tx) tx /= 2;
  while (LY % ty) ty /= 2;
  auto lg = [](int v) { int l = 0; while ((1 << l) < v) l++; return l; };
  g.TX = tx; g.logTX = lg(tx); g.NXT = LX / tx;
  g.TY = ty; g.logTY = lg(ty); g.NYT = LY / ty;
  g.Sz = g.TB;
  g.Syi = (int64_t)g.TB * LZ;
  g.Sxi = g.Syi * g.TY;
  g.Syt = g.Sxi * g.TX;
  g.Sxt = g.Syt * g.NYT;
  g.Sth = g.Sxt * g.NXT;
  // z-slab visit order of the batched kernels when every column is a whole number
  // (at least two) of 256-site blocks; NQ_RB_ORDER=0 keeps the storage order
  const int64_t col = (int64_t)g.TB * LZ / 2;
  const char* ro = getenv("NQ_RB_ORDER");
  const bool slabs = !(ro && ro[0] == '0');
  if (slabs && g.Vh % 256 == 0 && col % 256 == 0 && col >= 512) {
    g.rb_bpc = col / 256;
    g.rb_ncol = g.Vh / col;
  }
  return g;
}

}  // namespace nq

using namespace nq;

int64_t nq_gauge::bytes() const
{
  return (int64_t)(aos[0].bytes + aos[1].bytes + l64.bytes + l32.bytes + l32c.bytes + l16c.bytes + l64c.bytes);
}

namespace {
struct Guard {
  std::lock_guard<std::recursive_mutex> lk;
  explicit Guard(nq_ctx* c) : lk(c->mu) { (void)hipSetDevice(c->device); }
};
bool dims_ok(int T, int LX, int LY, int LZ)
{
  const int L[4] = {T, LX, LY, LZ};
  for (int i = 0; i < 4; i++)
    if (L[i] < 2 || (L[i] & 1)) return false;
  return (int64_t)T * LX * LY * LZ < (int64_t)1 << 31;
}
}  // namespace

extern "C" {

const char* nq_last_error(void) { return g_err; }

int nq_device_count(int* n)
{
  if (!n) return fail(NQ_EINVAL, "null pointer");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *n = 0;
    (void)hipGetLastError();
    return fail(NQ_EDEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *n = c;
  return NQ_OK;
}

int nq_ctx_create(int T, int LX, int LY, int LZ, int device, nq_ctx** out)
{
  if (!out) return fail(NQ_EINVAL, "null output");
  *out = nullptr;
  if (!dims_ok(T, LX, LY, LZ)) return fail(NQ_EINVAL, "lattice extents must be even and >= 2 (got %d %d %d %d)", T, LX, LY, LZ);
  int n = 0;
  NQ_TRY(nq_device_count(&n));
  if (device < 0 || device >= n) return fail(NQ_EDEVICE, "device %d not available (%d devices)", device, n);
  NQ_HIP(hipSetDevice(device));
  auto* c = new nq_ctx;
  c->device = device;
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->ncu = 0;
  // t tile of the device site order: the largest even divisor of T up to 64, or
  // up to NQ_TTILE when set.  Whole t rows in one tile keep every t neighbour in
  // the lines the wavefront already reads: a t tile boundary sends a quarter of
  // the lanes' t-neighbour loads to lines of a far tile (of another XCD), which
  // costs ~10% of the hop at 32^3x64 with tiles of 8.
  int tbmax = 64;
  const char* ett = getenv("NQ_TTILE");
  if (ett) tbmax = atoi(ett);
  if (tbmax < 2) tbmax = 64;
  // x/y patch of the device site order (NQ_XYTILE "TXxTY", powers of two; default
  // 4x2: the x and y neighbours inside a patch are a few thousand sites apart;
  // measured at 32^3x64 with the persistent FP64 hop, 4x2 269 us, 4x1 271 us, 8x1
  // 287 us, 2x1 280 us)
  int tx = 4, ty = 2;
  const char* exy = getenv("NQ_XYTILE");
  if (exy) sscanf(exy, "%dx%d", &tx, &ty);
  c->geo = make_geo(T, LX, LY, LZ, tbmax, tx, ty);
  const Geo& g = c->geo;
  int rc = NQ_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = fail(NQ_EDEVICE, "stream creation failed");
  const int64_t nchunks = (g.Vh + 63) / 64;
  c->max_partials = nchunks * 16;
  c->nstate = 64;
  if (rc == NQ_OK && hipMalloc(&c->partial, sizeof(double) * c->max_partials) != hipSuccess) rc = fail(NQ_EOOM, "partials");
  if (rc == NQ_OK && hipMalloc(&c->state, sizeof(CgState) * c->nstate) != hipSuccess) rc = fail(NQ_EOOM, "state");
  if (rc == NQ_OK && hipHostMalloc(&c->state_host, sizeof(CgState) * c->nstate, hipHostMallocDefault) != hipSuccess)
    rc = fail(NQ_EOOM, "pinned state");
  if (rc == NQ_OK && hipMemset(c->state, 0, sizeof(CgState) * c->nstate) != hipSuccess) rc = fail(NQ_EDEVICE, "memset");
  if (rc == NQ_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(NQ_EDEVICE, "sync");
  // site-order tiles chosen by the autotuner (cached per device and lattice), unless
  // the caller fixed them or the lattice is small enough for the defaults
  if (rc == NQ_OK && !ett && !exy) rc = tune_site_tiles(c);
  if (rc != NQ_OK) {
    nq_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return NQ_OK;
}

int nq_ctx_site_tiles(const nq_ctx* c, int* tb, int* tx, int* ty)
{
  if (!c || !tb || !tx || !ty) return fail(NQ_EINVAL, "null argument");
  *tb = c->geo.TB;
  *tx = c->geo.TX;
  *ty = c->geo.TY;
  return NQ_OK;
}

void nq_ctx_destroy(nq_ctx* c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (auto& t : c->temps) dev_free(t.second);
  if (c->partial) (void)hipFree(c->partial);
  if (c->state) (void)hipFree(c->state);
  if (c->state_host) (void)hipHostFree(c->state_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int nq_ctx_sync(nq_ctx* c)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  Guard G(c);
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_ctx_set_profiling(nq_ctx* c, int on)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  Guard G(c);
  if (!on) prof_flush(c);
  c->profiling = on != 0;
  return NQ_OK;
}

int nq_ctx_reset_kernel_stats(nq_ctx* c)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  Guard G(c);
  prof_flush(c);
  c->stats.clear();
  return NQ_OK;
}

int nq_ctx_kernel_count(nq_ctx* c, int* n)
{
  if (!c || !n) return fail(NQ_EINVAL, "null argument");
  Guard G(c);
  prof_flush(c);
  *n = (int)c->stats.size();
  return NQ_OK;
}

int nq_ctx_kernel_stat(nq_ctx* c, int i, char* name, int name_len, int64_t* launches, double* total_ms,
                       double* bytes_per_launch)
{
  if (!c) return fail(NQ_EINVAL, "null ctx");
  Guard G(c);
  prof_flush(c);
  if (i < 0 || i >= (int)c->stats.size()) return fail(NQ_EINVAL, "kernel index out of range");
  auto it = c->stats.begin();
  std::advance(it, i);
  if (name && name_len > 0) snprintf(name, name_len, "%s", it->first.c_str());
  if (launches) *launches = it->second.launches;
  if (total_ms) *total_ms = it->second.total_ms;
  if (bytes_per_launch) *bytes_per_launch = it->second.launches ? it->second.total_bytes / it->second.launches : 0.0;
  return NQ_OK;
}

// ------------------------------------------------------------------ fields
int nq_field_create(nq_ctx* c, int kind, int coverage, nq_field** out)
{
  if (!c || !out) return fail(NQ_EINVAL, "null argument");
  *out = nullptr;
  if (kind != NQ_SPINCOLOR && kind != NQ_COLOR) return fail(NQ_EINVAL, "bad field kind %d", kind);
  if (coverage < NQ_EVEN || coverage > NQ_LX) return fail(NQ_EINVAL, "bad coverage %d", coverage);
  Guard G(c);
  auto* f = new nq_field;
  f->ctx = c;
  f->kind = kind;
  f->coverage = coverage;
  const size_t per = (size_t)f->ncomp() * c->geo.stride * 16;
  int rc = dev_alloc(f->buf, per * (coverage == NQ_LX ? 2 : 1));
  if (rc == NQ_OK && hipMemsetAsync(f->buf.p, 0, f->buf.bytes, c->stream) != hipSuccess) rc = fail(NQ_EDEVICE, "memset");
  if (rc == NQ_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(NQ_EDEVICE, "sync");
  if (rc != NQ_OK) {
    dev_free(f->buf);
    delete f;
    return rc;
  }
  *out = f;
  return NQ_OK;
}

void nq_field_destroy(nq_field* f)
{
  if (!f) return;
  {
    Guard G(f->ctx);
    (void)hipStreamSynchronize(f->ctx->stream);
    dev_free(f->buf);
  }
  delete f;
}

int nq_field_upload(nq_field* f, const double* host)
{
  if (!f || !host) return fail(NQ_EINVAL, "null argument");
  nq_ctx* c = f->ctx;
  Guard G(c);
  const int nd = 2 * f->ncomp();
  const int64_t nsites = f->coverage == NQ_LX ? 2 * c->geo.Vh : c->geo.Vh;
  void* stg = nullptr;
  NQ_TRY(ctx_temp(c, "staging", sizeof(double) * nd * nsites, &stg));
  NQ_HIP(hipMemcpyAsync(stg, host, sizeof(double) * nd * nsites, hipMemcpyHostToDevice, c->stream));
  if (f->coverage == NQ_LX) {
    NQ_TRY(launch_aos_to_soa(c, (const double*)stg, f->ncomp(), 0, 1, view_of(f, 0)));
    NQ_TRY(launch_aos_to_soa(c, (const double*)stg, f->ncomp(), 1, 1, view_of(f, 1)));
  } else {
    NQ_TRY(launch_aos_to_soa(c, (const double*)stg, f->ncomp(), f->coverage, 0, view_of(f, 0)));
  }
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_field_download(const nq_field* f, double* host)
{
  if (!f || !host) return fail(NQ_EINVAL, "null argument");
  nq_ctx* c = f->ctx;
  Guard G(c);
  const int nd = 2 * f->ncomp();
  const int64_t nsites = f->coverage == NQ_LX ? 2 * c->geo.Vh : c->geo.Vh;
  void* stg = nullptr;
  NQ_TRY(ctx_temp(c, "staging", sizeof(double) * nd * nsites, &stg));
  if (f->coverage == NQ_LX) {
    NQ_TRY(launch_soa_to_aos(c, view_of(f, 0), f->ncomp(), 0, 1, (double*)stg));
    NQ_TRY(launch_soa_to_aos(c, view_of(f, 1), f->ncomp(), 1, 1, (double*)stg));
  } else {
    NQ_TRY(launch_soa_to_aos(c, view_of(f, 0), f->ncomp(), f->coverage, 0, (double*)stg));
  }
  NQ_HIP(hipMemcpyAsync(host, stg, sizeof(double) * nd * nsites, hipMemcpyDeviceToHost, c->stream));
  NQ_HIP(hipStreamSynchronize(c->stream));
  return NQ_OK;
}

int nq_field_copy(nq_field* dst, const nq_field* src)
{
  if (!dst || !src) return fail(NQ_EINVAL, "null argument");
  if (dst->ctx != src->ctx || dst->kind != src->kind || dst->coverage != src->coverage)
    return fail(NQ_EINVAL, "field copy: shape mismatch");
  Guard G(dst->ctx);
  NQ_HIP(hipMemcpyAsync(dst->buf.p, src->buf.p, src->buf.bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
  return NQ_OK;
}

int nq_field_norm2(const nq_field* f, double* out)
{
  if (!f || !out) return fail(NQ_EINVAL, "null argument");
  Guard G(f->ctx);
  SpinorView v = f->coverage == NQ_LX ? view_lx_flat(f) : view_of(f, 0);
  return reduce_norm2(f->ctx, v, out);
}

// ------------------------------------------------------------------ gauge
static double det3(const double* U, double* im)
{
  std::complex<double> m[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = {U[(i * 3 + j) * 2], U[(i * 3 + j) * 2 + 1]};
  const std::complex<double> d = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                                 m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  *im = d.imag();
  return d.real();
}

int nq_gauge_create_eo(nq_ctx* c, const double* ce, const double* co, nq_gauge** out)
{
  if (!c || !ce || !co || !out) return fail(NQ_EINVAL, "null argument");
  *out = nullptr;
  Guard G(c);
  const int64_t Vh = c->geo.Vh;
  auto* g = new nq_gauge;
  g->ctx = c;
  // 12-real compression needs det(U_mu) constant per direction (boundary phases are)
  bool comp = true;
  double dref[4][2];
  for (int mu = 0; mu < 4; mu++) dref[mu][0] = det3(ce + mu * 18, &dref[mu][1]);
  for (int p = 0; p < 2 && comp; p++) {
    const double* cf = p ? co : ce;
    for (int64_t e = 0; e < Vh && comp; e++)
      for (int mu = 0; mu < 4; mu++) {
        double im, re = det3(cf + (e * 4 + mu) * 18, &im);
        if (std::fabs(re - dref[mu][0]) > 1e-12 || std::fabs(im - dref[mu][1]) > 1e-12) { comp = false; break; }
      }
  }
  g->compressible = comp;
  std::memcpy(g->det, dref, sizeof dref);
  int rc = dev_alloc(g->aos[0], sizeof(double) * 72 * Vh);
  if (rc == NQ_OK) rc = dev_alloc(g->aos[1], sizeof(double) * 72 * Vh);
  if (rc == NQ_OK && hipMemcpyAsync(g->aos[0].p, ce, sizeof(double) * 72 * Vh, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = fail(NQ_EDEVICE, "gauge upload");
  if (rc == NQ_OK && hipMemcpyAsync(g->aos[1].p, co, sizeof(double) * 72 * Vh, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = fail(NQ_EDEVICE, "gauge upload");
  if (rc == NQ_OK) rc = gauge_build(g, LK_F64_18);
  if (rc == NQ_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(NQ_EDEVICE, "sync");
  if (rc != NQ_OK) {
    nq_gauge_destroy(g);
    return rc;
  }
  g->fp_eo[0] = conf_fingerprint(ce, 72 * Vh);
  g->fp_eo[1] = conf_fingerprint(co, 72 * Vh);
  *out = g;
  return NQ_OK;
}

int nq_gauge_create_lx(nq_ctx* c, const double* lx, nq_gauge** out)
{
  if (!c || !lx || !out) return fail(NQ_EINVAL, "null argument");
  const Geo& g = c->geo;
  std::vector<double> e(72 * g.Vh), o(72 * g.Vh);
  NQ_TRY(nq_host_split_lx(g.T, g.LX, g.LY, g.LZ, 72, lx, e.data(), o.data()));
  NQ_TRY(nq_gauge_create_eo(c, e.data(), o.data(), out));
  (*out)->fp_lx = conf_fingerprint(lx, 144 * g.Vh);
  (*out)->from_lx = true;
  return NQ_OK;
}

int nq_host_conf_fingerprint(const double* data, int64_t ndoubles, uint64_t* out)
{
  if (!data || !out || ndoubles < 0) return fail(NQ_EINVAL, "bad argument");
  *out = conf_fingerprint(data, ndoubles);
  return NQ_OK;
}

int nq_gauge_matches_lx(const nq_gauge* g, const double* lx, int* same)
{
  if (!g || !lx || !same) return fail(NQ_EINVAL, "null argument");
  if (!g->from_lx) return fail(NQ_EINVAL, "gauge handle was made from eo buffers");
  *same = conf_fingerprint(lx, 144 * g->ctx->geo.Vh) == g->fp_lx;
  return NQ_OK;
}

int nq_gauge_matches_eo(const nq_gauge* g, const double* ce, const double* co, int* same)
{
  if (!g || !ce || !co || !same) return fail(NQ_EINVAL, "null argument");
  const int64_t n = 72 * g->ctx->geo.Vh;
  *same = conf_fingerprint(ce, n) == g->fp_eo[0] && conf_fingerprint(co, n) == g->fp_eo[1];
  return NQ_OK;
}

void nq_gauge_destroy(nq_gauge* g)
{
  if (!g) return;
  {
    Guard G(g->ctx);
    (void)hipStreamSynchronize(g->ctx->stream);
    dev_free(g->aos[0]);
    dev_free(g->aos[1]);
    dev_free(g->l64);
    dev_free(g->l32);
    dev_free(g->l32c);
    dev_free(g->l16c);
    dev_free(g->l64c);
  }
  delete g;
}

int64_t nq_gauge_device_bytes(const nq_gauge* g) { return g ? g->bytes() : 0; }

int nq_gauge_plaquette(nq_ctx* c, const nq_gauge* g, double* plaq)
{
  if (!c || !g || !plaq) return fail(NQ_EINVAL, "null argument");
  if (g->ctx != c) return fail(NQ_EINVAL, "gauge handle belongs to another context");
  Guard G(c);
  double ts[2];
  NQ_TRY(launch_plaquette(c, (const double*)g->aos[0].p, (const double*)g->aos[1].p, ts));
  const double norm = 2.0 * (double)c->geo.Vh * 9.0;  // V sites x 3 planes x trace of 3
  *plaq = (ts[0] / norm + ts[1] / norm) / 2;
  return NQ_OK;
}

}  // extern "C"

namespace nq {

static DevBuf& link_buf(nq_gauge* g, int lk)
{
  switch (lk) {
  case LK_F64_18: return g->l64;
  case LK_F32_18: return g->l32;
  case LK_F32_12: return g->l32c;
  case LK_F64_12: return g->l64c;
  default: return g->l16c;
  }
}
static size_t link_elem(int lk) { return (size_t)lk_eb(lk); }
static int link_nc(int lk) { return lk_nc(lk); }

int gauge_build(nq_gauge* g, int lk)
{
  DevBuf& b = link_buf(g, lk);
  if (b.p) return NQ_OK;
  if (lk_nc(lk) == 6 && !g->compressible)
    return fail(NQ_EINVAL, "12-real link compression needs det(U_mu) constant per direction");
  nq_ctx* c = g->ctx;
  const size_t bytes = 2 * 8 * (size_t)link_nc(lk) * c->geo.stride * link_elem(lk);
  NQ_TRY(dev_alloc(b, bytes));
  NQ_TRY(launch_build_links(c, (const double*)g->aos[0].p, (const double*)g->aos[1].p, lk, b.p, g->det));
  return NQ_OK;
}

int link_view(const nq_gauge* gc, int lk, LinkView* v)
{
  nq_gauge* g = const_cast<nq_gauge*>(gc);
  NQ_TRY(gauge_build(g, lk));
  DevBuf& b = link_buf(g, lk);
  const size_t per = 8 * (size_t)link_nc(lk) * g->ctx->geo.stride * link_elem(lk);
  v->prec = lk;
  v->data[0] = b.p;
  v->data[1] = (const char*)b.p + per;
  for (int p = 0; p < 2; p++)
    for (int mu = 0; mu < 4; mu++) {
      v->det[p][2 * mu][0] = g->det[mu][0];
      v->det[p][2 * mu][1] = g->det[mu][1];
      v->det[p][2 * mu + 1][0] = g->det[mu][0];
      v->det[p][2 * mu + 1][1] = -g->det[mu][1];
    }
  return NQ_OK;
}

}  // namespace nq
