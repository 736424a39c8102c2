// Propagator-solve server: continuous batching of eo twisted-mass solves.
//
// The caller of this path in nissa is the propagator driver, which solves 12
// spin-colour sources per (mass, kappa) on one gauge configuration
// (projects/ib/prop.cpp:41-106,925-964, through inv_tmD_cg_eoprec).  Here such
// requests arrive over time for several lattice volumes:
//  * every volume has an engine: its own context (device stream), its gauge field
//    resident in HBM, and one host thread that drives its batches;
//  * a batch group is a TmBatch of 16 slots; a queued source takes any free slot of
//    a group with the same kappa and inner precision, at the next chunk boundary,
//    while the other slots keep iterating (continuous batching); sources with
//    different twisted masses share the launches;
//  * a new group is created only when its device footprint fits the HBM budget
//    (admission by footprint); otherwise the sources wait for a slot;
//  * engines of different volumes run on their own streams and threads, so their
//    kernels overlap on the device.
// A source's arithmetic is independent of the other slots and of the slot index
// (TmBatch), so its solution is bit-identical alone and in any mix.  Requests with a
// cSW are twisted clover solves (inv_tmclovD_cg_eoprec, as the driver calls it when
// its input sets cSW, projects/ib/prop.cpp:93-96): the engine builds the clover term
// of its gauge field once per cSW, and such sources batch per (kappa, mu, cSW).

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "nq_internal.hpp"
#include "tm_batch.hpp"

using namespace nq;

namespace {

struct Req {
  int64_t id = 0;
  nq_serve_request r{};
  std::vector<const double*> src;
  std::vector<double*> sol;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int remaining = 0;
  int status = NQ_OK;
  int iters_max = 0, nrel = 0;
  double max_chk = 0;
  bool done = false;
};

struct RhsRef {
  std::shared_ptr<Req> req;
  int k = 0;
};

struct Group {
  TmBatch b;
  double kappa = 0;
  int prec = 0;
  double csw = 0;  // > 0: a twisted clover group, whose sources share one mu
  double mu = 0;
  RhsRef slot[TmBatch::kMaxSlots];
  bool pending[TmBatch::kMaxSlots] = {};  // assigned, not started yet
  nq_field* src[TmBatch::kMaxSlots] = {};
  nq_field* sol[TmBatch::kMaxSlots] = {};
  int64_t bytes = 0;
  bool ready = false;
  bool draining = false;  // takes no new sources: it is to be replaced by a group of
                          // another kappa / precision once its running sources finish
  ~Group()
  {
    for (int i = 0; i < TmBatch::kMaxSlots; i++) {
      nq_field_destroy(src[i]);
      nq_field_destroy(sol[i]);
    }
  }
  bool has_free() const
  {
    for (int i = 0; i < TmBatch::kMaxSlots; i++)
      if (!slot[i].req) return true;
    return false;
  }
  bool busy() const
  {
    for (int i = 0; i < TmBatch::kMaxSlots; i++)
      if (slot[i].req) return true;
    return false;
  }
  int nbusy() const
  {
    int n = 0;
    for (int i = 0; i < TmBatch::kMaxSlots; i++) n += slot[i].req != nullptr;
    return n;
  }
  // the parameters a source must share with the group's (kappa, precision; cSW and,
  // for twisted clover, mu)
  bool same(const Req& q) const
  {
    return kappa == q.r.kappa && prec == q.r.prec && csw == q.r.csw && (csw == 0 || mu == q.r.mu);
  }
  bool fits(const Req& q) const { return !draining && same(q) && has_free(); }
};

struct Engine {
  int id = 0;
  int dims[4] = {};
  nq_ctx* ctx = nullptr;
  nq_gauge* g = nullptr;
  std::deque<RhsRef> queue;
  // batch groups by index: group i owns the solver states [16 i, 16 i + 16) of the
  // context and the temporaries named "tb<i>.", so a group is replaced in place
  std::unique_ptr<Group> groups[TmBatch::kMaxGroups];
  std::map<double, nq_clover*> clovers;  // clover term of the resident gauge per cSW (made on first use)
  int64_t gauge_bytes = 0;
  std::thread th;
};

}  // namespace

struct nq_server {
  int device = 0;
  int64_t budget = 0;
  int max_groups = 1;
  std::mutex mu;
  std::condition_variable cv_done;
  std::vector<std::unique_ptr<Engine>> engines;
  std::vector<std::unique_ptr<std::condition_variable>> cv_work;
  std::map<int64_t, std::shared_ptr<Req>> reqs;
  int64_t next_id = 1;
  int64_t used = 0;         // resident gauges + admitted groups
  int64_t gauges = 0;       // resident gauges alone
  bool stop = false;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int64_t group_bytes(const nq_ctx* c, int prec, bool clover)
{
  const int64_t s = c->geo.stride;
  const int64_t lx = 2 * 12 * 16 * s;  // one FP64 lx spincolor field
  int64_t b = TmBatch::kMaxSlots * (TmBatch::bytes_per_slot(c->geo, prec, clover) + 2 * lx) + (64 << 20);
  // clover: the group's inverse blocks (FP64, and a sloppy copy when mixed: 8 bytes per
  // complex in FP32, 4 in FP16) and, counted with every group, the clover term of both
  // parities with its sloppy odd copy
  const int64_t cs = prec == NQ_PREC_F64 ? 0 : (prec == NQ_PREC_F32 ? 8 : 4);
  if (clover) b += 72 * 16 * s + 72 * cs * s + 2 * 36 * 16 * s + 36 * cs * s + 12 * 16 * s;
  return b;
}

// point source: 1 at lx site `site` and spin-colour component sc (id*3+ic), 0 elsewhere
int point_source(nq_ctx* c, nq_field* f, int64_t site, int sc)
{
  NQ_TRY(launch_zero(c, view_lx_flat(f)));
  const Geo& g = c->geo;
  int64_t r = site;
  const int z = (int)(r % g.LZ);
  r /= g.LZ;
  const int y = (int)(r % g.LY);
  r /= g.LY;
  const int x = (int)(r % g.LX);
  const int t = (int)(r / g.LX);
  const int64_t dlx = device_rank(g, t, x, y, z);
  const int par = (t + x + y + z) & 1;
  static const double one[2] = {1.0, 0.0};
  char* p = (char*)f->buf.p + ((int64_t)(par * 12 + sc) * g.stride + (dlx >> 1)) * 16;
  NQ_HIP(hipMemcpyAsync(p, one, sizeof one, hipMemcpyHostToDevice, c->stream));
  return NQ_OK;
}

void finish_source(nq_server* s, Req& q, int k, int rc, const nq_solve_stats& st)
{
  q.iters_max = std::max(q.iters_max, st.iters);
  q.nrel += st.reliable_updates;
  q.max_chk = std::max(q.max_chk, st.rel_res_sq);
  if (rc != NQ_OK && q.status == NQ_OK) q.status = rc;
  (void)k;
  if (--q.remaining == 0) {
    q.t_done = s->now();
    q.done = true;
  }
}

// Admission of the queued sources of one engine (called under s->mu).  Every queued
// source is tried in order; one that cannot start now stays queued and the scan goes
// on past it.  A source goes into a free slot of a group with its kappa and
// precision; otherwise into a new group at a free index when the group's footprint
// fits the budget; otherwise, when every index is taken, one group is marked
// draining: it takes no new sources, and once its running sources have finished it
// is released and replaced by a group for the waiting source, so a request of a new
// kappa or precision cannot wait forever behind a stream of compatible ones.  A
// source whose group can never fit (resident gauges plus one group above the
// budget) is refused with NQ_EBUSY.
struct Admission {
  std::vector<std::pair<Group*, int>> starts;
  std::vector<int> fresh;                        // group indices to set up
  std::vector<std::unique_ptr<Group>> released;  // idle groups to free (outside the lock)
};

void admit(nq_server* s, Engine* E, Admission* A)
{
  nq_ctx* c = E->ctx;
  const int ng = std::min(s->max_groups, TmBatch::kMaxGroups);
  bool drain_marked = false;
  for (auto it = E->queue.begin(); it != E->queue.end();) {
    const Req& q = *it->req;
    Group* sel = nullptr;
    for (int gi = 0; gi < ng && !sel; gi++)
      if (E->groups[gi] && E->groups[gi]->fits(q)) sel = E->groups[gi].get();
    const int64_t need = group_bytes(c, q.r.prec, q.r.csw > 0);
    if (!sel && s->gauges + need > s->budget) {
      // no group of this precision fits even with every other group released
      RhsRef ref = *it;
      it = E->queue.erase(it);
      if (ref.req->t_admit < 0) ref.req->t_admit = s->now();
      finish_source(s, *ref.req, ref.k, NQ_EBUSY, nq_solve_stats{});
      if (ref.req->done) s->cv_done.notify_all();
      continue;
    }
    if (!sel) {
      // a group of this kappa and precision that is merely full: wait for its slots
      // (or take a free index), never displace another group for it
      bool compat = false;
      for (int i = 0; i < ng; i++) compat |= E->groups[i] && !E->groups[i]->draining && E->groups[i]->same(q);
      int gi = -1;
      for (int i = 0; i < ng && gi < 0; i++)
        if (!E->groups[i]) gi = i;
      // otherwise an idle group (draining or not) gives way
      for (int i = 0; i < ng && gi < 0 && !compat; i++)
        if (E->groups[i] && !E->groups[i]->busy()) gi = i;
      const int64_t freed = gi >= 0 && E->groups[gi] ? E->groups[gi]->bytes : 0;
      if (gi >= 0 && s->used - freed + need <= s->budget) {
        if (E->groups[gi]) {
          s->used -= freed;
          A->released.push_back(std::move(E->groups[gi]));
        }
        auto gp = std::make_unique<Group>();
        gp->kappa = q.r.kappa;
        gp->prec = q.r.prec;
        gp->csw = q.r.csw;
        gp->mu = q.r.mu;
        gp->bytes = need;
        s->used += need;
        sel = gp.get();
        E->groups[gi] = std::move(gp);
        A->fresh.push_back(gi);
      } else if (gi < 0 && !compat && !drain_marked) {
        // every index holds a busy group of other parameters: drain the one with the
        // fewest sources (unless one drains already) so that it can be replaced
        bool any = false;
        for (int i = 0; i < ng; i++) any |= E->groups[i] && E->groups[i]->draining;
        if (!any) {
          int best = -1;
          for (int i = 0; i < ng; i++)
            if (E->groups[i] && (best < 0 || E->groups[i]->nbusy() < E->groups[best]->nbusy())) best = i;
          if (best >= 0) E->groups[best]->draining = true;
        }
        drain_marked = true;
      }
    }
    if (!sel) {
      ++it;  // waits for a slot, for a group to drain or for memory
      continue;
    }
    int i = 0;
    while (sel->slot[i].req) i++;
    sel->slot[i] = *it;
    sel->pending[i] = true;
    if (it->req->t_admit < 0) it->req->t_admit = s->now();
    A->starts.push_back({sel, i});
    it = E->queue.erase(it);
  }
}

void engine_loop(nq_server* s, Engine* E)
{
  (void)hipSetDevice(s->device);
  nq_ctx* c = E->ctx;
  std::condition_variable& cv = *s->cv_work[E->id];
  for (;;) {
    Admission A;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop) return;
      admit(s, E, &A);
      bool active = !A.starts.empty() || !A.released.empty();
      for (auto& gp : E->groups) active |= gp && gp->busy();
      if (!active) {
        if (E->queue.empty()) cv.wait(lk, [&] { return s->stop || !E->queue.empty(); });
        else cv.wait_for(lk, std::chrono::milliseconds(5));  // memory held by other volumes
        continue;
      }
    }
    // released groups go first (batches, graphs, fields); the temporaries of an index
    // that gets a new group are freed too, since a group of another precision names
    // some of them differently and would leave the old ones allocated
    A.released.clear();
    for (int gi : A.fresh) (void)ctx_temp_release(c, "tb" + std::to_string(gi) + ".");
    // new groups: device set-up (group index = its slot in the engine)
    for (int gi : A.fresh) {
      Group* gp = E->groups[gi].get();
      nq_clover* cl = nullptr;
      int rc = NQ_OK;
      if (gp->csw > 0) {
        auto it = E->clovers.find(gp->csw);
        if (it == E->clovers.end()) {
          rc = nq_clover_create(c, E->g, gp->csw, &cl);
          if (rc == NQ_OK) E->clovers[gp->csw] = cl;
        } else {
          cl = it->second;
        }
      }
      if (rc == NQ_OK) rc = gp->b.setup(c, E->g, gp->kappa, nullptr, gp->prec, gi, cl, gp->mu);
      for (int i = 0; rc == NQ_OK && i < TmBatch::kMaxSlots; i++) {
        rc = nq_field_create(c, NQ_SPINCOLOR, NQ_LX, &gp->src[i]);
        if (rc == NQ_OK) rc = nq_field_create(c, NQ_SPINCOLOR, NQ_LX, &gp->sol[i]);
      }
      gp->ready = rc == NQ_OK;
    }
    // start the admitted sources
    for (auto& [gp, i] : A.starts) {
      RhsRef ref = gp->slot[i];
      Req& q = *ref.req;
      int rc = gp->ready ? NQ_OK : NQ_EOOM;
      if (rc == NQ_OK)
        rc = q.r.src_kind == 1 ? point_source(c, gp->src[i], q.r.site0, ref.k) : nq_field_upload(gp->src[i], q.src[ref.k]);
      if (rc == NQ_OK)
        rc = gp->b.start(i, q.r.mu, q.r.residue, q.r.niter, view_of(gp->src[i], 0), view_of(gp->src[i], 1));
      gp->pending[i] = false;
      if (rc != NQ_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        finish_source(s, q, ref.k, rc, nq_solve_stats{});
        gp->slot[i] = RhsRef{};
        if (q.done) s->cv_done.notify_all();
      }
    }
    // iterate every group that has work, then poll and retire finished sources
    for (auto& gp : E->groups)
      if (gp && gp->ready && gp->b.nactive() > 0) (void)gp->b.run(std::min(gp->b.suggested_chunk(), 32));
    for (auto& gp : E->groups) {
      if (!gp || !gp->ready || gp->b.nactive() == 0) continue;
      std::vector<int> fin;
      int prc = gp->b.poll(&fin);
      if (prc != NQ_OK) {
        // device error: fail every source of the group
        fin.clear();
        for (int i = 0; i < TmBatch::kMaxSlots; i++)
          if (gp->slot[i].req && gp->b.active(i)) fin.push_back(i);
      }
      for (int i : fin) {
        RhsRef ref = gp->slot[i];
        Req& q = *ref.req;
        nq_solve_stats st{};
        double chk = 0;
        int rc = gp->b.finish(i, view_of(gp->src[i], 0), view_of(gp->src[i], 1), view_of(gp->sol[i], 0),
                              view_of(gp->sol[i], 1), &st, &chk);
        if (prc != NQ_OK) rc = prc;
        if (rc != NQ_ENAN && q.sol[ref.k]) {
          const int d = nq_field_download(gp->sol[i], q.sol[ref.k]);
          if (rc == NQ_OK) rc = d;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        finish_source(s, q, ref.k, rc, st);
        gp->slot[i] = RhsRef{};
        if (q.done) s->cv_done.notify_all();
      }
    }
  }
}

}  // namespace

extern "C" {

int nq_server_create(int device, int64_t hbm_budget_bytes, int max_groups_per_volume, nq_server** out)
{
  if (!out) return fail(NQ_EINVAL, "null argument");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return fail(NQ_EDEVICE, "no device %d", device);
  auto s = std::make_unique<nq_server>();
  s->device = device;
  s->max_groups = std::max(1, std::min(max_groups_per_volume, TmBatch::kMaxGroups));
  if (hbm_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    NQ_HIP(hipSetDevice(device));
    NQ_HIP(hipMemGetInfo(&fr, &tot));
    hbm_budget_bytes = (int64_t)fr - ((int64_t)4 << 30);
  }
  s->budget = hbm_budget_bytes;
  *out = s.release();
  return NQ_OK;
}

int nq_server_add_volume(nq_server* s, int T, int LX, int LY, int LZ, const double* conf_lx, int* volume_id)
{
  if (!s || !conf_lx || !volume_id) return fail(NQ_EINVAL, "null argument");
  auto E = std::make_unique<Engine>();
  E->dims[0] = T, E->dims[1] = LX, E->dims[2] = LY, E->dims[3] = LZ;
  NQ_TRY(nq_ctx_create(T, LX, LY, LZ, s->device, &E->ctx));
  int rc = nq_gauge_create_lx(E->ctx, conf_lx, &E->g);
  if (rc != NQ_OK) {
    nq_ctx_destroy(E->ctx);
    return rc;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  E->gauge_bytes = nq_gauge_device_bytes(E->g);
  s->used += E->gauge_bytes;
  s->gauges += E->gauge_bytes;
  E->id = (int)s->engines.size();
  *volume_id = E->id;
  s->cv_work.push_back(std::make_unique<std::condition_variable>());
  Engine* e = E.get();
  s->engines.push_back(std::move(E));
  e->th = std::thread(engine_loop, s, e);
  return NQ_OK;
}

int nq_server_submit(nq_server* s, const nq_serve_request* r, int64_t* id)
{
  if (!s || !r) return fail(NQ_EINVAL, "null argument");
  if (r->nsrc < 1) return fail(NQ_EINVAL, "request without sources");
  if (r->src_kind == 1 && r->nsrc > 12) return fail(NQ_EINVAL, "point sources: at most 12 spin-colour components");
  if (r->src_kind == 0 && !r->src) return fail(NQ_EINVAL, "host sources missing");
  if (r->src_kind != 0 && r->src_kind != 1) return fail(NQ_EINVAL, "unknown source kind %d", r->src_kind);
  if (r->prec < NQ_PREC_F64 || r->prec > NQ_PREC_F16H) return fail(NQ_EINVAL, "unknown precision %d", r->prec);
  if (!(r->residue > 0) || r->niter <= 0 || !(r->kappa != 0)) return fail(NQ_EINVAL, "bad kappa/residue/niter");
  if (!(r->csw >= 0)) return fail(NQ_EINVAL, "bad cSW %g", r->csw);
  auto q = std::make_shared<Req>();
  q->r = *r;
  for (int k = 0; k < r->nsrc; k++) {
    if (r->src_kind == 0 && !r->src[k]) return fail(NQ_EINVAL, "null source %d", k);
    q->src.push_back(r->src_kind == 0 ? r->src[k] : nullptr);
    q->sol.push_back(r->sol ? r->sol[k] : nullptr);
  }
  q->r.src = nullptr;
  q->r.sol = nullptr;
  q->remaining = r->nsrc;
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->volume < 0 || r->volume >= (int)s->engines.size()) return fail(NQ_EINVAL, "unknown volume %d", r->volume);
  Engine* E = s->engines[r->volume].get();
  const int64_t V = (int64_t)E->dims[0] * E->dims[1] * E->dims[2] * E->dims[3];
  if (r->src_kind == 1 && (r->site0 < 0 || r->site0 >= V))
    return fail(NQ_EINVAL, "point source site %lld outside the volume (%lld sites)", (long long)r->site0, (long long)V);
  q->id = s->next_id++;
  q->t_submit = s->now();
  s->reqs[q->id] = q;
  for (int k = 0; k < r->nsrc; k++) E->queue.push_back(RhsRef{q, k});
  s->cv_work[r->volume]->notify_one();
  if (id) *id = q->id;
  return NQ_OK;
}

int nq_server_wait(nq_server* s, int64_t id, double timeout_s, nq_serve_result* out)
{
  if (!s || !out) return fail(NQ_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(NQ_EINVAL, "unknown request %lld", (long long)id);
  std::shared_ptr<Req> q = it->second;
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  if (!s->cv_done.wait_until(lk, until, [&] { return q->done; })) return fail(NQ_EBUSY, "request %lld still running", (long long)id);
  out->id = q->id;
  out->status = q->status;
  out->nsrc = q->r.nsrc;
  out->iters_max = q->iters_max;
  out->reliable_updates = q->nrel;
  out->max_rel_res_sq = q->max_chk;
  out->t_submit = q->t_submit;
  out->t_admit = q->t_admit;
  out->t_done = q->t_done;
  s->reqs.erase(id);
  return NQ_OK;
}

double nq_server_clock(const nq_server* s) { return s ? s->now() : 0.0; }

int64_t nq_server_device_bytes(const nq_server* s) { return s ? s->used : 0; }

void nq_server_destroy(nq_server* s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& cv : s->cv_work) cv->notify_all();
  }
  for (auto& E : s->engines)
    if (E->th.joinable()) E->th.join();
  for (auto& E : s->engines) {
    for (auto& gp : E->groups) gp.reset();
    for (auto& kv : E->clovers) nq_clover_destroy(kv.second);
    nq_gauge_destroy(E->g);
    nq_ctx_destroy(E->ctx);
  }
  delete s;
}

}  // extern "C"
