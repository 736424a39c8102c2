// Slot-based batched even-odd twisted-mass solver.
//
// Up to 16 sources that share the gauge field and kappa are solved together, each
// with its own twisted mass (inv_tmD_cg_eoprec, cg_invert_tmD_eoprec.cpp:102-150, for
// every slot): eq. 8 prepares the odd source of a slot, the K K^+ CG runs for all
// slots in the same launches (per-slot coefficient rows and solver states), and
// eqs. 10-11 rebuild the full solution of a finished slot.  A slot can be started
// while the others iterate, so new work joins a running batch at the next chunk
// boundary (continuous batching).  The arithmetic of a slot does not depend on the
// other slots or on its slot index, so a source gives the same bits alone and in
// any mix.
//
// With a clover handle the batch solves the twisted clover eo system instead
// (inv_tmclovD_cg_eoprec, cg_invert_tmclovD_eoprec.cpp:146-232, for every slot) at one
// twisted mass for the whole group: K = g5 (D_oo - 1/4 H D_ee^-1 H) with the clover
// blocks applied to every slot by batched block kernels that read each site's block
// once, and the inverse blocks of the group's (kappa, mu) kept by the group.
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>

#include "nq_internal.hpp"

namespace nq {

class TmBatch {
 public:
  TmBatch() = default;
  TmBatch(const TmBatch&) = delete;
  TmBatch& operator=(const TmBatch&) = delete;
  ~TmBatch();
  static constexpr int kMaxSlots = 16;
  static constexpr int kMaxGroups = 3;  // state ranges [16*group, 16*group+16) of the ctx

  // prec: NQ_PREC_*; group selects the state range and the names of the temporaries;
  // cl (optional): twisted clover at twisted mass clover_mu for every slot
  int setup(nq_ctx* ctx, const nq_gauge* g, double kappa, const double* anis, int prec, int group,
            nq_clover* cl = nullptr, double clover_mu = 0);
  // eq. 8 for the source (src_e, src_o FP64 views of one lx field) into `slot`, then
  // the start of its CG (x = guess or 0, r = b - K K^+ x); `residue` is nissa's squared
  // relative residue
  int start(int slot, double mu, double residue, int niter, const SpinorView& src_e, const SpinorView& src_o,
            const SpinorView* guess = nullptr);
  // enqueue `iters` iterations over the active slots (no host synchronisation)
  int run(int iters);
  // wait for the stream, run the FP64 reliable updates that are due, and append the
  // slots that have finished (converged, out of iterations or NaN) to `finished`
  int poll(std::vector<int>* finished);
  // eqs. 10-11 into (sol_e, sol_o) and the FP64 check |D sol - src|^2/|src|^2 with the
  // 18-real links (guess_out, optional: receives the eq. 9 solution); returns
  // NQ_ENOCONV / NQ_ENAN for a slot that did not converge
  int finish(int slot, const SpinorView& src_e, const SpinorView& src_o, const SpinorView& sol_e,
             const SpinorView& sol_o, nq_solve_stats* st, double* check, const SpinorView* guess_out = nullptr);
  int suggested_chunk() const { return chunk_; }
  bool active(int slot) const { return active_[slot]; }
  int nactive() const;
  int prec() const { return prec_; }
  double kappa() const { return kappa_; }
  bool clover() const { return cl_ != nullptr; }
  // device bytes of the per-slot work vectors (clover: two more per precision)
  static int64_t bytes_per_slot(const Geo& g, int prec, bool clover = false);

 private:
  int hi() const;
  int lk64() const;  // links of the FP64 work: the tm solves' (12-real when mixed), FP64 18-real for clover
  int upload_row(int slot, double mu);
  int enqueue(int nr, int iters);
  int run_graphs(int nr, int iters);
  // K(dag) on nr slots of the batch vectors: out = K in (clover), with the last hop's
  // epilogue from `last` (the coefficient table of the K out hop)
  int clov_kern_batched(const LinkView& U, double anis, int dag, const SpinorView& out, const SpinorView& in,
                        const SpinorView& tmp, const SpinorView& te2, const SpinorView& uo, const SpinorView& inv,
                        const SpinorView& clo, HopEpi last, int nr, int pred);

  nq_ctx* ctx_ = nullptr;
  const nq_gauge* g_ = nullptr;
  double kappa_ = 0, anis_ = 1;
  bool has_anis_ = false;
  int prec_ = 0, sp_ = 0, lk_ = 0, group_ = 0;
  bool mixed_ = false;
  CgState* S_ = nullptr;
  double* coef_ = nullptr;
  std::vector<double> hcoef_;
  SpinorView b_, x64_, r64_, t64_, tmp64_, xs_, rs_, p_, t1s_, tmps_, et_;
  bool active_[kMaxSlots] = {};
  double mu_[kMaxSlots] = {};
  int niter_[kMaxSlots] = {};
  int launched_[kMaxSlots] = {};  // iterations enqueued since the slot started
  int chunk_ = 8;
  nq_clover* cl_ = nullptr;
  double clmu_ = 0;
  SpinorView clinv64_, clinv_s_, clo64_, clo_s_;  // even inverse blocks (group's), odd clover blocks
  SpinorView te2_64_, uo64_, te2s_, uos_;         // batched even / odd block products
  SpinorView cte2_;                               // one even FP64 field (single-slot K in finish)
  std::map<int, hipGraphExec_t> graphs_;  // (slots, 2^k iterations) -> captured chunk
  std::set<int> eager_done_;              // slot counts that ran eagerly once
};

}  // namespace nq
