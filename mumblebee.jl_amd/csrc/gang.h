// Lockstep batching of independent IPM solves on one GPU ("gang").
//
// Each member is a Solver whose solve() runs on a host thread of its own, as a solo solve does,
// and takes every decision of the reference's step! for its own instance. Its device operations
// do not launch: they go to the gang as ops. The gang advances in ticks. A tick starts when every
// running member has handed in its next op; the member that completes the set leads the tick and
// issues the ops on the gang's single stream, then releases the members whose ops it issued:
//  * ops of one kind with equal parameters become one launch over the combined tables of all
//    members (tile lists, cone tables, fronts and elimination-tree levels concatenated, each
//    item carrying its member's slot in the instance table); the members of the launch have
//    S_ACTIVE = 1 in their scalar block and the others 0, which every kernel checks;
//  * the expensive ops (factorization, triangular solves) wait until every running member has
//    reached one, so that they meet: all large fronts of a level go into one persistent
//    factorization launch (tasks of all fronts interleaved by progress, each front's tasks in
//    their own topological order) and one pipelined solve launch;
//  * scalar read-backs of a tick become one gather kernel and one device-to-host copy;
//  * an op without a combined form, or of a member that cannot share it (vectors too long for
//    the batched dot products, PSD cones above the LDS order), runs as the member's own launch
//    sequence on the gang's stream.
// A member's launches stay in its program order (each waits for the tick that issues the
// previous one), and members touch disjoint device memory, so a member's results are those of
// its solo solve: the combined kernels compute each instance exactly as the solo ones do (fixed
// summation orders everywhere), and the tests compare bit for bit. Members join between ticks
// (continuous batching: a finished member's slot takes the next request) and leave when their
// IPM loop ends.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "kernels_big.h"
#include "solver.h"

namespace mb {

constexpr int GANG_MAX = GANG_SLOTS;  // members at a time (bits of the active mask; kernel argument arrays)

enum GKind : int {
  GK_NONE = 0,
  GK_SOLO,      // the member's own launch sequence (solo lambda)
  GK_READBACK,  // the scalar block into the member's host mirror
  GK_LEAVE,
  GK_SET2,      // sc[s1] = v1, sc[s2] = v2 (s2 < 0: one)
  GK_SMU,
  GK_CGSC,
  GK_LIN,
  GK_AXPBY,
  GK_AXPBYPCZ,
  GK_SPMV,
  GK_SPMVNM,
  GK_DOTS,
  GK_SCALE,     // scale! with ‖Q + cone blocks‖²
  GK_MATVEC,    // y += a (cone blocks) x
  GK_CORR0,
  GK_CORR,
  GK_MAXSTEP,
  GK_ZEROFREE,
  GK_FACTOR,    // assembly + factorization (expensive)
  GK_LSOLVE,    // forward solve (expensive)
  GK_USOLVE,    // backward solve (expensive)
  GK_COUNT
};

struct GOp {
  int kind = GK_NONE;
  bool park = false;       // expensive: waits until every running member has reached one
  bool solo_only = false;  // this member's op runs as its own launches
  int a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double v[3] = {0.0, 0.0, 0.0};  // per-member values (SET2: v1 v2; SMU: ν, μ, μt)
  Coef c[5] = {};
  LinSet N, M;
  SpmvN sn{};
  SpmvM sm{};
  DotBatch db;
  std::function<void()> solo;
  // completion (written by the tick's leader)
  std::atomic<bool> done{false};
  std::string err;
};

struct GangStats {
  double ticks = 0, launches = 0, ops = 0, combined_ops = 0, readbacks = 0, joins = 0;
  double factor_ms = 0, factor_flops = 0, factor_launches = 0, fronts_per_factor = 0;
  double dag_ms = 0, dag_launches = 0, dag_flops = 0;
};

class Gang {
 public:
  // kmax: members at a time (≤ GANG_MAX); dag_grid: workgroups of the persistent factorization
  // (0: one per CU); profile: time the combined factorizations with events
  Gang(int kmax, int dag_grid, bool profile);
  ~Gang();
  hipStream_t stream() const { return st_; }
  int capacity() const { return kmax_; }
  int members();
  // member side (called on the member's thread)
  void join(Solver* s);   // blocks until admitted between ticks
  void submit(Solver* s, GOp& op);
  void leave(Solver* s);
  void stats(GangStats& out);

 private:
  struct BigCombo;  // cached device data of one set of large fronts launched together
  void tick();      // the leader: admissions, then the ops of the tick
  void admit();
  void rebuild();
  void set_mask(uint64_t mask);
  void issue_group(const std::vector<int>& slots);
  void issue_solo(int slot);
  void readbacks(const std::vector<int>& slots);
  void combined_factor(const std::vector<int>& slots, bool fp32);
  void combined_solve(const std::vector<int>& slots, int kv, bool forward);
  BigCombo& combo(const std::vector<std::pair<int, int>>& fronts);
  bool can_combine(int slot, const GOp& op) const;
  void detach(int slot);

  int kmax_, dag_grid_;
  bool profile_;
  hipStream_t st_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr, evleave_ = nullptr;
  // membership and tick state (under mu_)
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<Solver*> mem_;       // by slot, nullptr when vacant
  std::vector<GOp*> op_;           // the op a member handed in and that is not yet issued
  std::vector<Solver*> joining_;
  int nrun_ = 0, narr_ = 0;
  bool leading_ = false;
  bool dirty_ = true;              // the combined tables need a rebuild
  uint64_t dev_mask_ = 0;          // S_ACTIVE = 1 on the device for these slots
  GangStats st_acc_;
  // combined device tables (rebuilt when the membership changes; leaving only clears a slot)
  DevBuf d_inst_, d_dead_, d_tn_, d_tm_, d_sp_[3], d_cones_, d_cone_ofs_, d_coneout_, d_psd_, d_dense_, d_soc_,
      d_dpart_, d_cone_front_, d_fr_, d_childlist_, d_gather_, d_flags_, d_items_, d_steps_, d_ea_, d_ea0_,
      d_tick_, d_stick_;
  double* h_gather_ = nullptr;     // pinned: GANG_MAX scalar blocks
  long long sbase_ = 0;            // solve ticket base
  int ntn_ = 0, ntm_ = 0, nsp_[3] = {0, 0, 0}, ncone_ = 0, npsd_ = 0, ndense_ = 0, nsoc_ = 0, dense_nmax_ = 0,
      cone_nmax_ = 0;
  struct GLevel {
    int items_off = 0, nsmall = 0;                     // small fronts (solve items)
    std::vector<int> ea_off, ea_n, ea_tiles;           // extend-add rounds
    std::vector<int> pan_off, pan_n, pan_blk, upd_off, upd_n, upd_blk;
    std::vector<std::pair<int, int>> big;              // (slot, member's large-front index)
  };
  std::vector<GLevel> lv_;
  std::vector<int> fofs_;                              // first combined front of each slot
  std::map<std::vector<int>, std::unique_ptr<BigCombo>> combos_;
  std::vector<InstDev> hinst_;                         // the instance table as uploaded
  InstDev dead_{};
  // launch data of fronts whose member left, freed once the stream has passed its launches
  std::vector<std::unique_ptr<BigCombo>> graveyard_;
  bool pending_timing_ = false;                        // a timed factorization awaits its read-back
};

}  // namespace mb
