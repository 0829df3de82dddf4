// Lockstep batching of independent IPM solves on one GPU ("gang").
//
// Each member is a Solver whose IPM loop takes every decision of the reference's step! for its own
// instance, as a solo solve does. The loops run as fibers of the gang's scheduler thread, and
// their device operations do not launch: they go to the gang as ops, in a queue per member. The
// gang advances in ticks: while every running member has an op queued, the scheduler takes the
// front op of each and issues them on the gang's single stream:
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
// A member hands in ops without waiting for them and switches back to the scheduler only at a
// read-back, a launch sequence with host results, or its leave; the scheduler then runs the other
// members' fibers and the ticks. A member's launches stay in its program order, and members touch
// disjoint device memory, so a member's results are those of
// its solo solve: the combined kernels compute each instance exactly as the solo ones do (fixed
// summation orders everywhere), and the tests compare bit for bit. Members join between ticks
// (continuous batching: a finished member's slot takes the next request) and leave when their
// IPM loop ends.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <ucontext.h>

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
  bool blocking = false;   // the member waits for it (read-back, own launches with host results, leave)
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
  double ticks = 0, launches = 0, ops = 0, combined_ops = 0, readbacks = 0, joins = 0, members = 0;
  double factor_ms = 0, factor_flops = 0, factor_launches = 0, fronts_per_factor = 0;
  double dag_ms = 0, dag_launches = 0, dag_flops = 0;
  // the scheduler thread's time (ms): admitting members (tables rebuilt and uploaded), running the
  // members' fibers (their host logic), issuing ticks, and waiting in read-backs for the device
  double admit_ms = 0, fiber_ms = 0, tick_ms = 0, wait_ms = 0;
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
  // solve!(s) as a member: waits for a free slot, runs the IPM loop in lockstep with the other
  // members and returns its status (any thread; blocks until the solve ended)
  int run(Solver* s);
  // on a member's fiber: queue an op (heap-allocated; the gang deletes it once issued, or, for a
  // blocking op, returns after it was issued and deletes it then; an error of an earlier op of
  // the member is thrown here)
  void submit(Solver* s, GOp* op);
  void stats(GangStats& out);

 private:
  struct BigCombo;  // cached device data of one set of large fronts launched together
  struct Fiber;
  struct Job;
  static constexpr size_t FIBER_STACK = size_t(8) << 20;  // reserved, committed as touched
  static void fiber_entry(unsigned lo, unsigned hi);
  void sched_main();
  void fail_fronts(const std::string& what);
  void tick();      // the front op of every member: one tick
  void admit();
  bool try_admit();  // admit(); on failure the waiting jobs fail. False if nothing was waiting
  bool factor_issued_ = false;  // the last tick issued a combined factorization and no read-back
  void rebuild();
  void set_mask(uint64_t mask);
  void issue_group(const std::vector<int>& slots);
  void issue_solo(int slot);
  void readbacks(const std::vector<int>& slots);
  void combined_factor(const std::vector<int>& slots, bool fp32);
  void combined_solve(const std::vector<int>& slots, int kv, bool forward, bool fp32);
  BigCombo& combo(const std::vector<std::pair<int, int>>& fronts);
  bool can_combine(int slot, const GOp& op) const;
  void detach(int slot);
  bool ready() const;  // every running member has an op queued

  int kmax_, dag_grid_;
  bool profile_;
  hipStream_t st_ = nullptr;
  hipEvent_t evleave_ = nullptr;
  // profile: event pairs around each combined factorization kernel, read at the next read-back
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed_;
  std::vector<hipEvent_t> evpool_;
  hipEvent_t event();
  // the scheduler thread runs the members' fibers and the ticks; other threads only hand in jobs
  // and collect them (under mu_)
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::deque<Job*> joining_;
  bool stop_ = false;
  std::thread sched_;
  ucontext_t sched_ctx_;
  GangStats st_pub_;               // the statistics as of the last scheduler round (under mu_)
  // scheduler-thread state
  std::vector<Solver*> mem_;       // by slot, nullptr when vacant
  std::vector<std::unique_ptr<Fiber>> fib_;
  std::vector<std::deque<GOp*>> q_; // each member's ops handed in and not yet issued
  std::vector<GOp*> op_;           // the tick's front ops
  std::vector<std::string> err_;   // the first error of a member's ops that nobody waited for
  int nrun_ = 0;
  bool dirty_ = true;              // the combined tables need a rebuild
  uint64_t dev_mask_ = 0;          // S_ACTIVE = 1 on the device for these slots
  GangStats st_acc_;
  // combined device tables, packed in one block (rebuilt when members join; a leaving member's
  // slot is cleared in place)
  struct Tabs {
    InstDev* inst = nullptr;
    Tile *tn = nullptr, *tm = nullptr, *sp[3] = {nullptr, nullptr, nullptr};
    ConeDesc* cones = nullptr;
    int *cone_ofs = nullptr, *psd = nullptr, *dense = nullptr, *soc = nullptr, *cone_front = nullptr;
    double *coneout = nullptr, *dpart = nullptr;
    FrontDesc* fr = nullptr;
    int *childlist = nullptr, *items = nullptr, *ea = nullptr, *ea0 = nullptr, *flags = nullptr;
    StepDesc* steps = nullptr;
    GangAsm* asmv = nullptr;
  } T_;
  long long fill_blocks_ = 0, union_blocks_ = 0;
  struct Pack;
  std::unique_ptr<DevBuf> upload(const Pack& p);
  void passed();
  std::unique_ptr<DevBuf> tab_;
  std::vector<std::unique_ptr<DevBuf>> old_tabs_;
  std::vector<std::pair<void*, size_t>> pin_free_, pin_busy_;  // pinned staging blocks
  DevBuf d_dead_, d_gather_, d_tick_, d_stick_;
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
};

}  // namespace mb
