#include "gang.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <queue>
#include <stdexcept>
#include <thread>

#include <sys/mman.h>
#include <ucontext.h>

namespace mb {

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +           \
                                                   hipGetErrorString(e_) + " at " #x);     \
  } while (0)

// device data of one set of large fronts launched together: their BigFront copies (instance
// index = gang slot), front-table indices, and the interleaved ticket orders of the
// factorization's tasks and of the two solves' blocks
struct Gang::BigCombo {
  std::unique_ptr<DevBuf> buf;
  BigFront* bs = nullptr;
  int* bfront = nullptr;
  int2 *odag = nullptr, *ofs = nullptr, *obs = nullptr;
  int nf = 0, ndag = 0, nfs = 0, nbs = 0, max2tr = 0, maxtr = 0, maxch = 0;
  DagQueues q{};
  double flops = 0.0;
};

// Several host arrays packed into one device block (parts 256-byte aligned), copied with one
// asynchronous transfer: no stream synchronisation, so the device goes on with the running launches
struct Gang::Pack {
  std::vector<char> h;
  template <class T>
  size_t add(const std::vector<T>& v) {
    const size_t off = (h.size() + 255) & ~size_t(255);
    h.resize(off + std::max<size_t>(v.size(), 1) * sizeof(T), 0);
    if (!v.empty()) std::memcpy(h.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
  size_t zeros(size_t bytes) {  // a part the device writes (or needs zeroed)
    const size_t off = (h.size() + 255) & ~size_t(255);
    h.resize(off + std::max<size_t>(bytes, 8), 0);
    return off;
  }
};

// the packed arrays into a new device block, through a pinned staging block; both stay alive
// until a stream synchronisation shows that every launch before has passed
std::unique_ptr<DevBuf> Gang::upload(const Pack& p) {
  auto d = std::make_unique<DevBuf>();
  d->alloc(p.h.size());
  void* hp = nullptr;
  for (size_t i = 0; i < pin_free_.size(); ++i)
    if (pin_free_[i].second >= p.h.size()) {
      hp = pin_free_[i].first;
      pin_busy_.push_back(pin_free_[i]);
      pin_free_.erase(pin_free_.begin() + (long)i);
      break;
    }
  if (!hp) {
    const size_t c = std::max<size_t>(p.h.size(), size_t(1) << 20);
    HIPCHK(hipHostMalloc(&hp, c, hipHostMallocDefault));
    pin_busy_.push_back({hp, c});
  }
  std::memcpy(hp, p.h.data(), p.h.size());
  HIPCHK(hipMemcpyAsync(d->p, hp, p.h.size(), hipMemcpyHostToDevice, st_));
  return d;
}

// after a synchronisation of the gang's stream: what earlier launches used can go
void Gang::passed() {
  graveyard_.clear();
  old_tabs_.clear();
  for (auto& b : pin_busy_) pin_free_.push_back(b);
  pin_busy_.clear();
}

namespace {

// items i = 0..c_j−1 of lists j merged by progress (i + ½)/c_j, ties to the smaller j: every list
// keeps its own order, and lists of different lengths advance together
std::vector<int2> interleave(const std::vector<int>& counts) {
  std::vector<int2> out;
  size_t tot = 0;
  for (int c : counts) tot += (size_t)std::max(c, 0);
  out.reserve(tot);
  using E = std::pair<double, int>;
  std::priority_queue<E, std::vector<E>, std::greater<E>> q;
  std::vector<int> next(counts.size(), 0);
  for (size_t j = 0; j < counts.size(); ++j)
    if (counts[j] > 0) q.push(E{0.5 / counts[j], (int)j});
  while (!q.empty()) {
    const int j = q.top().second;
    q.pop();
    out.push_back(int2{j, next[j]});
    if (++next[j] < counts[j]) q.push(E{(next[j] + 0.5) / counts[j], j});
  }
  return out;
}

bool same_d(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }
bool same_coef(const Coef& a, const Coef& b) { return same_d(a.c, b.c) && a.slot == b.slot && a.inv == b.inv; }
bool same_lin(const Lin& a, const Lin& b) {
  return a.kx == b.kx && a.ky == b.ky && a.kw == b.kw && a.kz == b.kz && same_coef(a.a, b.a) &&
         same_coef(a.b, b.b) && same_coef(a.c, b.c);
}
bool same_linset(const LinSet& a, const LinSet& b) {
  if (a.n != b.n) return false;
  for (int k = 0; k < a.n; ++k)
    if (!same_lin(a.op[k], b.op[k])) return false;
  return true;
}
bool same_spmvn(const SpmvN& a, const SpmvN& b) {
  return a.kxq == b.kxq && a.kxt == b.kxt && a.kz == b.kz && a.kw == b.kw && a.kv == b.kv && a.ky == b.ky &&
         a.kyq == b.kyq && same_coef(a.aq, b.aq) && same_coef(a.at, b.at) && same_coef(a.c, b.c) &&
         same_coef(a.d, b.d) && same_coef(a.e, b.e);
}
bool same_spmvm(const SpmvM& a, const SpmvM& b) {
  return a.kx == b.kx && a.kz == b.kz && a.ky == b.ky && a.ky2 == b.ky2 && a.y2set == b.y2set &&
         same_coef(a.a, b.a) && same_coef(a.c, b.c) && same_coef(a.e, b.e);
}
bool same_dots(const DotBatch& a, const DotBatch& b) {
  if (a.n != b.n) return false;
  for (int k = 0; k < a.n; ++k) {
    const DotSpec &x = a.s[k], &y = b.s[k];
    if (x.space != y.space || x.kx != y.kx || x.ky != y.ky || x.mode != y.mode || x.kw0 != y.kw0 ||
        x.kw1 != y.kw1 || x.slot != y.slot || x.op != y.op)
      return false;
  }
  return true;
}

// ops that may share one launch: same kind, same parameters (per-member values apart)
bool same_op(const GOp& x, const GOp& y) {
  if (x.kind != y.kind) return false;
  for (int k = 0; k < 8; ++k)
    if (x.a[k] != y.a[k]) return false;
  for (int k = 0; k < 5; ++k)
    if (!same_coef(x.c[k], y.c[k])) return false;
  switch (x.kind) {
    case GK_LIN: return same_linset(x.N, y.N) && same_linset(x.M, y.M);
    case GK_SPMVNM: return same_spmvn(x.sn, y.sn) && same_spmvm(x.sm, y.sm);
    case GK_DOTS: return same_dots(x.db, y.db);
    default: return true;
  }
}

}  // namespace

Gang::Gang(int kmax, int dag_grid, bool profile)
    : kmax_(std::max(1, std::min(kmax, GANG_MAX))), dag_grid_(dag_grid), profile_(profile) {
  HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&evleave_, hipEventDisableTiming));
  if (dag_grid_ <= 0) {
    int dev = 0, ncu = 256;
    HIPCHK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
    dag_grid_ = ncu;
  }
  mem_.assign(kmax_, nullptr);
  op_.assign(kmax_, nullptr);
  q_.assign(kmax_, {});
  err_.assign(kmax_, {});
  fib_.resize(kmax_);
  HIPCHK(hipHostMalloc((void**)&h_gather_, sizeof(double) * GANG_MAX * NSCALAR, hipHostMallocDefault));
  d_gather_.alloc(sizeof(double) * GANG_MAX * NSCALAR);
  d_dead_.alloc(sizeof(double) * NSCALAR);
  d_tick_.alloc(64);
  d_stick_.alloc(64);
  HIPCHK(hipMemsetAsync(d_dead_.p, 0, d_dead_.bytes, st_));
  HIPCHK(hipMemsetAsync(d_tick_.p, 0, d_tick_.bytes, st_));
  HIPCHK(hipMemsetAsync(d_stick_.p, 0, d_stick_.bytes, st_));
  HIPCHK(hipStreamSynchronize(st_));
  dead_ = InstDev{};
  dead_.sc = d_dead_.as<double>();  // S_ACTIVE stays 0: every kernel skips a vacant slot
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  sched_ = std::thread([this, dev] {
    (void)hipSetDevice(dev);
    sched_main();
  });
}

Gang::~Gang() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (sched_.joinable()) sched_.join();
  (void)hipStreamSynchronize(st_);
  passed();
  combos_.clear();
  tab_.reset();
  for (auto& b : pin_free_) (void)hipHostFree(b.first);
  if (h_gather_) (void)hipHostFree(h_gather_);
  for (auto& p : timed_) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  for (hipEvent_t e : evpool_) (void)hipEventDestroy(e);
  (void)hipEventDestroy(evleave_);
  (void)hipStreamDestroy(st_);
}

hipEvent_t Gang::event() {
  if (evpool_.empty()) {
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = evpool_.back();
  evpool_.pop_back();
  return e;
}

int Gang::members() {
  std::lock_guard<std::mutex> lk(mu_);
  return (int)st_pub_.members;
}

void Gang::stats(GangStats& out) {
  std::lock_guard<std::mutex> lk(mu_);
  out = st_pub_;
}

// ------------------------------------------------------------------------------------------
// membership, fibers and ticks
// ------------------------------------------------------------------------------------------

// A member's IPM loop runs on a fiber (its own stack, switched with swapcontext) of the gang's
// scheduler thread: handing in an op is a function call, and a blocking op switches back to the
// scheduler, which runs the other members' fibers and the ticks. No OS thread waits on another
// within a solve.
struct Gang::Fiber {
  ucontext_t ctx;
  void* stack = nullptr;
  size_t stack_bytes = 0;
  Solver* s = nullptr;
  Job* job = nullptr;
  bool runnable = true;
  bool finished = false;
};

struct Gang::Job {
  Solver* s = nullptr;
  int status = -1;
  std::string err;
  bool done = false;
};

int Gang::run(Solver* s) {
  if (s->gang_) throw std::logic_error("solver is already in a gang");
  Job job;
  job.s = s;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("gang is shutting down");
    joining_.push_back(&job);
  }
  cv_.notify_all();
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return job.done; });
  if (!job.err.empty()) throw std::runtime_error(job.err);
  return job.status;
}

void Gang::fiber_entry(unsigned lo, unsigned hi) {
  Fiber* f = reinterpret_cast<Fiber*>((uintptr_t)lo | ((uintptr_t)hi << 32));
  Gang* g = f->s->gang_;
  Solver* s = f->s;
  try {
    int status = CONTINUE;
    while (status == CONTINUE) status = s->step();
    f->job->status = status;
  } catch (const std::exception& e) {
    f->job->err = e.what();
  }
  try {
    GOp* op = new GOp;
    op->kind = GK_LEAVE;
    op->blocking = true;
    g->submit(s, op);
  } catch (const std::exception& e) {
    if (f->job->err.empty()) f->job->err = e.what();
  }
  f->finished = true;
  swapcontext(&f->ctx, &g->sched_ctx_);  // not resumed again
}

// the waiting solvers take free slots (between ticks, on the scheduler thread); the tables are
// rebuilt and each new member gets a fiber
void Gang::admit() {
  // a slot is free once its last member's fiber has ended too
  std::vector<int> slots;
  for (int k = 0; k < kmax_; ++k)
    if (!mem_[k] && !fib_[k]) slots.push_back(k);
  std::vector<Job*> take;
  {
    std::lock_guard<std::mutex> lk(mu_);
    while (!joining_.empty() && take.size() < slots.size()) {
      take.push_back(joining_.front());
      joining_.pop_front();
    }
  }
  if (take.empty()) return;
  for (size_t t = 0; t < take.size(); ++t) {
    Job* j = take[t];
    Solver* s = j->s;
    const int k = slots[t];
    HIPCHK(hipStreamSynchronize(s->st_));  // its setup and starting point are on the device
    mem_[k] = s;
    s->gslot_ = k;
    s->own_st_ = s->st_;
    s->st_ = st_;
    s->gang_prof_ = s->prof_mask_;
    s->prof_mask_ = 0;  // per-kernel events are the gang's (its combined launches serve all members)
    s->gang_ = this;
    auto f = std::make_unique<Fiber>();
    f->s = s;
    f->job = j;
    f->stack_bytes = FIBER_STACK;
    f->stack = mmap(nullptr, f->stack_bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                    -1, 0);
    if (f->stack == MAP_FAILED) throw std::runtime_error("gang: no memory for a fiber stack");
    getcontext(&f->ctx);
    f->ctx.uc_stack.ss_sp = f->stack;
    f->ctx.uc_stack.ss_size = f->stack_bytes;
    f->ctx.uc_link = nullptr;
    const uintptr_t pf = reinterpret_cast<uintptr_t>(f.get());
    makecontext(&f->ctx, (void (*)())&Gang::fiber_entry, 2, (unsigned)(pf & 0xffffffffu), (unsigned)(pf >> 32));
    fib_[k] = std::move(f);
    ++nrun_;
    st_acc_.joins += 1;
  }
  rebuild();
}

bool Gang::try_admit() {
  std::string fail;
  try {
    admit();
  } catch (const std::exception& e) {
    fail = e.what();
  }
  if (fail.empty()) return true;
  // the waiting jobs cannot be admitted: they fail
  std::lock_guard<std::mutex> lk(mu_);
  for (Job* j : joining_) {
    j->err = fail;
    j->done = true;
  }
  joining_.clear();
  cv_done_.notify_all();
  return false;
}

void Gang::sched_main() {
  for (;;) {
    bool fibers = false;
    for (auto& f : fib_) fibers = fibers || f != nullptr;
    if (!fibers) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !joining_.empty(); });
      if (joining_.empty()) return;  // stop_ with nothing left
    }
    const auto ta = std::chrono::steady_clock::now();
    const bool failed = !try_admit();
    const auto tf = std::chrono::steady_clock::now();
    st_acc_.admit_ms += std::chrono::duration<double, std::milli>(tf - ta).count();
    if (failed && nrun_ == 0) continue;
    // every runnable member runs until it waits on an op (or ends)
    for (int k = 0; k < kmax_; ++k) {
      Fiber* f = fib_[k].get();
      if (!f || !f->runnable) continue;
      f->runnable = false;
      swapcontext(&sched_ctx_, &f->ctx);
      if (f->finished) {
        Job* j = f->job;
        munmap(f->stack, f->stack_bytes);
        fib_[k].reset();
        std::lock_guard<std::mutex> lk(mu_);
        j->done = true;
        cv_done_.notify_all();
      }
    }
    const auto tt = std::chrono::steady_clock::now();
    st_acc_.fiber_ms += std::chrono::duration<double, std::milli>(tt - tf).count();
    // ticks while every member has an op queued
    while (ready()) {
      factor_issued_ = false;
      try {
        tick();
      } catch (const std::exception& e) {
        fail_fronts(e.what());
      }
      // solvers waiting to join are admitted while the combined factorization just issued runs
      // (their tables are rebuilt and uploaded behind it) rather than with the device idle after
      // a read-back; they start at the next round
      if (factor_issued_) {
        const auto ta2 = std::chrono::steady_clock::now();
        (void)try_admit();
        st_acc_.admit_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta2).count();
      }
    }
    st_acc_.tick_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tt).count();
    {
      std::lock_guard<std::mutex> lk(mu_);
      st_pub_ = st_acc_;
      st_pub_.members = nrun_;
    }
  }
}

// a failure outside any one op (table rebuild, read-back): the front ops fail, and the members
// that were leaving are gone
void Gang::fail_fronts(const std::string& what) {
  for (int j = 0; j < kmax_; ++j) {
    if (!mem_[j] || q_[j].empty()) continue;
    GOp* o = q_[j].front();
    q_[j].pop_front();
    if (o->kind == GK_LEAVE) {
      detach(j);
      mem_[j] = nullptr;
      --nrun_;
    }
    if (o->blocking) {
      o->err = what;
      o->done.store(true, std::memory_order_relaxed);
      if (fib_[j]) fib_[j]->runnable = true;
    } else {
      if (err_[j].empty()) err_[j] = what;
      delete o;
    }
  }
}

bool Gang::ready() const {
  if (nrun_ == 0) return false;
  for (int k = 0; k < kmax_; ++k)
    if (mem_[k] && q_[k].empty()) return false;
  return true;
}

// on a member's fiber: queue the op; a blocking op switches to the scheduler until it was issued
void Gang::submit(Solver* s, GOp* op) {
  const int k = s->gslot_;
  op->done.store(false, std::memory_order_relaxed);
  q_[k].push_back(op);
  if (!op->blocking) return;
  Fiber* f = fib_[k].get();
  swapcontext(&f->ctx, &sched_ctx_);
  std::string err = op->err;
  delete op;
  if (err.empty() && s->gang_ == this && !err_[k].empty()) std::swap(err, err_[k]);
  if (!err.empty()) throw std::runtime_error(err);
}

void Gang::tick() {
  std::vector<int> cheap, parked;
  bool all = true;
  for (int k = 0; k < kmax_; ++k) {
    op_[k] = nullptr;
    if (!mem_[k]) continue;
    if (q_[k].empty()) {
      all = false;
      continue;
    }
    op_[k] = q_[k].front();
    (op_[k]->park ? parked : cheap).push_back(k);
  }
  // Cheap ops first. With every member parked: the triangular solves, and the factorizations only
  // once every member waits on one, so that the factorizations of all members meet in one launch
  // (a member that joined late, or refined more, catches up by the next factorization)
  std::vector<int> issue;
  if (!cheap.empty()) {
    issue = cheap;
  } else if (all) {
    for (int k : parked)
      if (op_[k]->kind != GK_FACTOR) issue.push_back(k);
    if (issue.empty()) issue = parked;
  }
  if (issue.empty()) return;
  st_acc_.ticks += 1;
  // groups of ops with one combined form, in slot order
  std::vector<int> rb, lv, rest;
  for (int k : issue) {
    const int kind = op_[k]->kind;
    if (kind == GK_READBACK) rb.push_back(k);
    else if (kind == GK_LEAVE) lv.push_back(k);
    else rest.push_back(k);
  }
  std::vector<char> used(kmax_, 0);
  bool fac = false;
  for (size_t i = 0; i < rest.size(); ++i) {
    const int k = rest[i];
    if (used[k]) continue;
    std::vector<int> g{k};
    used[k] = 1;
    if (can_combine(k, *op_[k]))
      for (size_t j = i + 1; j < rest.size(); ++j) {
        const int k2 = rest[j];
        if (!used[k2] && can_combine(k2, *op_[k2]) && same_op(*op_[k], *op_[k2])) {
          g.push_back(k2);
          used[k2] = 1;
        }
      }
    st_acc_.ops += (double)g.size();
    const int kind = op_[k]->kind;
    const bool heavy = kind == GK_FACTOR || kind == GK_LSOLVE || kind == GK_USOLVE;
    fac = fac || kind == GK_FACTOR;
    // a lone cheap op runs as the member's own launch; a lone factorization or solve still takes
    // the combined path, whose persistent kernels use every CU
    if (g.size() == 1 && !(heavy && can_combine(k, *op_[k]))) {
      issue_solo(k);
    } else {
      st_acc_.combined_ops += (double)g.size();
      try {
        issue_group(g);
      } catch (const std::exception& e) {
        for (int j : g) op_[j]->err = e.what();
      }
    }
  }
  if (!rb.empty()) readbacks(rb);
  factor_issued_ = fac && rb.empty();
  for (int k : lv) {
    Solver* s = mem_[k];
    set_mask(dev_mask_ & ~(1ull << k));
    hinst_[k] = dead_;
    launch_gang_setinst(st_, T_.inst, k, dead_);
    launch_set_scalar(st_, s->d_sc_.as<double>(), S_ACTIVE, 1.0);
    HIPCHK(hipEventRecord(evleave_, st_));
    HIPCHK(hipStreamWaitEvent(s->own_st_, evleave_, 0));  // its own stream starts after its last op here
    detach(k);
    // cached launch data may name this slot's fronts: kept until the stream has passed them
    for (auto& kv : combos_) graveyard_.push_back(std::move(kv.second));
    combos_.clear();
  }
  for (int k : issue) {
    GOp* o = op_[k];
    op_[k] = nullptr;
    q_[k].pop_front();
    if (o->kind == GK_LEAVE) {
      mem_[k] = nullptr;
      --nrun_;
      for (GOp* r : q_[k]) delete r;  // nothing follows a leave
      q_[k].clear();
      if (o->err.empty()) o->err.swap(err_[k]);  // an unreported failure of its last ops
      err_[k].clear();
    }
    if (o->blocking) {
      o->done.store(true, std::memory_order_relaxed);
      if (fib_[k]) fib_[k]->runnable = true;
    } else {
      if (!o->err.empty() && err_[k].empty()) err_[k] = o->err;
      delete o;
    }
  }
}

// the member's own stream and profiling back (its slot is cleared by the caller)
void Gang::detach(int k) {
  Solver* s = mem_[k];
  s->st_ = s->own_st_;
  s->gang_ = nullptr;
  s->gslot_ = -1;
  s->prof_mask_ = s->gang_prof_;
}

bool Gang::can_combine(int slot, const GOp& op) const {
  (void)slot;
  if (op.solo_only) return false;
  switch (op.kind) {
    case GK_SET2: case GK_SMU: case GK_CGSC: case GK_LIN: case GK_AXPBY: case GK_AXPBYPCZ: case GK_SPMV:
    case GK_SPMVNM: case GK_DOTS: case GK_SCALE: case GK_MATVEC: case GK_CORR0: case GK_CORR: case GK_MAXSTEP:
    case GK_ZEROFREE: case GK_FACTOR: case GK_LSOLVE: case GK_USOLVE:
      return true;
    default:
      return false;
  }
}

void Gang::set_mask(uint64_t mask) {
  if (mask == dev_mask_) return;
  uint64_t occ = 0;
  for (int k = 0; k < kmax_; ++k)
    if (mem_[k]) occ |= 1ull << k;
  launch_gang_active(st_, T_.inst, kmax_, mask & occ, occ);
  dev_mask_ = mask & occ;
  st_acc_.launches += 1;
}

void Gang::issue_solo(int k) {
  Solver* s = mem_[k];
  GOp& op = *op_[k];
  if (!((dev_mask_ >> k) & 1ull)) set_mask(dev_mask_ | (1ull << k));
  Solver::set_issuing(s);
  try {
    if (op.solo) op.solo();
  } catch (const std::exception& e) {
    op.err = e.what();
  }
  Solver::set_issuing(nullptr);
}

void Gang::readbacks(const std::vector<int>& slots) {
  uint64_t mask = 0;
  for (int k : slots) mask |= 1ull << k;
  launch_gang_gather(st_, T_.inst, kmax_, mask, d_gather_.as<double>());
  HIPCHK(hipMemcpyAsync(h_gather_, d_gather_.p, sizeof(double) * NSCALAR * kmax_, hipMemcpyDeviceToHost, st_));
  const auto tw = std::chrono::steady_clock::now();
  HIPCHK(hipStreamSynchronize(st_));
  st_acc_.wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
  passed();  // the stream has passed every launch that could read them
  for (int k : slots) std::memcpy(mem_[k]->hsc_, h_gather_ + (size_t)k * NSCALAR, sizeof(double) * NSCALAR);
  st_acc_.readbacks += 1;
  st_acc_.launches += 1;
  for (auto& p : timed_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) {
      st_acc_.factor_ms += ms;
      st_acc_.dag_ms += ms;
    }
    evpool_.push_back(p.first);
    evpool_.push_back(p.second);
  }
  timed_.clear();
}

// ------------------------------------------------------------------------------------------
// combined tables
// ------------------------------------------------------------------------------------------

void Gang::rebuild() {
  // cached launch data names the old front table's indices
  for (auto& kv : combos_) graveyard_.push_back(std::move(kv.second));
  combos_.clear();
  const int K = kmax_;
  hinst_.assign(K, dead_);
  std::vector<Tile> tn, tm, sp[3];
  std::vector<ConeDesc> cones;
  std::vector<int> cone_ofs(K + 1, 0), psd, dense, soc, cone_front, childlist;
  std::vector<FrontDesc> fr;
  fofs_.assign(K + 1, 0);
  cone_nmax_ = 0;
  dense_nmax_ = 0;
  int H = 0;
  for (int k = 0; k < K; ++k) {
    Solver* s = mem_[k];
    cone_ofs[k] = (int)cones.size();
    fofs_[k] = (int)fr.size();
    if (!s) continue;
    hinst_[k] = s->idev_;
    const HostLists& h = s->hl_;
    for (const Tile& t : h.tn) tn.push_back(Tile{k, t.start});
    for (const Tile& t : h.tm) tm.push_back(Tile{k, t.start});
    for (int w = 0; w < 3; ++w)
      for (const Tile& t : h.sp[w]) sp[w].push_back(Tile{k, t.start});
    const int c0 = (int)cones.size(), f0 = (int)fr.size(), ch0 = (int)childlist.size();
    for (ConeDesc C : s->cones_) {
      C.inst = k;
      cones.push_back(C);
    }
    for (int v : h.psd) psd.push_back(v + c0);
    for (int v : h.dense) dense.push_back(v + c0);
    for (int v : h.soc) soc.push_back(v + c0);
    for (int f : s->cone_front_) cone_front.push_back(f + f0);
    for (FrontDesc F : s->fr_) {
      F.inst = k;
      if (F.parent >= 0) F.parent += f0;
      F.child0 += ch0;
      fr.push_back(F);
    }
    for (int c : h.childlist) childlist.push_back(c + f0);
    cone_nmax_ = std::max(cone_nmax_, s->cone_nmax_);
    dense_nmax_ = std::max(dense_nmax_, s->dense_nmax_);
    H = std::max(H, (int)s->levels_.size());
  }
  cone_ofs[K] = (int)cones.size();
  fofs_[K] = (int)fr.size();
  ntn_ = (int)tn.size();
  ntm_ = (int)tm.size();
  for (int w = 0; w < 3; ++w) nsp_[w] = (int)sp[w].size();
  ncone_ = (int)cones.size();
  npsd_ = (int)psd.size();
  ndense_ = (int)dense.size();
  nsoc_ = (int)soc.size();
  // levels: level l of every member (levels count from the leaves)
  lv_.assign(H, GLevel{});
  std::vector<int> items, ea, ea0;
  std::vector<StepDesc> steps;
  for (int l = 0; l < H; ++l) {
    GLevel& G = lv_[l];
    G.items_off = (int)items.size();
    size_t R = 0, S = 0;
    for (int k = 0; k < K; ++k) {
      Solver* s = mem_[k];
      if (!s || l >= (int)s->levels_.size()) continue;
      const Level& L = s->levels_[l];
      for (int i = L.items_off; i < L.items_off + L.nsmall; ++i) items.push_back(s->hl_.solve_items[i] + fofs_[k]);
      R = std::max(R, L.ea_off.size());
      S = std::max(S, L.pan_off.size());
      for (int b : L.big) G.big.push_back({k, b});
    }
    G.nsmall = (int)items.size() - G.items_off;
    for (size_t r = 0; r < R; ++r) {
      const int off = (int)ea.size();
      int tiles = 0;
      for (int k = 0; k < K; ++k) {
        Solver* s = mem_[k];
        if (!s || l >= (int)s->levels_.size()) continue;
        const Level& L = s->levels_[l];
        if (r >= L.ea_off.size()) continue;
        for (int i = L.ea_off[r]; i < L.ea_off[r] + L.ea_n[r]; ++i) {
          ea.push_back(s->hl_.ea_items[i] + fofs_[k]);
          ea0.push_back(s->hl_.ea_tile0[i] + tiles);
        }
        tiles += L.ea_tiles[r];
      }
      G.ea_off.push_back(off);
      G.ea_n.push_back((int)ea.size() - off);
      G.ea_tiles.push_back(tiles);
    }
    for (size_t q = 0; q < S; ++q) {
      for (int part = 0; part < 2; ++part) {
        const int off = (int)steps.size();
        int blk = 0;
        for (int k = 0; k < K; ++k) {
          Solver* s = mem_[k];
          if (!s || l >= (int)s->levels_.size()) continue;
          const Level& L = s->levels_[l];
          if (q >= L.pan_off.size()) continue;
          const int o = part == 0 ? L.pan_off[q] : L.upd_off[q];
          const int n = part == 0 ? L.pan_n[q] : L.upd_n[q];
          for (int i = o; i < o + n; ++i) {
            StepDesc d = s->hl_.steps[i];
            d.front += fofs_[k];
            d.tile0 += blk;
            steps.push_back(d);
          }
          blk += part == 0 ? L.pan_blk[q] : L.upd_blk[q];
        }
        (part == 0 ? G.pan_off : G.upd_off).push_back(off);
        (part == 0 ? G.pan_n : G.upd_n).push_back((int)steps.size() - off);
        (part == 0 ? G.pan_blk : G.upd_blk).push_back(blk);
      }
    }
  }
  // the members' assembly arrays (vacant slots: empty)
  std::vector<GangAsm> asmv(K);
  long long fb = 0, ub = 0;
  for (int k = 0; k < K; ++k) {
    GangAsm& A = asmv[k];
    A = GangAsm{nullptr, 0, nullptr, nullptr, nullptr, 0, fb, ub};
    Solver* s = mem_[k];
    if (!s) continue;
    A.F = s->d_F_.as<double>();
    const bool part = s->fill_free_ && s->static_ok_;  // no clearing; only the δ-dependent entries
    A.nF = part ? 0 : s->fsize_;
    A.pos = s->d_upos_.as<int64_t>();
    A.l = s->d_ul_.as<double>();
    A.q = s->d_uq_.as<double>();
    A.nent = part ? s->nunion_dyn_ : s->nunion_;
    fb += (A.nF / 2 + TILE_THREADS - 1) / TILE_THREADS + (A.nF & 1 ? 1 : 0);
    ub += (A.nent + TILE_THREADS - 1) / TILE_THREADS;
  }
  fill_blocks_ = fb;
  union_blocks_ = ub;
  Pack P;
  const size_t o_asm = P.add(asmv);
  const size_t o_inst = P.add(hinst_), o_tn = P.add(tn), o_tm = P.add(tm), o_sp0 = P.add(sp[0]),
               o_sp1 = P.add(sp[1]), o_sp2 = P.add(sp[2]), o_cones = P.add(cones), o_cofs = P.add(cone_ofs),
               o_cout = P.zeros(sizeof(double) * 3 * std::max(ncone_, 1)), o_psd = P.add(psd), o_dense = P.add(dense),
               o_soc = P.add(soc), o_dpart = P.zeros(sizeof(double) * DENSE_SPLIT * std::max(ndense_, 1)),
               o_cf = P.add(cone_front), o_fr = P.add(fr), o_cl = P.add(childlist), o_items = P.add(items),
               o_steps = P.add(steps), o_ea = P.add(ea), o_ea0 = P.add(ea0),
               o_flags = P.zeros(sizeof(int) * std::max<size_t>(steps.size(), 1));
  if (tab_) old_tabs_.push_back(std::move(tab_));  // launches queued before may still read it
  tab_ = upload(P);
  char* b = static_cast<char*>(tab_->p);
  T_.inst = reinterpret_cast<InstDev*>(b + o_inst);
  T_.tn = reinterpret_cast<Tile*>(b + o_tn);
  T_.tm = reinterpret_cast<Tile*>(b + o_tm);
  T_.sp[0] = reinterpret_cast<Tile*>(b + o_sp0);
  T_.sp[1] = reinterpret_cast<Tile*>(b + o_sp1);
  T_.sp[2] = reinterpret_cast<Tile*>(b + o_sp2);
  T_.cones = reinterpret_cast<ConeDesc*>(b + o_cones);
  T_.cone_ofs = reinterpret_cast<int*>(b + o_cofs);
  T_.coneout = reinterpret_cast<double*>(b + o_cout);
  T_.psd = reinterpret_cast<int*>(b + o_psd);
  T_.dense = reinterpret_cast<int*>(b + o_dense);
  T_.soc = reinterpret_cast<int*>(b + o_soc);
  T_.dpart = reinterpret_cast<double*>(b + o_dpart);
  T_.cone_front = reinterpret_cast<int*>(b + o_cf);
  T_.fr = reinterpret_cast<FrontDesc*>(b + o_fr);
  T_.childlist = reinterpret_cast<int*>(b + o_cl);
  T_.items = reinterpret_cast<int*>(b + o_items);
  T_.steps = reinterpret_cast<StepDesc*>(b + o_steps);
  T_.ea = reinterpret_cast<int*>(b + o_ea);
  T_.ea0 = reinterpret_cast<int*>(b + o_ea0);
  T_.flags = reinterpret_cast<int*>(b + o_flags);
  T_.asmv = reinterpret_cast<GangAsm*>(b + o_asm);
  // every member takes part until a launch says otherwise
  uint64_t occ = 0;
  for (int k = 0; k < K; ++k)
    if (mem_[k]) occ |= 1ull << k;
  launch_gang_active(st_, T_.inst, K, occ, occ);
  dev_mask_ = occ;
  dirty_ = false;
}

Gang::BigCombo& Gang::combo(const std::vector<std::pair<int, int>>& fronts) {
  std::vector<int> key;
  for (auto& p : fronts) {
    key.push_back(p.first);
    key.push_back(p.second);
  }
  auto it = combos_.find(key);
  if (it != combos_.end()) return *it->second;
  auto c = std::make_unique<BigCombo>();
  std::vector<BigFront> bs;
  std::vector<int> bfront, cdag, cfs, cbs;
  for (auto& p : fronts) {
    Solver* s = mem_[p.first];
    const auto& bh = *s->big_[p.second];
    BigFront B = bh.B;
    B.inst = p.first;
    bs.push_back(B);
    bfront.push_back(fofs_[p.first] + bh.front);
    cdag.push_back(B.ntask);
    cfs.push_back(B.T);
    cbs.push_back(B.Tr);
    c->max2tr = std::max(c->max2tr, 2 * B.Tr);
    c->maxtr = std::max(c->maxtr, B.Tr);
    c->maxch = std::max(c->maxch, (B.s + NB - 1) / NB);
    const double r = B.r, sz = B.s;
    c->flops += r * r * r / 3.0 + sz * r * r + sz * sz * r;
  }
  // factorization tasks: with at least DAG_QUEUES fronts, one queue per XCD, the fronts dealt to
  // the queues by work (largest first, to the least loaded queue); fewer fronts share one queue
  std::vector<int2> odag;
  const int nf = (int)bs.size();
  c->q.n = nf >= DAG_QUEUES ? DAG_QUEUES : 1;
  {
    std::vector<std::vector<int>> qf(c->q.n);
    std::vector<double> load(c->q.n, 0.0);
    std::vector<int> byw(nf);
    for (int f = 0; f < nf; ++f) byw[f] = f;
    std::stable_sort(byw.begin(), byw.end(), [&](int a, int b) { return cdag[a] > cdag[b]; });
    for (int f : byw) {
      const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      qf[q].push_back(f);
      load[q] += cdag[f];
    }
    c->q.off[0] = 0;
    for (int q = 0; q < c->q.n; ++q) {
      std::sort(qf[q].begin(), qf[q].end());
      std::vector<int> cnt;
      for (int f : qf[q]) cnt.push_back(cdag[f]);
      for (const int2& e : interleave(cnt)) odag.push_back(int2{qf[q][e.x], e.y});
      c->q.off[q + 1] = (int)odag.size();
    }
  }
  const std::vector<int2> ofs = interleave(cfs), obs = interleave(cbs);
  c->nf = (int)bs.size();
  c->ndag = (int)odag.size();
  c->nfs = (int)ofs.size();
  c->nbs = (int)obs.size();
  Pack P;
  const size_t o_bs = P.add(bs), o_bf = P.add(bfront), o_dag = P.add(odag), o_fs = P.add(ofs), o_bsv = P.add(obs);
  c->buf = upload(P);
  char* b = static_cast<char*>(c->buf->p);
  c->bs = reinterpret_cast<BigFront*>(b + o_bs);
  c->bfront = reinterpret_cast<int*>(b + o_bf);
  c->odag = reinterpret_cast<int2*>(b + o_dag);
  c->ofs = reinterpret_cast<int2*>(b + o_fs);
  c->obs = reinterpret_cast<int2*>(b + o_bsv);
  BigCombo& ref = *c;
  combos_[key] = std::move(c);
  return ref;
}

// ------------------------------------------------------------------------------------------
// combined launches
// ------------------------------------------------------------------------------------------

void Gang::combined_factor(const std::vector<int>& slots, bool fp32) {
  uint64_t mask = 0;
  for (int k : slots) mask |= 1ull << k;
  set_mask(mask);
  const InstDev* I = T_.inst;
  const FrontDesc* F = T_.fr;
  // assembly: the members' fronts cleared, BᵀB + δQ scattered and the cone blocks added, each
  // step one launch for all
  launch_gang_assemble(st_, I, T_.asmv, kmax_, fill_blocks_, union_blocks_);
  launch_assemble_cones(st_, I, T_.cones, T_.cone_front, F, ncone_, cone_nmax_);
  st_acc_.launches += 3;
  const int epoch = next_epoch();
  double flops = 0.0;
  int nf = 0;
  for (size_t l = 0; l < lv_.size(); ++l) {
    const GLevel& G = lv_[l];
    for (size_t r = 0; r < G.ea_off.size(); ++r)
      launch_extend_add(st_, F, T_.ea + G.ea_off[r], T_.ea0 + G.ea_off[r], G.ea_n[r],
                        G.ea_tiles[r], I);
    for (size_t q = 0; q < G.pan_off.size(); ++q) {
      launch_panel(st_, F, T_.steps + G.pan_off[q], G.pan_n[q], G.pan_blk[q], I,
                   T_.flags + G.pan_off[q], epoch, nullptr);
      if (G.upd_blk[q] > 0) launch_update(st_, F, T_.steps + G.upd_off[q], G.upd_n[q], G.upd_blk[q], I);
    }
    std::vector<std::pair<int, int>> big;
    for (auto& p : G.big)
      if ((mask >> p.first) & 1ull) big.push_back(p);
    if (!big.empty()) {
      BigCombo& c = combo(big);
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (profile_) {  // the factorization kernel alone, read at the next read-back
        e0 = event();
        e1 = event();
        timed_.push_back({e0, e1});
      }
      launch_dag_factor_multi(st_, c.bs, c.nf, c.odag, c.q, d_tick_.as<int>(), I, dag_grid_, epoch, fp32,
                              c.max2tr, e0, e1);
      flops += c.flops;
      nf += c.nf;
      st_acc_.dag_launches += 1;
      st_acc_.launches += 2;
    }
    st_acc_.launches += (double)(G.ea_off.size() + 2 * G.pan_off.size());
  }
  st_acc_.factor_launches += 1;
  st_acc_.factor_flops += flops;
  st_acc_.dag_flops += flops;
  st_acc_.fronts_per_factor += nf;
}

void Gang::combined_solve(const std::vector<int>& slots, int kv, bool forward, bool fp32) {
  uint64_t mask = 0;
  for (int k : slots) mask |= 1ull << k;
  set_mask(mask);
  const InstDev* I = T_.inst;
  const FrontDesc* F = T_.fr;
  const int H = (int)lv_.size();
  for (int i = 0; i < H; ++i) {
    const int l = forward ? i : H - 1 - i;
    const GLevel& G = lv_[l];
    if (forward) launch_fsolve(st_, F, T_.items + G.items_off, G.nsmall, T_.childlist, I, kv);
    else launch_bsolve(st_, F, T_.items + G.items_off, G.nsmall, I, kv);
    st_acc_.launches += G.nsmall > 0;
    // the large fronts (the group's factors are all of one precision)
    {
      const int prec = fp32 ? 1 : 0;
      std::vector<std::pair<int, int>> big;
      for (auto& p : G.big)
        if ((mask >> p.first) & 1ull) big.push_back(p);
      if (big.empty()) continue;
      BigCombo& c = combo(big);
      const int epoch = next_epoch();
      if (forward) {
        launch_big_fsolve_multi(st_, c.bs, c.bfront, c.nf, c.ofs, c.nfs,
                                d_stick_.as<unsigned long long>(), F, T_.childlist, I, kv, epoch, sbase_,
                                prec == 1);
        sbase_ += c.nfs;
      } else {
        launch_big_bsolve_multi(st_, c.bs, c.bfront, c.nf, c.obs, c.nbs,
                                d_stick_.as<unsigned long long>(), F, I, kv, epoch, sbase_, prec == 1, c.maxtr,
                                c.maxch);
        sbase_ += c.nbs;
      }
      st_acc_.launches += 2;
    }
  }
}

void Gang::issue_group(const std::vector<int>& g) {
  const GOp& op = *op_[g[0]];
  uint64_t mask = 0;
  for (int k : g) mask |= 1ull << k;
  const InstDev* I = T_.inst;
  const ConeDesc* C = T_.cones;
  if (op.kind == GK_FACTOR) {
    combined_factor(g, op.a[0] != 0);
    return;
  }
  if (op.kind == GK_LSOLVE || op.kind == GK_USOLVE) {
    combined_solve(g, op.a[0], op.kind == GK_LSOLVE, op.a[1] != 0);
    return;
  }
  set_mask(mask);
  st_acc_.launches += 1;
  switch (op.kind) {
    case GK_SET2:
    case GK_SMU: {
      GangVals v{};
      for (int k : g) {
        v.a[k] = op_[k]->v[0];
        v.b[k] = op_[k]->v[1];
        v.c[k] = op_[k]->v[2];
      }
      if (op.kind == GK_SET2) launch_gang_set2(st_, I, kmax_, op.a[0], op.a[1], v);
      else launch_gang_smu(st_, I, kmax_, v);
      break;
    }
    case GK_CGSC:
      launch_cg_scalars(st_, I, kmax_, op.a[0]);
      break;
    case GK_LIN:
      launch_lin(st_, I, T_.tn, ntn_, T_.tm, ntm_, op.N, op.M);
      break;
    case GK_AXPBY:
      if (op.a[0] == 0) launch_axpby(st_, I, T_.tn, ntn_, 0, op.a[1], op.a[2], op.a[3], op.c[0], op.c[1]);
      else launch_axpby(st_, I, T_.tm, ntm_, 1, op.a[1], op.a[2], op.a[3], op.c[0], op.c[1]);
      break;
    case GK_AXPBYPCZ:
      if (op.a[0] == 0)
        launch_axpbypcz(st_, I, T_.tn, ntn_, 0, op.a[1], op.a[2], op.a[3], op.a[4], op.c[0], op.c[1],
                        op.c[2]);
      else
        launch_axpbypcz(st_, I, T_.tm, ntm_, 1, op.a[1], op.a[2], op.a[3], op.a[4], op.c[0], op.c[1],
                        op.c[2]);
      break;
    case GK_SPMV:
      launch_spmv(st_, I, T_.sp[op.a[0]], nsp_[op.a[0]], op.a[0], op.a[1], op.a[2], op.c[0], op.c[1]);
      break;
    case GK_SPMVNM: {
      const int ntm = op.a[0];
      launch_spmv_nm(st_, I, ntm, T_.sp[ntm], op.a[1] ? nsp_[ntm] : 0, T_.sp[0],
                     op.a[2] ? nsp_[0] : 0, op.sn, op.sm);
      break;
    }
    case GK_DOTS:
      launch_dots(st_, I, 0, op.db, kmax_);
      break;
    case GK_SCALE:
      launch_cone_scale(st_, I, C, ncone_, T_.cone_ofs, kmax_, T_.coneout, T_.psd,
                        npsd_, BigPsd{}, T_.soc, nsoc_);
      launch_cone_hnorm2(st_, I, C, ncone_, T_.cone_ofs, kmax_, T_.coneout + 2 * ncone_, S_TMP1,
                         T_.dense, ndense_, dense_nmax_, T_.dpart, BigPsd{}, nullptr);
      st_acc_.launches += 3;
      break;
    case GK_MATVEC:
      launch_cone_matvec(st_, I, C, ncone_, op.a[0], op.a[1], op.c[0]);
      launch_cone_matvec_dense(st_, I, C, T_.dense, ndense_, dense_nmax_, op.a[0], op.a[1], op.c[0]);
      st_acc_.launches += 1;
      break;
    case GK_CORR0:
      launch_cone_corr0(st_, I, C, ncone_, op.a[0]);
      launch_psd_corr0(st_, I, C, T_.psd, npsd_, op.a[0]);
      st_acc_.launches += 1;
      break;
    case GK_CORR:
      launch_cone_corr(st_, I, C, ncone_, op.a[0], op.a[1], op.a[2]);
      launch_psd_corr(st_, I, C, T_.psd, npsd_, op.a[0], op.a[1], op.a[2]);
      st_acc_.launches += 1;
      break;
    case GK_MAXSTEP:
      launch_cone_maxstep(st_, I, C, ncone_, T_.cone_ofs, kmax_, op.a[0], op.a[1], op.c[0].c,
                          T_.coneout, T_.psd, npsd_, BigPsd{});
      st_acc_.launches += 2;
      break;
    case GK_ZEROFREE:
      launch_zerofree(st_, I, C, ncone_, op.a[0]);
      break;
    default:
      throw std::logic_error("gang: op without a combined form");
  }
}

}  // namespace mb
