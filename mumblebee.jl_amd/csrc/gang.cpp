#include "gang.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <queue>
#include <stdexcept>
#include <thread>

#include <immintrin.h>

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
  DevBuf bs, bfront, odag, ofs, obs;
  int nf = 0, ndag = 0, nfs = 0, nbs = 0, max2tr = 0, maxtr = 0, maxch = 0;
  double flops = 0.0;
};

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

void spin_until(const std::atomic<bool>& f) {
  for (int i = 0; !f.load(std::memory_order_acquire); ++i) {
    if (i < 4096) _mm_pause();
    else std::this_thread::yield();
  }
}

}  // namespace

Gang::Gang(int kmax, int dag_grid, bool profile)
    : kmax_(std::max(1, std::min(kmax, GANG_MAX))), dag_grid_(dag_grid), profile_(profile) {
  HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&ev0_));
  HIPCHK(hipEventCreate(&ev1_));
  HIPCHK(hipEventCreateWithFlags(&evleave_, hipEventDisableTiming));
  if (dag_grid_ <= 0) {
    int dev = 0, ncu = 256;
    HIPCHK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
    dag_grid_ = ncu;
  }
  mem_.assign(kmax_, nullptr);
  op_.assign(kmax_, nullptr);
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
}

Gang::~Gang() {
  (void)hipStreamSynchronize(st_);
  graveyard_.clear();
  combos_.clear();
  if (h_gather_) (void)hipHostFree(h_gather_);
  (void)hipEventDestroy(ev0_);
  (void)hipEventDestroy(ev1_);
  (void)hipEventDestroy(evleave_);
  (void)hipStreamDestroy(st_);
}

int Gang::members() {
  std::lock_guard<std::mutex> lk(mu_);
  return nrun_;
}

void Gang::stats(GangStats& out) {
  std::lock_guard<std::mutex> lk(mu_);
  out = st_acc_;
}

// ------------------------------------------------------------------------------------------
// membership and ticks
// ------------------------------------------------------------------------------------------

void Gang::join(Solver* s) {
  if (s->gang_) throw std::logic_error("solver is already in a gang");
  HIPCHK(hipStreamSynchronize(s->st_));  // its setup and starting point are on the device
  std::unique_lock<std::mutex> lk(mu_);
  joining_.push_back(s);
  for (;;) {
    if (s->gang_ == this) return;
    if (!leading_ && nrun_ == 0) {
      // no tick will come: admit the waiting solvers here
      leading_ = true;
      lk.unlock();
      std::string err;
      try {
        admit();
      } catch (const std::exception& e) {
        err = e.what();
      }
      lk.lock();
      leading_ = false;
      cv_.notify_all();
      if (!err.empty()) throw std::runtime_error(err);
      continue;
    }
    cv_.wait_for(lk, std::chrono::microseconds(200));
  }
}

// the waiting solvers take free slots (the leader, between ticks); the tables are rebuilt
void Gang::admit() {
  std::vector<Solver*> take;
  {
    std::lock_guard<std::mutex> lk(mu_);
    std::vector<Solver*> rest;
    for (Solver* s : joining_) {
      int k = -1;
      for (int j = 0; j < kmax_; ++j) {
        bool used = mem_[j] != nullptr;
        for (Solver* t : take)
          if (t->gslot_ == j) used = true;
        if (!used) {
          k = j;
          break;
        }
      }
      if (k < 0) {
        rest.push_back(s);
        continue;
      }
      s->gslot_ = k;
      take.push_back(s);
    }
    joining_ = rest;
  }
  if (take.empty()) return;
  HIPCHK(hipStreamSynchronize(st_));
  for (Solver* s : take) {
    mem_[s->gslot_] = s;
    s->own_st_ = s->st_;
    s->st_ = st_;
    s->gang_prof_ = s->prof_mask_;
    s->prof_mask_ = 0;  // per-kernel events are the gang's (its combined launches serve all members)
  }
  dirty_ = true;
  rebuild();
  std::lock_guard<std::mutex> lk(mu_);
  for (Solver* s : take) {
    s->gang_ = this;
    ++nrun_;
    st_acc_.joins += 1;
  }
}

void Gang::leave(Solver* s) {
  if (s->gang_ != this) return;
  GOp op;
  op.kind = GK_LEAVE;
  submit(s, op);
}

void Gang::submit(Solver* s, GOp& op) {
  const int k = s->gslot_;
  op.done.store(false, std::memory_order_relaxed);
  bool lead = false;
  {
    std::lock_guard<std::mutex> lk(mu_);
    op_[k] = &op;
    ++narr_;
    if (!leading_ && narr_ == nrun_) lead = leading_ = true;
  }
  while (lead) {
    try {
      tick();
    } catch (const std::exception& e) {
      // a failure outside any one op (table rebuild, read-back): every pending op fails, and the
      // members that were leaving are gone
      std::lock_guard<std::mutex> lk(mu_);
      for (int j = 0; j < kmax_; ++j)
        if (op_[j]) {
          GOp* o = op_[j];
          op_[j] = nullptr;
          --narr_;
          if (o->kind == GK_LEAVE) {
            detach(j);
            mem_[j] = nullptr;
            --nrun_;
          }
          o->err = e.what();
          o->done.store(true, std::memory_order_release);
        }
    }
    // every running member waits on a handed-in op: this thread leads the next tick too
    std::lock_guard<std::mutex> lk(mu_);
    lead = nrun_ > 0 && narr_ == nrun_;
    leading_ = lead;
    cv_.notify_all();
  }
  spin_until(op.done);
  if (!op.err.empty()) throw std::runtime_error(op.err);
}

void Gang::tick() {
  {
    bool wait;
    {
      std::lock_guard<std::mutex> lk(mu_);
      wait = !joining_.empty();
    }
    if (wait) admit();
  }
  std::vector<int> cheap, parked;
  bool all;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (int k = 0; k < kmax_; ++k)
      if (op_[k]) (op_[k]->park ? parked : cheap).push_back(k);
    all = narr_ == nrun_;
  }
  std::vector<int> issue;
  if (!cheap.empty()) issue = cheap;
  else if (all) issue = parked;
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
  for (int k : lv) {
    Solver* s = mem_[k];
    set_mask(dev_mask_ & ~(1ull << k));
    hinst_[k] = dead_;
    launch_gang_setinst(st_, d_inst_.as<InstDev>(), k, dead_);
    launch_set_scalar(st_, s->d_sc_.as<double>(), S_ACTIVE, 1.0);
    HIPCHK(hipEventRecord(evleave_, st_));
    HIPCHK(hipStreamWaitEvent(s->own_st_, evleave_, 0));  // its own stream starts after its last op here
    detach(k);
    // cached launch data may name this slot's fronts: kept until the stream has passed them
    for (auto& kv : combos_) graveyard_.push_back(std::move(kv.second));
    combos_.clear();
  }
  std::lock_guard<std::mutex> lk(mu_);
  for (int k : issue) {
    GOp* o = op_[k];
    op_[k] = nullptr;
    --narr_;
    if (o->kind == GK_LEAVE) {
      mem_[k] = nullptr;
      --nrun_;
    }
    o->done.store(true, std::memory_order_release);
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
  launch_gang_active(st_, d_inst_.as<InstDev>(), kmax_, mask & occ, occ);
  dev_mask_ = mask & occ;
  st_acc_.launches += 1;
}

void Gang::issue_solo(int k) {
  Solver* s = mem_[k];
  GOp& op = *op_[k];
  if (!((dev_mask_ >> k) & 1ull)) set_mask(dev_mask_ | (1ull << k));
  s->issuing_ = true;
  try {
    if (op.solo) op.solo();
  } catch (const std::exception& e) {
    op.err = e.what();
  }
  s->issuing_ = false;
}

void Gang::readbacks(const std::vector<int>& slots) {
  uint64_t mask = 0;
  for (int k : slots) mask |= 1ull << k;
  launch_gang_gather(st_, d_inst_.as<InstDev>(), kmax_, mask, d_gather_.as<double>());
  HIPCHK(hipMemcpyAsync(h_gather_, d_gather_.p, sizeof(double) * NSCALAR * kmax_, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  graveyard_.clear();  // the stream has passed every launch that could read them
  for (int k : slots) std::memcpy(mem_[k]->hsc_, h_gather_ + (size_t)k * NSCALAR, sizeof(double) * NSCALAR);
  st_acc_.readbacks += 1;
  st_acc_.launches += 1;
  if (profile_ && pending_timing_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev0_, ev1_) == hipSuccess) {
      st_acc_.factor_ms += ms;
      st_acc_.dag_ms += ms;
    }
    pending_timing_ = false;
  }
}

// ------------------------------------------------------------------------------------------
// combined tables
// ------------------------------------------------------------------------------------------

void Gang::rebuild() {
  HIPCHK(hipStreamSynchronize(st_));
  graveyard_.clear();
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
  auto up_tiles = [&](DevBuf& d, const std::vector<Tile>& v) { d.upload(v.empty() ? std::vector<Tile>{Tile{0, 0}} : v, st_); };
  auto up_int = [&](DevBuf& d, const std::vector<int>& v) { d.upload(v.empty() ? std::vector<int>{0} : v, st_); };
  d_inst_.upload(hinst_, st_);
  up_tiles(d_tn_, tn);
  up_tiles(d_tm_, tm);
  for (int w = 0; w < 3; ++w) up_tiles(d_sp_[w], sp[w]);
  d_cones_.upload(cones.empty() ? std::vector<ConeDesc>(1) : cones, st_);
  up_int(d_cone_ofs_, cone_ofs);
  d_coneout_.alloc(sizeof(double) * 3 * std::max(ncone_, 1));
  up_int(d_psd_, psd);
  up_int(d_dense_, dense);
  up_int(d_soc_, soc);
  d_dpart_.alloc(sizeof(double) * DENSE_SPLIT * std::max(ndense_, 1));
  up_int(d_cone_front_, cone_front);
  d_fr_.upload(fr.empty() ? std::vector<FrontDesc>(1) : fr, st_);
  up_int(d_childlist_, childlist);
  up_int(d_items_, items);
  d_steps_.upload(steps.empty() ? std::vector<StepDesc>(1) : steps, st_);
  up_int(d_ea_, ea);
  up_int(d_ea0_, ea0);
  d_flags_.alloc(sizeof(int) * std::max<size_t>(steps.size(), 1));
  HIPCHK(hipMemsetAsync(d_flags_.p, 0, d_flags_.bytes, st_));
  // every member takes part until a launch says otherwise
  uint64_t occ = 0;
  for (int k = 0; k < K; ++k)
    if (mem_[k]) occ |= 1ull << k;
  launch_gang_active(st_, d_inst_.as<InstDev>(), K, occ, occ);
  dev_mask_ = occ;
  HIPCHK(hipStreamSynchronize(st_));
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
  const std::vector<int2> odag = interleave(cdag), ofs = interleave(cfs), obs = interleave(cbs);
  c->nf = (int)bs.size();
  c->ndag = (int)odag.size();
  c->nfs = (int)ofs.size();
  c->nbs = (int)obs.size();
  c->bs.upload(bs, st_);
  c->bfront.upload(bfront, st_);
  c->odag.upload(odag, st_);
  c->ofs.upload(ofs, st_);
  c->obs.upload(obs, st_);
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
  const InstDev* I = d_inst_.as<InstDev>();
  const FrontDesc* F = d_fr_.as<FrontDesc>();
  // assembly: each member's fill and union map, then the cone blocks of all in one launch
  for (int k : slots) {
    Solver* s = mem_[k];
    launch_fill(st_, s->d_F_.as<double>(), (size_t)s->fsize_, 0.0);
    launch_assemble_union(st_, s->d_inst_.as<InstDev>(), 0, s->d_upos_.as<int64_t>(), s->d_ul_.as<double>(),
                          s->d_uq_.as<double>(), s->nunion_, 1);
  }
  launch_assemble_cones(st_, I, d_cones_.as<ConeDesc>(), d_cone_front_.as<int>(), F, ncone_, cone_nmax_);
  st_acc_.launches += 2.0 * slots.size() + 1;
  const int epoch = next_epoch();
  if (profile_) HIPCHK(hipEventRecord(ev0_, st_));
  double flops = 0.0;
  int nf = 0;
  for (size_t l = 0; l < lv_.size(); ++l) {
    const GLevel& G = lv_[l];
    for (size_t r = 0; r < G.ea_off.size(); ++r)
      launch_extend_add(st_, F, d_ea_.as<int>() + G.ea_off[r], d_ea0_.as<int>() + G.ea_off[r], G.ea_n[r],
                        G.ea_tiles[r], I);
    for (size_t q = 0; q < G.pan_off.size(); ++q) {
      launch_panel(st_, F, d_steps_.as<StepDesc>() + G.pan_off[q], G.pan_n[q], G.pan_blk[q], I,
                   d_flags_.as<int>() + G.pan_off[q], epoch, nullptr);
      if (G.upd_blk[q] > 0) launch_update(st_, F, d_steps_.as<StepDesc>() + G.upd_off[q], G.upd_n[q], G.upd_blk[q], I);
    }
    std::vector<std::pair<int, int>> big;
    for (auto& p : G.big)
      if ((mask >> p.first) & 1ull) big.push_back(p);
    if (!big.empty()) {
      BigCombo& c = combo(big);
      launch_dag_factor_multi(st_, c.bs.as<BigFront>(), c.nf, c.odag.as<int2>(), c.ndag, d_tick_.as<int>(), I,
                              dag_grid_, epoch, fp32, c.max2tr);
      flops += c.flops;
      nf += c.nf;
      st_acc_.dag_launches += 1;
      st_acc_.launches += 2;
    }
    st_acc_.launches += (double)(G.ea_off.size() + 2 * G.pan_off.size());
  }
  if (profile_) {
    HIPCHK(hipEventRecord(ev1_, st_));
    pending_timing_ = true;
  }
  st_acc_.factor_launches += 1;
  st_acc_.factor_flops += flops;
  st_acc_.dag_flops += flops;
  st_acc_.fronts_per_factor += nf;
}

void Gang::combined_solve(const std::vector<int>& slots, int kv, bool forward) {
  uint64_t mask = 0;
  for (int k : slots) mask |= 1ull << k;
  set_mask(mask);
  const InstDev* I = d_inst_.as<InstDev>();
  const FrontDesc* F = d_fr_.as<FrontDesc>();
  const int H = (int)lv_.size();
  for (int i = 0; i < H; ++i) {
    const int l = forward ? i : H - 1 - i;
    const GLevel& G = lv_[l];
    if (forward) launch_fsolve(st_, F, d_items_.as<int>() + G.items_off, G.nsmall, d_childlist_.as<int>(), I, kv);
    else launch_bsolve(st_, F, d_items_.as<int>() + G.items_off, G.nsmall, I, kv);
    st_acc_.launches += G.nsmall > 0;
    // the large fronts, by the precision of their member's current factor
    for (int prec = 0; prec < 2; ++prec) {
      std::vector<std::pair<int, int>> big;
      for (auto& p : G.big)
        if (((mask >> p.first) & 1ull) && (mem_[p.first]->fac_fp32_ ? 1 : 0) == prec) big.push_back(p);
      if (big.empty()) continue;
      BigCombo& c = combo(big);
      const int epoch = next_epoch();
      if (forward) {
        launch_big_fsolve_multi(st_, c.bs.as<BigFront>(), c.bfront.as<int>(), c.nf, c.ofs.as<int2>(), c.nfs,
                                d_stick_.as<unsigned long long>(), F, d_childlist_.as<int>(), I, kv, epoch, sbase_,
                                prec == 1);
        sbase_ += c.nfs;
      } else {
        launch_big_bsolve_multi(st_, c.bs.as<BigFront>(), c.bfront.as<int>(), c.nf, c.obs.as<int2>(), c.nbs,
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
  const InstDev* I = d_inst_.as<InstDev>();
  const ConeDesc* C = d_cones_.as<ConeDesc>();
  if (op.kind == GK_FACTOR) {
    combined_factor(g, op.a[0] != 0);
    return;
  }
  if (op.kind == GK_LSOLVE || op.kind == GK_USOLVE) {
    combined_solve(g, op.a[0], op.kind == GK_LSOLVE);
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
      launch_lin(st_, I, d_tn_.as<Tile>(), ntn_, d_tm_.as<Tile>(), ntm_, op.N, op.M);
      break;
    case GK_AXPBY:
      if (op.a[0] == 0) launch_axpby(st_, I, d_tn_.as<Tile>(), ntn_, 0, op.a[1], op.a[2], op.a[3], op.c[0], op.c[1]);
      else launch_axpby(st_, I, d_tm_.as<Tile>(), ntm_, 1, op.a[1], op.a[2], op.a[3], op.c[0], op.c[1]);
      break;
    case GK_AXPBYPCZ:
      if (op.a[0] == 0)
        launch_axpbypcz(st_, I, d_tn_.as<Tile>(), ntn_, 0, op.a[1], op.a[2], op.a[3], op.a[4], op.c[0], op.c[1],
                        op.c[2]);
      else
        launch_axpbypcz(st_, I, d_tm_.as<Tile>(), ntm_, 1, op.a[1], op.a[2], op.a[3], op.a[4], op.c[0], op.c[1],
                        op.c[2]);
      break;
    case GK_SPMV:
      launch_spmv(st_, I, d_sp_[op.a[0]].as<Tile>(), nsp_[op.a[0]], op.a[0], op.a[1], op.a[2], op.c[0], op.c[1]);
      break;
    case GK_SPMVNM: {
      const int ntm = op.a[0];
      launch_spmv_nm(st_, I, ntm, d_sp_[ntm].as<Tile>(), op.a[1] ? nsp_[ntm] : 0, d_sp_[0].as<Tile>(),
                     op.a[2] ? nsp_[0] : 0, op.sn, op.sm);
      break;
    }
    case GK_DOTS:
      launch_dots(st_, I, 0, op.db, kmax_);
      break;
    case GK_SCALE:
      launch_cone_scale(st_, I, C, ncone_, d_cone_ofs_.as<int>(), kmax_, d_coneout_.as<double>(), d_psd_.as<int>(),
                        npsd_, BigPsd{}, d_soc_.as<int>(), nsoc_);
      launch_cone_hnorm2(st_, I, C, ncone_, d_cone_ofs_.as<int>(), kmax_, d_coneout_.as<double>() + 2 * ncone_, S_TMP1,
                         d_dense_.as<int>(), ndense_, dense_nmax_, d_dpart_.as<double>(), BigPsd{}, nullptr);
      st_acc_.launches += 3;
      break;
    case GK_MATVEC:
      launch_cone_matvec(st_, I, C, ncone_, op.a[0], op.a[1], op.c[0]);
      launch_cone_matvec_dense(st_, I, C, d_dense_.as<int>(), ndense_, dense_nmax_, op.a[0], op.a[1], op.c[0]);
      st_acc_.launches += 1;
      break;
    case GK_CORR0:
      launch_cone_corr0(st_, I, C, ncone_, op.a[0]);
      launch_psd_corr0(st_, I, C, d_psd_.as<int>(), npsd_, op.a[0]);
      st_acc_.launches += 1;
      break;
    case GK_CORR:
      launch_cone_corr(st_, I, C, ncone_, op.a[0], op.a[1], op.a[2]);
      launch_psd_corr(st_, I, C, d_psd_.as<int>(), npsd_, op.a[0], op.a[1], op.a[2]);
      st_acc_.launches += 1;
      break;
    case GK_MAXSTEP:
      launch_cone_maxstep(st_, I, C, ncone_, d_cone_ofs_.as<int>(), kmax_, op.a[0], op.a[1], op.c[0].c,
                          d_coneout_.as<double>(), d_psd_.as<int>(), npsd_, BigPsd{});
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
