// IPM driver over device-resident state (see solver.h).
#include "solver.h"
#include "gang.h"
#include "tdc.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>

namespace mb {

int next_epoch() {
  static std::atomic<int> e{0};
  return ++e;
}

static const double EPS = std::numeric_limits<double>::epsilon();
static const double NaN = std::numeric_limits<double>::quiet_NaN();
// mixed precision: a step below this with an FP32 factor sends the rest of the solve to FP64
static const double MIXED_MIN_STEP = 0.05;
static const double FORCING_FRAC = 0.1;  // solver/solver.jl:3

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +           \
                                                   hipGetErrorString(e_) + " at " #x);     \
  } while (0)

// device bytes allocated through DevBuf by the solver being built on this thread
static thread_local size_t* g_dev_counter = nullptr;

// Uploads of the solver being set up on this thread. Each is copied into a pinned host image and
// recorded; so is each buffer the setup zeroes. flush() moves the image to the device in one copy
// per pinned chunk and then places every piece (copies and zero fills) with one kernel. A setup
// used to issue some 50 copies and 20 fills, each a small blit kernel on the solver's stream; while
// other solvers' persistent kernels held every CU, each of them waited for a free one in turn.
namespace {
void* pinned_alloc(size_t c);
void pinned_free(void* p, size_t c);
size_t size_class(size_t b);
void* dev_alloc(size_t b);
void dev_free(void* p, size_t b);
struct Stager {
  static constexpr size_t CHUNK = size_t(4) << 20;
  hipStream_t st = nullptr;
  std::vector<std::pair<char*, size_t>> chunks;  // pinned block, its size class
  std::vector<size_t> base;                      // each chunk's offset in the image
  size_t used = 0;                               // bytes used in the last chunk
  std::vector<StageSeg> segs;
  std::vector<std::pair<void*, size_t>> images;  // device images of earlier flushes (until finish)
  void add(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const size_t span = (bytes + 255) / 256 * 256;  // pieces start 256-byte aligned in the image
    if (chunks.empty() || used + span > chunks.back().second) {
      const size_t c = size_class(std::max(CHUNK, span));
      base.push_back(chunks.empty() ? 0 : base.back() + used);
      chunks.emplace_back(static_cast<char*>(pinned_alloc(c)), c);
      used = 0;
    }
    std::memcpy(chunks.back().first + used, src, bytes);
    segs.push_back(StageSeg{static_cast<char*>(dst), (long long)(base.back() + used), (long long)bytes, 0});
    used += span;
  }
  void zero(void* dst, size_t bytes) {
    if (bytes) segs.push_back(StageSeg{static_cast<char*>(dst), -1, (long long)bytes, 0});
  }
  // the pieces recorded so far, in stream order before anything queued after this call: the zero
  // fills in one launch, then the copies in another (a setup zeroes a buffer before it writes
  // parts of it; the pieces of one launch are placed in no particular order)
  void flush() {
    if (segs.empty()) return;
    std::vector<StageSeg> tab[2];  // copies, zero fills
    long long nblk[2] = {0, 0};
    for (const StageSeg& g : segs) {
      const int z = g.src < 0 ? 1 : 0;
      tab[z].push_back(g);
      tab[z].back().blk0 = nblk[z];
      nblk[z] += (g.bytes + STAGE_BLOCK - 1) / STAGE_BLOCK;
    }
    // the piece tables go up with the image, in a chunk of their own
    const size_t n0 = tab[0].size(), n1 = tab[1].size();
    const size_t tb = (n0 + n1) * sizeof(StageSeg);
    const size_t img = chunks.empty() ? 0 : base.back() + used;
    const size_t tc = size_class(std::max<size_t>(tb, 256));
    char* tp = static_cast<char*>(pinned_alloc(tc));
    if (n0) std::memcpy(tp, tab[0].data(), n0 * sizeof(StageSeg));
    if (n1) std::memcpy(tp + n0 * sizeof(StageSeg), tab[1].data(), n1 * sizeof(StageSeg));
    const size_t total = img + tb;
    char* d = static_cast<char*>(dev_alloc(total));
    images.emplace_back(d, total);
    for (size_t k = 0; k < chunks.size(); ++k) {
      const size_t n = k + 1 < chunks.size() ? base[k + 1] - base[k] : used;
      if (n) HIPCHK(hipMemcpyAsync(d + base[k], chunks[k].first, n, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d + img, tp, tb, hipMemcpyHostToDevice, st));
    const StageSeg* dt = reinterpret_cast<const StageSeg*>(d + img);
    launch_stage_scatter(st, d, dt + n0, (int)n1, nblk[1]);
    launch_stage_scatter(st, d, dt, (int)n0, nblk[0]);
    // the pinned blocks are read by the copies: kept until finish() has synchronised
    chunks.emplace_back(tp, tc);
    retired.insert(retired.end(), chunks.begin(), chunks.end());
    chunks.clear();
    base.clear();
    used = 0;
    segs.clear();
  }
  std::vector<std::pair<char*, size_t>> retired;
  // flush, wait for the device, then the pinned and device blocks go back to the caches
  void finish() {
    hipError_t e = hipSuccess;
    try {
      flush();
    } catch (...) {
      e = hipErrorUnknown;
    }
    if (!retired.empty() || !images.empty()) {
      const hipError_t w = hipStreamSynchronize(st);
      if (e == hipSuccess) e = w;
    }
    release();
    HIPCHK(e);
  }
  void release() {
    for (auto& c : retired) pinned_free(c.first, c.second);
    for (auto& c : chunks) pinned_free(c.first, c.second);
    for (auto& m : images) dev_free(m.first, m.second);
    retired.clear();
    chunks.clear();
    images.clear();
    segs.clear();
    base.clear();
    used = 0;
  }
  ~Stager() {
    if (!retired.empty() || !images.empty() || !chunks.empty()) {
      (void)hipStreamSynchronize(st);
      release();
    }
  }
};
thread_local Stager* g_stager = nullptr;
}  // namespace

// Device and pinned-host memory cache. hipFree (and hipHostFree) waits for every kernel on the
// device, whichever stream it runs on: a hipFree issued while a 50 ms kernel ran on another
// stream returned when that kernel ended. A solver that freed its buffers while other solvers
// ran would stall all of them, so freed blocks are kept in size classes and handed to later
// allocations of the same class. Classes: powers of two up to 1 MiB, then eight per octave (at
// most 12.5 % over the request). A block returns to the cache only when its solver's stream is
// idle (the solver synchronizes its stream before its buffers are released). Out of memory, the
// cache is emptied into the runtime and the allocation retried.
namespace {
struct MemCache {
  std::mutex mu;
  std::map<size_t, std::vector<void*>> dev;  // class bytes → free blocks
  std::vector<double*> scal;                 // pinned scalar blocks (NSCALAR doubles)
  size_t dev_cached = 0;
};
MemCache& memcache() {
  static MemCache* c = new MemCache();  // outlives every solver, including static ones
  return *c;
}
size_t size_class(size_t b) {
  size_t c = 256;
  while (c < b) c <<= 1;
  if (c <= (size_t(1) << 20)) return c;
  const size_t step = c / 16;  // c/2 < b ≤ c: eight classes over the octave
  return (b + step - 1) / step * step;
}
void* dev_alloc(size_t b) {
  const size_t c = size_class(b);
  MemCache& mc = memcache();
  {
    std::lock_guard<std::mutex> lk(mc.mu);
    auto it = mc.dev.find(c);
    if (it != mc.dev.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      mc.dev_cached -= c;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    mb::release_cached_memory();
    e = hipMalloc(&p, c);
  }
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at hipMalloc");
  return p;
}
void dev_free(void* p, size_t b) {
  MemCache& mc = memcache();
  std::lock_guard<std::mutex> lk(mc.mu);
  mc.dev[size_class(b)].push_back(p);
  mc.dev_cached += size_class(b);
}
double* scal_alloc() {
  MemCache& mc = memcache();
  {
    std::lock_guard<std::mutex> lk(mc.mu);
    if (!mc.scal.empty()) {
      double* p = mc.scal.back();
      mc.scal.pop_back();
      return p;
    }
  }
  double* p = nullptr;
  HIPCHK(hipHostMalloc((void**)&p, NSCALAR * sizeof(double), hipHostMallocDefault));
  return p;
}
void scal_free(double* p) {
  MemCache& mc = memcache();
  std::lock_guard<std::mutex> lk(mc.mu);
  mc.scal.push_back(p);
}
}  // namespace

size_t release_cached_memory() {
  MemCache& mc = memcache();
  std::lock_guard<std::mutex> lk(mc.mu);
  size_t n = mc.dev_cached;
  for (auto& kv : mc.dev)
    for (void* p : kv.second) (void)hipFree(p);
  mc.dev.clear();
  mc.dev_cached = 0;
  return n;
}

size_t cached_memory() {
  MemCache& mc = memcache();
  std::lock_guard<std::mutex> lk(mc.mu);
  return mc.dev_cached;
}

DevBuf::~DevBuf() {
  if (p) dev_free(p, bytes);
}
void DevBuf::alloc(size_t b) {
  if (p) {
    if (g_stager) {  // a staged piece may still name the old block: place them all first
      g_stager->flush();
      HIPCHK(hipStreamSynchronize(g_stager->st));
    }
    dev_free(p, bytes);
    p = nullptr;
  }
  bytes = b;
  if (b) p = dev_alloc(b);
  if (g_dev_counter && b) *g_dev_counter += size_class(b);  // what the allocator hands out
}
// Every copy and fill of a solver's buffers goes on the solver's own stream. The solver's stream
// is non-blocking, so work on the null stream (hipMemcpy, hipMemset: both may return before the
// device has finished them) is not ordered before the solver's kernels.
template <class T>
void DevBuf::upload(const std::vector<T>& v, hipStream_t st) {
  if (g_stager) {  // during device setup: staged, placed with the others (alloc flushes first)
    alloc(std::max<size_t>(v.size() * sizeof(T), sizeof(T)));
    g_stager->add(p, v.data(), v.size() * sizeof(T));
    return;
  }
  HIPCHK(hipStreamSynchronize(st));  // the old buffer is no longer in use
  alloc(std::max<size_t>(v.size() * sizeof(T), sizeof(T)));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // synchronous: callers pass temporaries
}

// zeros in the whole buffer, staged during device setup
void DevBuf::clear(hipStream_t st) {
  if (!p) return;
  if (g_stager) g_stager->zero(p, bytes);
  else HIPCHK(hipMemsetAsync(p, 0, bytes, st));
}

// the element types other translation units upload (the gang's combined tables)
template void DevBuf::upload<int>(const std::vector<int>&, hipStream_t);
template void DevBuf::upload<Tile>(const std::vector<Tile>&, hipStream_t);
template void DevBuf::upload<InstDev>(const std::vector<InstDev>&, hipStream_t);
template void DevBuf::upload<ConeDesc>(const std::vector<ConeDesc>&, hipStream_t);
template void DevBuf::upload<FrontDesc>(const std::vector<FrontDesc>&, hipStream_t);
template void DevBuf::upload<StepDesc>(const std::vector<StepDesc>&, hipStream_t);
template void DevBuf::upload<BigFront>(const std::vector<BigFront>&, hipStream_t);
template void DevBuf::upload<int2>(const std::vector<int2>&, hipStream_t);

// pinned host blocks for the setup uploads, kept in size classes like the device blocks
namespace {
std::mutex g_pin_mu;
std::map<size_t, std::vector<void*>> g_pinned;
void* pinned_alloc(size_t c) {
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.find(c);
    if (it != g_pinned.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      return p;
    }
  }
  void* p = nullptr;
  HIPCHK(hipHostMalloc(&p, c, hipHostMallocDefault));
  return p;
}
void pinned_free(void* p, size_t c) {
  std::lock_guard<std::mutex> lk(g_pin_mu);
  g_pinned[c].push_back(p);
}
}  // namespace

// ------------------------------------------------------------------------------------------
// Ruiz equilibration (scaling.jl:52-184)
// ------------------------------------------------------------------------------------------
void equilibrate(BSM& B, BSM& Q, std::vector<double>& f, std::vector<double>& g, int itmax, double tol,
                 std::vector<double>& pscl, std::vector<double>& yscl) {
  const i64 nrow = B.nrow(), nvtx = B.nvtx;
  pscl.assign(B.ncol(), 1.0);
  yscl.assign(nrow, 1.0);
  std::vector<double> vscl(nvtx, 1.0), vnrm(nvtx), rnrm(nrow), vswp(nvtx, 1.0), rswp(nrow, 1.0);
  for (int it = 0; it < itmax; ++it) {
    std::fill(rnrm.begin(), rnrm.end(), 0.0);
    for (i64 v = 0; v < nvtx; ++v) {
      double nv = 0.0;
      // the blocks of a vertex's arcs lie one after another in val
      for (i64 b = Q.xblk[Q.xsrc[v]], be = Q.xblk[Q.xsrc[v + 1]]; b < be; ++b) nv = std::max(nv, std::fabs(Q.val[b]));
      i64 nc = B.ncols(v);
      for (i64 e = B.xsrc[v]; e < B.xsrc[v + 1]; ++e) {
        i64 u = B.tgt[e], nr = B.nrows(u), r0 = B.xrow[u];
        for (i64 jl = 0; jl < nc; ++jl)
          for (i64 il = 0; il < nr; ++il) {
            double a = std::fabs(B.val[B.xblk[e] + jl * nr + il]);
            nv = std::max(nv, a);
            rnrm[r0 + il] = std::max(rnrm[r0 + il], a);
          }
      }
      vnrm[v] = nv;
    }
    double d = 0.0;
    for (i64 v = 0; v < nvtx; ++v) {
      double n = std::sqrt(vnrm[v]);
      d = std::max(d, std::fabs(1.0 - n));
      if (n != 0.0) {
        vswp[v] = 1.0 / n;
        vscl[v] *= vswp[v];
      }
    }
    for (i64 i = 0; i < nrow; ++i) {
      double n = std::sqrt(rnrm[i]);
      d = std::max(d, std::fabs(1.0 - n));
      if (n != 0.0) {
        rswp[i] = 1.0 / n;
        yscl[i] *= rswp[i];
      }
    }
    for (i64 v = 0; v < nvtx; ++v) {
      double sv = vswp[v];
      i64 nc = B.ncols(v);
      for (i64 e = B.xsrc[v]; e < B.xsrc[v + 1]; ++e) {
        i64 u = B.tgt[e], nr = B.nrows(u), r0 = B.xrow[u];
        for (i64 jl = 0; jl < nc; ++jl)
          for (i64 il = 0; il < nr; ++il) B.val[B.xblk[e] + jl * nr + il] *= rswp[r0 + il] * sv;
      }
      for (i64 e = Q.xsrc[v]; e < Q.xsrc[v + 1]; ++e) {
        const double s = vswp[Q.tgt[e]] * sv;
        const i64 b0 = Q.xblk[e], b1 = Q.xblk[e + 1];
        if (b1 == b0 + 1) {
          Q.val[b0] *= s;
          continue;
        }
        for (i64 b = b0; b < b1; ++b) Q.val[b] *= s;
      }
    }
    if (d < tol) break;
  }
  for (i64 v = 0; v < nvtx; ++v)
    for (i64 j = B.xcol[v]; j < B.xcol[v + 1]; ++j) {
      pscl[j] = vscl[v];
      f[j] *= vscl[v];
    }
  for (i64 i = 0; i < nrow; ++i) g[i] *= yscl[i];
}

static double norm2(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

static double scalenorm(const std::vector<double>& v, const std::vector<double>& s) {
  double o = 0.0;
  for (size_t i = 0; i < v.size(); ++i) {
    double q = v[i] / s[i];
    o += q * q;
  }
  return std::sqrt(o);
}

static int cone_degree(const Cone& c, i64 n) {
  switch (c.kind) {
    case CONE_POS: return (int)n;
    case CONE_FREE: return 0;
    case CONE_SOC: return 2;
    case CONE_PSD: return (int)((std::sqrt(1.0 + 8.0 * (double)n) - 1.0) / 2.0 + 0.5);
    default: return 3;
  }
}

// ------------------------------------------------------------------------------------------
// construction
// ------------------------------------------------------------------------------------------

Solver::Solver(const Problem& prob, const Settings& st, const double* p0, const double* d0, const double* y0)
    : set_(st) {
  host_setup(prob);
  device_setup(p0, d0, y0);
}

Solver::Solver(const Problem& prob, const Settings& st, Deferred) : set_(st) { host_setup(prob); }

static double wall_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void bsm_to_csr(const BSM& A, bool transpose, std::vector<int>& ptr, std::vector<int>& col,
                       std::vector<double>& val);

void Solver::host_setup(const Problem& prob) {
  const double t0 = wall_ms();
  for (const Cone& c : prob.K)
    if (c.kind < CONE_POS || c.kind > CONE_POW)
      throw std::invalid_argument("cone kind not supported by this build's device path");
  for (i64 v = 0; v < prob.B.nvtx; ++v) {
    const int k = prob.K[v].kind;
    if ((k == CONE_EXP || k == CONE_POW) && prob.B.ncols(v) != 3)
      throw std::invalid_argument("exponential and power cones have 3 columns");
    if (k == CONE_POW && !(prob.K[v].param > 0.0 && prob.K[v].param < 1.0))
      throw std::invalid_argument("power cone exponent outside (0, 1)");
    if (k == CONE_PSD) {
      i64 nc = prob.B.ncols(v), d = cone_degree(prob.K[v], nc);
      if (d * (d + 1) / 2 != nc) throw std::invalid_argument("PSD cone size is not d(d+1)/2");
      if (d > PSD_BIG_MAX) throw std::invalid_argument("PSD cone larger than the device path's order limit (4096)");
    }
  }
  n_ = prob.B.ncol();
  m_ = prob.B.nrow();
  // the device copies index rows, columns and nonzeros with 32-bit ints
  const i64 imax = std::numeric_limits<int32_t>::max();
  if (n_ > imax || m_ > imax || prob.B.nbnz() > imax || prob.Q.nbnz() > imax)
    throw std::invalid_argument("problem too large for 32-bit device indices (n, m and the stored entries of B "
                                "and Q must each be below 2^31)");
  nu_ = 0;
  for (i64 v = 0; v < prob.B.nvtx; ++v) nu_ += cone_degree(prob.K[v], prob.B.ncols(v));
  mu_t_ = prob.mu;
  // symbkkt (kkt.jl:23-53)
  sym_ = analyse_kkt(prob.B, prob.Q, set_.amalgamate != 0);
  const std::vector<i64>& vp = sym_.vperm;
  P_.Q = symselectvtxs(prob.Q, vp);
  P_.B = selectvtxs(prob.B, vp);
  std::vector<i64> colp;
  for (i64 v : vp)
    for (i64 j = prob.B.xcol[v]; j < prob.B.xcol[v + 1]; ++j) colp.push_back(j);
  P_.f.resize(n_);
  P_.P2.resize(n_);
  for (i64 j = 0; j < n_; ++j) {
    P_.f[j] = prob.f[colp[j]];
    P_.P2[j] = prob.P2[colp[j]];
  }
  P_.g = prob.g;
  P_.P1 = prob.P1;
  for (i64 v : vp) P_.K.push_back(prob.K[v]);
  P_.mu = prob.mu;
  if (set_.scale_max_iter > 0) {
    equilibrate(P_.B, P_.Q, P_.f, P_.g, set_.scale_max_iter, 1e-3, pscl_, yscl_);
  } else {
    pscl_.assign(n_, 1.0);
    yscl_.assign(m_, 1.0);
  }
  nB_ = norm2(P_.B.val);
  nH_ = norm2(P_.Q.val);
  nf_ = norm2(P_.f);
  ng_ = norm2(P_.g);
  sg_ = scalenorm(P_.g, yscl_);
  sf_ = scalenorm(P_.f, pscl_);
  bsm_to_csr(P_.B, false, brp_, brc_, brv_);
  bsm_to_csr(P_.B, true, bcp_, bcr_, bcv_);
  gram_lower(n_, brp_, brc_, brv_, bcp_, bcr_, bcv_, Lp_, Li_, Lx_);
  host_ms_ = wall_ms() - t0;
}

// The arc of Q's diagonal block of vertex v, or −1 where there is none or it holds only zeros (the
// block a PSD vertex always has in Q's layout, which for a large cone would be an n×n block of
// zeros on the device): the cone kernels read no block as zero.
i64 Solver::qdiag_arc(i64 v) const {
  for (i64 e = P_.Q.xsrc[v]; e < P_.Q.xsrc[v + 1]; ++e)
    if (P_.Q.tgt[e] == v) {
      for (i64 t = P_.Q.xblk[e]; t < P_.Q.xblk[e + 1]; ++t)
        if (P_.Q.val[t] != 0.0) return e;
      return -1;
    }
  return -1;
}

// An upper bound of what device_setup allocates: every buffer of build_device and
// build_schedule, from the symbolic structure alone, each rounded up to the size class the
// allocator hands out (the assembly map is bounded by the stored entries of BᵀB and Q).
double Solver::estimate_device_bytes() const {
  double tot = 0.0;
  auto buf = [&](double bytes) { tot += (double)size_class((size_t)std::max(bytes, 8.0)); };
  const double n = (double)std::max<i64>(n_, 1), m = (double)std::max<i64>(m_, 1);
  const i64 NF = sym_.nfront(), N = P_.B.nvtx;
  double fsize = 0.0, ssum = 0.0, nsteps = 0.0;
  int nbig = 0;
  std::vector<double> bigs;  // bytes of each large front's own buffers
  for (i64 f = 0; f < NF; ++f) {
    const double r = (double)(sym_.xcol[sym_.front_vtx[f + 1]] - sym_.xcol[sym_.front_vtx[f]]);
    double sc = 0.0;
    for (i64 u : sym_.front_sep[f]) sc += (double)(sym_.xcol[u + 1] - sym_.xcol[u]);
    const double ld = r + sc;
    fsize += ld * ld;
    ssum += sc;
    if (ld < set_.big_front_ld) {
      nsteps += 2.0 * std::ceil(r / NB);
      continue;
    }
    ++nbig;
    const double T = std::ceil(r / NB) + std::ceil(sc / NB), Tr = std::ceil(r / NB), nt = T * (T + 1) / 2;
    bigs.push_back(4.0 * (T + 1));                           // tile bounds
    bigs.push_back(16.0 * nt);                               // tasks
    bigs.push_back(4.0 * nt);                                // tile flags
    bigs.push_back(8.0 * NB * NB * std::max(1.0, Tr));       // inverted diagonal blocks
    if (set_.precision == 1) {
      bigs.push_back(4.0 * ld * std::max(1.0, r));           // FP32 factor columns
      bigs.push_back(4.0 * NB * NB * std::max(1.0, Tr));     // FP32 inverted diagonal blocks
    }
    bigs.push_back(4.0 * std::max(1.0, Tr));                 // stage flags
    bigs.push_back(16.0 * NB * std::max(1.0, T));            // solve hand-off words
    bigs.push_back(8.0 * NB * std::max(1.0, Tr) * std::max(1.0, std::ceil(sc / NB)));  // solve partials
    bigs.push_back(8.0 * NB * NB * std::max(1.0, Tr));       // M tiles
    bigs.push_back(8.0 * NB * NB * std::max(1.0, Tr));       // N tiles
    bigs.push_back(8.0 * ld * std::max(1.0, r));             // FP64 factor columns
    bigs.push_back(8.0 * NB * NB * std::max(1.0, Tr));       // staged sub-diagonal tiles
  }
  for (double b : bigs) buf(b);
  double dsz = 0.0, csz = 0.0, qsz = 0.0, wsz = 0.0;
  int nbig_psd = 0;
  for (i64 v = 0; v < N; ++v) {
    const double c = (double)P_.B.ncols(v);
    const int k = P_.K[v].kind;
    const double d = k == CONE_PSD ? (double)cone_degree(P_.K[v], P_.B.ncols(v)) : 0.0;
    const bool big = k == CONE_PSD && d > PSD_MAX;
    dsz += k == CONE_POS ? c : (k == CONE_FREE || big ? 0.0 : c * c);
    csz += k == CONE_SOC ? 5 + c
                         : (k == CONE_PSD ? (big ? (double)psd_big_cache((long long)d) : 2 * d * d + d)
                                          : ((k == CONE_EXP || k == CONE_POW) ? tdc::CACHE : 0));
    if (big) {
      wsz += (double)psd_big_ws((long long)d);
      ++nbig_psd;
    }
    if (qdiag_arc(v) >= 0) qsz += c * c;  // a Q diagonal block, where there is one
  }
  buf(8.0 * wsz);
  buf(16.0 * DENSE_SPLIT * std::max(nbig_psd, 1));
  const double nnzB = (double)P_.B.val.size(), nnzQ = (double)P_.Q.val.size();
  const double nunion = (double)Lx_.size() + nnzQ;
  buf(8.0 * NV_COUNT * n); buf(8.0 * MV_COUNT * m); buf(8.0 * NSCALAR); buf(24.0 * n);
  buf(8.0 * fsize); buf(8.0 * dsz); buf(8.0 * csz); buf(8.0 * qsz);
  buf(8.0 * ssum); buf(4.0 * ssum); buf(4.0 * ssum); buf(4.0 * NF);
  buf((double)sizeof(FrontDesc) * NF); buf((double)sizeof(ConeDesc) * N);
  buf(8.0 * DENSE_SPLIT * std::max<i64>(N, 1)); buf(4.0 * N); buf(4.0 * N); buf(4.0 * N); buf(4.0 * N); buf(8.0);
  buf(4.0 * (m + 1)); buf(4.0 * nnzB); buf(8.0 * nnzB);  // B
  buf(4.0 * (n + 1)); buf(4.0 * nnzB); buf(8.0 * nnzB);  // Bᵀ
  buf(4.0 * (n + 1)); buf(4.0 * nnzQ); buf(8.0 * nnzQ);  // Q
  buf(8.0 * (m + 1)); buf(8.0 * (n + 1)); buf(8.0 * (n + 1));  // row tiles
  buf(8.0 * nunion); buf(8.0 * nunion); buf(8.0 * nunion);     // assembly map
  buf(8.0 * (n / 256 + 1)); buf(8.0 * (m / 256 + 1)); buf(8.0); buf(8.0);
  buf(8.0 * (std::max({n / 256 + 1, m / 256 + 1, (double)NF, 1.0}) + 16)); buf(8.0);
  buf(24.0 * std::max<i64>(N, 1)); buf((double)sizeof(InstDev));
  buf(8.0 * std::max(nbig, 1)); buf(8.0 * std::max(nbig, 1));
  buf((double)sizeof(StepDesc) * nsteps); buf(4.0 * NF); buf(4.0 * NF); buf(4.0 * NF);
  return tot + (1 << 20);
}

void Solver::device_setup(const double* p0, const double* d0, const double* y0) {
  const double t0 = wall_ms();
  g_dev_counter = &dev_bytes_;
  Stager stage;
  try {
    HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    hsc_ = scal_alloc();
    HIPCHK(hipEventCreate(&ev0_));
    HIPCHK(hipEventCreate(&ev1_));
    dev_parts_[0] = wall_ms() - t0;
    stage.st = st_;
    g_stager = &stage;
    build_device();
    dev_parts_[1] = wall_ms() - t0 - dev_parts_[0];
    build_schedule();
    g_stager = nullptr;
    stage.finish();
    if (fill_free_) {  // the front cleared and every entry of the pattern written once
      launch_fill(st_, d_F_.as<double>(), (size_t)fsize_, 0.0);
      launch_assemble_union(st_, d_inst_.as<InstDev>(), 0, d_upos_.as<int64_t>(), d_ul_.as<double>(),
                            d_uq_.as<double>(), nunion_, 1);
      static_ok_ = true;
    }
    dev_parts_[2] = wall_ms() - t0 - dev_parts_[0] - dev_parts_[1];
  } catch (...) {
    g_stager = nullptr;
    g_dev_counter = nullptr;
    throw;
  }
  g_dev_counter = nullptr;
  reinit(p0, d0, y0);
  device_ms_ = wall_ms() - t0;
  dev_parts_[3] = device_ms_ - dev_parts_[0] - dev_parts_[1] - dev_parts_[2];
}

Solver::~Solver() {
  if (st_) (void)hipStreamSynchronize(st_);
  for (hipStream_t a : aux_) (void)hipStreamSynchronize(a);
  for (hipEvent_t e : auxev_) (void)hipEventDestroy(e);
  for (hipStream_t a : aux_) (void)hipStreamDestroy(a);
  if (hsc_) scal_free(hsc_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  for (hipEvent_t e : evpool_) (void)hipEventDestroy(e);
  if (st_) (void)hipStreamDestroy(st_);
}

static void bsm_to_csr(const BSM& A, bool transpose, std::vector<int>& ptr, std::vector<int>& col,
                       std::vector<double>& val) {
  // entries (i, j, x) of A; CSR over rows i (or over columns j if transpose). Stored zeros are
  // dropped. A is visited column by column (each column's arcs in target order, rows in order),
  // which is CSR order over columns, and leaves the columns of each CSR row sorted.
  const i64 nr = transpose ? A.ncol() : A.nrow();
  std::vector<int> cnt(nr + 1, 0);
  auto visit = [&](auto&& fn) {
    for (i64 v = 0; v < A.nvtx; ++v) {
      const i64 nc = A.ncols(v), c0 = A.xcol[v];
      if (nc == 1) {  // one column: scalar blocks of nonnegative and free vertices, among others
        for (i64 e = A.xsrc[v]; e < A.xsrc[v + 1]; ++e) {
          const i64 u = A.tgt[e], nrw = A.nrows(u), r0 = A.xrow[u];
          const double* blk = A.val.data() + A.xblk[e];
          for (i64 il = 0; il < nrw; ++il)
            if (blk[il] != 0.0) fn(r0 + il, c0, blk[il]);
        }
        continue;
      }
      for (i64 jl = 0; jl < nc; ++jl)
        for (i64 e = A.xsrc[v]; e < A.xsrc[v + 1]; ++e) {
          const i64 u = A.tgt[e], nrw = A.nrows(u), r0 = A.xrow[u];
          const double* blk = A.val.data() + A.xblk[e] + jl * nrw;
          for (i64 il = 0; il < nrw; ++il)
            if (blk[il] != 0.0) fn(r0 + il, c0 + jl, blk[il]);
        }
    }
  };
  if (transpose) visit([&](i64, i64 j, double) { cnt[j + 1]++; });
  else visit([&](i64 i, i64, double) { cnt[i + 1]++; });
  for (i64 r = 0; r < nr; ++r) cnt[r + 1] += cnt[r];
  ptr.assign(cnt.begin(), cnt.end());
  col.resize(cnt[nr]);
  val.resize(cnt[nr]);
  if (transpose) {
    // column order is CSR order: one pass, no scatter
    i64 q = 0;
    visit([&](i64 i, i64, double x) {
      col[q] = (int)i;
      val[q] = x;
      ++q;
    });
  } else {
    visit([&](i64 i, i64 j, double x) {
      const int q = cnt[i]++;
      col[q] = (int)j;
      val[q] = x;
    });
  }
}

void Solver::build_device() {
  const double tb0 = wall_ms();
  const i64 N = P_.B.nvtx;
  const i64 NF = sym_.nfront();
  // fronts
  fr_.assign(NF, FrontDesc{});
  std::vector<i64>& fofs = fofs_;
  std::vector<std::vector<int>>& sepcols = sepcols_;
  std::vector<int>& colfront = colfront_;
  fofs.assign(NF + 1, 0);
  sepcols.assign(NF, {});
  colfront.assign(n_, 0);
  for (i64 f = 0; f < NF; ++f) {
    i64 res0 = sym_.xcol[sym_.front_vtx[f]];
    i64 r = sym_.xcol[sym_.front_vtx[f + 1]] - res0;
    for (i64 u : sym_.front_sep[f])
      for (i64 j = sym_.xcol[u]; j < sym_.xcol[u + 1]; ++j) sepcols[f].push_back((int)j);
    i64 s = (i64)sepcols[f].size();
    FrontDesc& F = fr_[f];
    F.r = (int)r;
    F.s = (int)s;
    F.ld = (int)(r + s);
    F.res0 = (int)res0;
    F.parent = (int)sym_.front_parent[f];
    F.inst = 0;
    fofs[f + 1] = fofs[f] + (r + s) * (r + s);
    for (i64 j = res0; j < res0 + r; ++j) colfront[j] = (int)f;
  }
  fsize_ = fofs[NF];
  auto fpos = [&](i64 i, i64 j) -> i64 {
    const i64 q = front_pos(i, j);
    if (q < 0) throw std::logic_error("symbolic structure misses an entry of M");
    return q;
  };
  // children lists, relidx, sepcols
  std::vector<std::vector<int>> children(NF);
  for (i64 f = 0; f < NF; ++f)
    if (fr_[f].parent >= 0) children[fr_[f].parent].push_back((int)f);
  std::vector<int> childlist, relidx, allsep;
  std::vector<i64> relofs(NF), sepofs(NF), uofs(NF + 1, 0);
  for (i64 f = 0; f < NF; ++f) {
    fr_[f].child0 = (int)childlist.size();
    fr_[f].nchild = (int)children[f].size();
    for (int c : children[f]) childlist.push_back(c);
    sepofs[f] = (i64)allsep.size();
    allsep.insert(allsep.end(), sepcols[f].begin(), sepcols[f].end());
    relofs[f] = (i64)relidx.size();
    int p = fr_[f].parent;
    for (int col : sepcols[f]) {
      if (p < 0) throw std::logic_error("root front with a separator");
      const FrontDesc& Pf = fr_[p];
      if (col >= Pf.res0 && col < Pf.res0 + Pf.r) {
        relidx.push_back(col - Pf.res0);
      } else {
        auto& sc = sepcols[p];
        auto it = std::lower_bound(sc.begin(), sc.end(), col);
        if (it == sc.end() || *it != col) throw std::logic_error("separator not contained in the parent");
        relidx.push_back(Pf.r + (int)(it - sc.begin()));
      }
    }
    uofs[f + 1] = uofs[f] + fr_[f].s;
  }
  bd_parts_[0] = wall_ms() - tb0;
  // union pattern of BᵀB and Q (lower triangle), column by column
  std::vector<int> qcp, qcr;
  std::vector<double> qcv;
  bsm_to_csr(P_.Q, true, qcp, qcr, qcv);  // Q by columns, rows sorted
  std::vector<int64_t> upos;
  std::vector<double> ul, uq;
  upos.reserve(Lx_.size() + qcv.size() / 2 + 1);
  ul.reserve(upos.capacity());
  uq.reserve(upos.capacity());
  for (i64 j = 0; j < n_; ++j) {
    i64 a = Lp_[j], ae = Lp_[j + 1];
    i64 b = std::lower_bound(qcr.begin() + qcp[j], qcr.begin() + qcp[j + 1], (int)j) - qcr.begin();
    const i64 be = qcp[j + 1];
    while (a < ae || b < be) {
      const i64 ia = a < ae ? Li_[a] : std::numeric_limits<i64>::max();
      const i64 ib = b < be ? qcr[b] : std::numeric_limits<i64>::max();
      const i64 i = std::min(ia, ib);
      const double lv = ia == i ? Lx_[a++] : 0.0, qv = ib == i ? qcv[b++] : 0.0;
      if (lv == 0.0 && qv == 0.0) continue;
      upos.push_back(fpos(i, j));
      ul.push_back(lv);
      uq.push_back(qv);
    }
  }
  nunion_ = (i64)upos.size();
  // Assembly without clearing the front first: one large front (its factor kept beside it, so
  // the factorization never writes it), no separator and no children (nothing else adds into
  // it), cone blocks only on the diagonal, and every diagonal entry assigned by the union scatter
  // (the cones add to the diagonal after it). Every entry the factorization reads is then either
  // assigned afresh or zero since the setup.
  {
    bool ok = NF == 1 && fr_[0].ld >= set_.big_front_ld && fr_[0].s == 0 && fr_[0].nchild == 0;
    for (i64 v = 0; ok && v < N; ++v) ok = P_.K[v].kind == CONE_POS || P_.K[v].kind == CONE_FREE;
    if (ok) {
      std::vector<char> diag(n_, 0);
      const i64 ld = fr_[0].ld;
      for (int64_t q : upos)
        if (q % (ld + 1) == 0) diag[q / (ld + 1)] = 1;
      for (i64 j = 0; ok && j < n_; ++j) ok = diag[j] != 0;
    }
    fill_free_ = ok;
  }
  // For such a front the entries that do not change with δ (no Q part, off the diagonal) are
  // written once; the assembly rewrites only the rest, which go first
  nunion_dyn_ = nunion_;
  if (fill_free_) {
    const i64 ld1 = fr_[0].ld + 1;
    std::vector<int64_t> p2;
    std::vector<double> l2, q2;
    p2.reserve(upos.size());
    l2.reserve(upos.size());
    q2.reserve(upos.size());
    for (int pass = 0; pass < 2; ++pass) {
      for (size_t e = 0; e < upos.size(); ++e) {
        const bool dyn = uq[e] != 0.0 || upos[e] % ld1 == 0;
        if (dyn != (pass == 0)) continue;
        p2.push_back(upos[e]);
        l2.push_back(ul[e]);
        q2.push_back(uq[e]);
      }
      if (pass == 0) nunion_dyn_ = (i64)p2.size();
    }
    upos.swap(p2);
    ul.swap(l2);
    uq.swap(q2);
  }
  bd_parts_[1] = wall_ms() - tb0;
  // ‖Q‖² off the diagonal blocks
  qoff2_ = 0.0;
  for (i64 v = 0; v < N; ++v)
    for (i64 e = P_.Q.xsrc[v]; e < P_.Q.xsrc[v + 1]; ++e)
      if (P_.Q.tgt[e] != v)
        for (i64 b = P_.Q.xblk[e]; b < P_.Q.xblk[e + 1]; ++b) qoff2_ += P_.Q.val[b] * P_.Q.val[b];
  // cones
  ncone_ = (int)N;
  std::vector<i64> dofs(N + 1, 0), cofs(N + 1, 0), qofs(N + 1, 0), wofs(N + 1, 0);
  for (i64 v = 0; v < N; ++v) {
    i64 nc = P_.B.ncols(v);
    int k = P_.K[v].kind;
    const i64 d = k == CONE_PSD ? cone_degree(P_.K[v], nc) : 0;
    const bool td = k == CONE_EXP || k == CONE_POW;
    const bool big = k == CONE_PSD && d > PSD_MAX;  // W ⊗ₛ W not stored; cache R, S, s, W
    dofs[v + 1] = dofs[v] + (k == CONE_POS ? nc : (k == CONE_FREE || big ? 0 : nc * nc));
    cofs[v + 1] = cofs[v] + (k == CONE_SOC ? 5 + nc
                                            : (k == CONE_PSD ? (big ? psd_big_cache(d) : 2 * d * d + d)
                                                             : (td ? tdc::CACHE : 0)));
    wofs[v + 1] = wofs[v] + (big ? psd_big_ws(d) : 0);
  }
  // the diagonal blocks of Q, for the cones that have one (‖H‖ reads them; the others read zero)
  std::vector<i64> qarc(N, -1);
  for (i64 v = 0; v < N; ++v) qarc[v] = qdiag_arc(v);
  for (i64 v = 0; v < N; ++v) {
    const i64 nc = P_.B.ncols(v);
    qofs[v + 1] = qofs[v] + (qarc[v] >= 0 ? nc * nc : 0);
  }
  std::vector<double> qdiag(qofs[N], 0.0);
  for (i64 v = 0; v < N; ++v)
    if (qarc[v] >= 0)
      std::copy(P_.Q.val.begin() + P_.Q.xblk[qarc[v]], P_.Q.val.begin() + P_.Q.xblk[qarc[v] + 1],
                qdiag.begin() + qofs[v]);
  cone_front_.assign(N, 0);
  for (i64 f = 0; f < NF; ++f)
    for (i64 v = sym_.front_vtx[f]; v < sym_.front_vtx[f + 1]; ++v) cone_front_[v] = (int)f;
  bd_parts_[2] = wall_ms() - tb0;
  // device allocations
  d_nv_.alloc(sizeof(double) * NV_COUNT * std::max<i64>(n_, 1));
  d_mv_.alloc(sizeof(double) * MV_COUNT * std::max<i64>(m_, 1));
  d_sc_.alloc(sizeof(double) * NSCALAR);
  d_scr_.alloc(sizeof(double) * 3 * std::max<i64>(n_, 1));
  d_F_.alloc(sizeof(double) * std::max<i64>(fsize_, 1));
  d_D_.alloc(sizeof(double) * std::max<i64>(dofs[N], 1));
  d_cache_.alloc(sizeof(double) * std::max<i64>(cofs[N], 1));
  d_bws_.alloc(sizeof(double) * std::max<i64>(wofs[N], 1));
  d_qdiag_.upload(qdiag, st_);
  d_uvec_.alloc(sizeof(double) * std::max<i64>(uofs[NF], 1));
  d_nv_.clear(st_);
  d_mv_.clear(st_);
  d_sc_.clear(st_);
  d_cache_.clear(st_);
  d_bws_.clear(st_);
  d_D_.clear(st_);
  d_sepcols_.upload(allsep, st_);
  d_relidx_.upload(relidx, st_);
  d_childlist_.upload(childlist, st_);
  hl_.childlist = childlist;
  for (i64 f = 0; f < NF; ++f) {
    fr_[f].F = d_F_.as<double>() + fofs[f];
    fr_[f].F32 = nullptr;
    fr_[f].sepcols = d_sepcols_.as<int>() + sepofs[f];
    fr_[f].relidx = d_relidx_.as<int>() + relofs[f];
    fr_[f].uvec = d_uvec_.as<double>() + uofs[f];
  }
  d_fr_.upload(fr_, st_);
  cones_.assign(N, ConeDesc{});
  for (i64 v = 0; v < N; ++v) {
    ConeDesc& C = cones_[v];
    C.kind = P_.K[v].kind;
    C.inst = 0;
    C.col0 = (int)P_.B.xcol[v];
    C.n = (int)P_.B.ncols(v);
    C.param = P_.K[v].param;
    const bool big = C.kind == CONE_PSD && cone_degree(P_.K[v], C.n) > PSD_MAX;
    C.D = big ? nullptr : d_D_.as<double>() + dofs[v];
    C.cache = d_cache_.as<double>() + cofs[v];
    C.ws = big ? d_bws_.as<double>() + wofs[v] : nullptr;
    C.Qd = qarc[v] >= 0 ? d_qdiag_.as<double>() + qofs[v] : nullptr;
    if (C.kind != CONE_FREE) cone_nmax_ = std::max(cone_nmax_, C.n);
  }
  d_cones_.upload(cones_, st_);
  {
    std::vector<int> psd, dense, td, bpsd, soc;
    for (i64 v = 0; v < N; ++v) {
      const int k = cones_[v].kind;
      const bool big = k == CONE_PSD && cones_[v].D == nullptr;
      if (k == CONE_PSD && !big) psd.push_back((int)v);
      if (big) {
        bpsd.push_back((int)v);
        bpsd_dmax_ = std::max(bpsd_dmax_, cone_degree(P_.K[v], cones_[v].n));
      }
      if ((k == CONE_PSD && !big) || k == CONE_SOC) dense.push_back((int)v);
      if (k == CONE_EXP || k == CONE_POW) td.push_back((int)v);
      if (k == CONE_SOC) soc.push_back((int)v);
    }
    npsd_ = (int)psd.size();
    nsoc_ = (int)soc.size();
    if (nsoc_) d_soclist_.upload(soc, st_);
    nbpsd_ = (int)bpsd.size();
    if (nbpsd_) d_bpsdlist_.upload(bpsd, st_);
    d_bpart_.alloc(sizeof(double) * 2 * DENSE_SPLIT * std::max(nbpsd_, 1));
    ndense_ = (int)dense.size();
    for (int v : dense) dense_nmax_ = std::max(dense_nmax_, cones_[v].n);
    d_dpart_.alloc(sizeof(double) * DENSE_SPLIT * std::max(ndense_, 1));
    ntd_ = (int)td.size();
    if (npsd_) d_psdlist_.upload(psd, st_);
    if (ndense_) d_denselist_.upload(dense, st_);
    if (ntd_) d_tdlist_.upload(td, st_);
    hl_.psd = psd;
    hl_.dense = dense;
    hl_.soc = soc;
    hl_.td = td;
  }
  d_cone_front_.upload(cone_front_, st_);
  d_cone_ofs_.upload(std::vector<int>{0, (int)N}, st_);
  bd_parts_[3] = wall_ms() - tb0;
  // sparse matrices
  std::vector<int> p, c;
  std::vector<double> x;
  int lanes[3];
  // lanes per row: the largest power of two ≤ the mean row length (1..64); tiles of 256/lanes rows
  auto plan = [&](int which, i64 nrow, const std::vector<int>& ptr) {
    const std::vector<int>& p = ptr;
    double avg = nrow > 0 ? (double)p[nrow] / (double)nrow : 0.0;
    int G = 1;
    while (G < 64 && 2.0 * G <= avg) G *= 2;
    lanes[which] = G;
    std::vector<Tile> t;
    for (i64 s = 0; s < nrow; s += TILE_THREADS / G) t.push_back(Tile{0, (int)s});
    for (i64 r = 0; r < nrow; ++r)  // rows the row tiles leave to a workgroup of their own
      if (p[r + 1] - p[r] > SPMV_LONG * G) t.push_back(Tile{0, (int)(-1 - r)});
    nsptile_[which] = (int)t.size();
    if (!t.empty()) d_sptiles_[which].upload(t, st_);
    hl_.sp[which] = std::move(t);
  };
  d_Bp_.upload(brp_, st_);
  d_Bc_.upload(brc_, st_);
  d_Bv_.upload(brv_, st_);
  plan(0, m_, brp_);
  d_Btp_.upload(bcp_, st_);
  d_Btc_.upload(bcr_, st_);
  d_Btv_.upload(bcv_, st_);
  plan(1, n_, bcp_);
  bsm_to_csr(P_.Q, false, p, c, x);
  d_Qp_.upload(p, st_);
  d_Qc_.upload(c, st_);
  d_Qv_.upload(x, st_);
  plan(2, n_, p);
  d_upos_.upload(upos, st_);
  d_ul_.upload(ul, st_);
  d_uq_.upload(uq, st_);
  bd_parts_[4] = wall_ms() - tb0;
  // tiles
  std::vector<Tile> tn, tm;
  for (i64 s = 0; s < n_; s += 256) tn.push_back(Tile{0, (int)s});
  for (i64 s = 0; s < m_; s += 256) tm.push_back(Tile{0, (int)s});
  ntile_n_ = (int)tn.size();
  ntile_m_ = (int)tm.size();
  d_ntiles_.upload(tn, st_);
  d_mtiles_.upload(tm, st_);
  hl_.tn = tn;
  hl_.tm = tm;
  d_ntofs_.upload(std::vector<int>{0, ntile_n_}, st_);
  d_mtofs_.upload(std::vector<int>{0, ntile_m_}, st_);
  d_partial_.alloc(sizeof(double) * (std::max<i64>({(i64)ntile_n_, (i64)ntile_m_, NF, 1}) + 16));
  d_dotcnt_.alloc(sizeof(int) * 2);
  d_dotcnt_.clear(st_);
  d_coneout_.alloc(sizeof(double) * 3 * std::max<i64>(N, 1));
  // instance table
  idev_.nv = d_nv_.as<double>();
  idev_.mv = d_mv_.as<double>();
  idev_.sc = d_sc_.as<double>();
  idev_.scratch = d_scr_.as<double>();
  idev_.F = d_F_.as<double>();
  idev_.B = CSRDev{d_Bp_.as<int>(), d_Bc_.as<int>(), d_Bv_.as<double>(), (int)m_, lanes[0]};
  idev_.Bt = CSRDev{d_Btp_.as<int>(), d_Btc_.as<int>(), d_Btv_.as<double>(), (int)n_, lanes[1]};
  idev_.Q = CSRDev{d_Qp_.as<int>(), d_Qc_.as<int>(), d_Qv_.as<double>(), (int)n_, lanes[2]};
  idev_.n = (int)n_;
  idev_.m = (int)m_;
  d_inst_.upload(std::vector<InstDev>{idev_}, st_);
  // constant vectors: f, g, pscl, yscl
  g_stager->add(idev_.nv + (size_t)NV_F * n_, P_.f.data(), n_ * 8);
  g_stager->add(idev_.nv + (size_t)NV_PSCL * n_, pscl_.data(), n_ * 8);
  g_stager->add(idev_.mv + (size_t)MV_G * m_, P_.g.data(), m_ * 8);
  g_stager->add(idev_.mv + (size_t)MV_YSCL * m_, yscl_.data(), m_ * 8);
  // staged like the rest: a kernel queued now would run before the staged zero fill of the scalars
  const double one = 1.0;
  g_stager->add(d_sc_.as<double>() + S_ACTIVE, &one, sizeof(double));
  hsc_[S_ACTIVE] = 1.0;
}

void Solver::build_schedule() {
  const int NF = (int)fr_.size();
  std::vector<int> h(NF, 0);
  for (int f = 0; f < NF; ++f)
    if (fr_[f].parent >= 0) h[fr_[f].parent] = std::max(h[fr_[f].parent], h[f] + 1);
  int H = 0;
  for (int f = 0; f < NF; ++f) H = std::max(H, h[f] + 1);
  levels_.assign(H, Level{});
  for (int f = 0; f < NF; ++f) levels_[h[f]].fronts.push_back(f);
  std::vector<std::vector<int>> kids(NF);
  for (int c = 0; c < NF; ++c)
    if (fr_[c].parent >= 0) kids[fr_[c].parent].push_back(c);
  std::vector<StepDesc> steps;
  std::vector<int> ea_items, ea_tile0, solve_items;
  // large fronts: persistent tile-DAG factorization and pipelined solves
  int nbig = 0;
  for (int f = 0; f < NF; ++f)
    if (fr_[f].ld >= set_.big_front_ld) ++nbig;
  d_bigtick_.alloc(sizeof(int) * 2 * std::max(nbig, 1));
  d_bigstick_.alloc(sizeof(long long) * std::max(nbig, 1));
  d_bigtick_.clear(st_);
  d_bigstick_.clear(st_);
  {
    int dev = 0, ncu = 256;
    HIPCHK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
    // the DAG kernel's grid in CUs (DAG_PER_CU workgroups on each); a serving slot may take a
    // share of the CUs so that several instances' factorizations run side by side
    dag_grid_ = set_.dag_grid > 0 ? std::min(set_.dag_grid, ncu) : ncu;
  }
  for (int l = 0; l < H; ++l) {
    Level& L = levels_[l];
    std::vector<int> small;
    for (int f : L.fronts) {
      if (fr_[f].ld < set_.big_front_ld) {
        small.push_back(f);
        continue;
      }
      auto bh = std::make_unique<BigHost>();
      const FrontDesc& Fd = fr_[f];
      std::vector<int> tb;
      for (int c = 0; c < Fd.r; c += NB) tb.push_back(c);
      for (int c = Fd.r; c < Fd.ld; c += NB) tb.push_back(c);
      tb.push_back(Fd.ld);
      int T = (int)tb.size() - 1, Tr = (Fd.r + NB - 1) / NB;
      // one task per lower tile. A diagonal task forms the sub-diagonal factor tile of the
      // previous column itself, so its own update loop stops one column early and that tile's
      // task only accumulates and stages it. The ticket order is a topological order of the
      // left-looking dependencies with the diagonal chain one column ahead of the other factor
      // tiles: diag j, then the factor tiles of column j−1 (rows ≥ j+1), then the staged tile
      // (j+1, j). Diag j needs only tickets before it (Dinv_{j−1}, the staged tile (j, j−1),
      // factor tiles of columns ≤ j−2); so does every other task. A grid of few workgroups then
      // still claims each diagonal task early enough to finish its updates before the chain
      // reaches it.
      std::vector<int4> tasks;
      for (int j = 0; j < Tr; ++j) {
        tasks.push_back(int4{0, j, j, std::max(j - 1, 0)});
        if (j > 0)
          for (int i = j + 1; i < T; ++i) tasks.push_back(int4{1, i, j - 1, j - 1});
        if (j + 1 < Tr) tasks.push_back(int4{3, j + 1, j, j});
      }
      if (Tr > 0)
        for (int i = Tr; i < T; ++i) tasks.push_back(int4{1, i, Tr - 1, Tr - 1});
      for (int j = Tr; j < T; ++j)
        for (int i = j; i < T; ++i) tasks.push_back(int4{2, i, j, Tr});
      if ((int)tasks.size() != T * (T + 1) / 2) throw std::logic_error("tile task list incomplete");
      bh->tb.upload(tb, st_);
      bh->tasks.upload(tasks, st_);
      bh->state.alloc(sizeof(int) * std::max(1, T * (T + 1) / 2));
      bh->state.clear(st_);
      bh->dinv.alloc(sizeof(double) * NB * NB * std::max(1, Tr));
      if (set_.precision == 1) {  // FP32 factor columns (ld × r) and inverted diagonal blocks
        bh->f32.alloc(sizeof(float) * (size_t)Fd.ld * std::max(1, Fd.r));
        bh->dinv32.alloc(sizeof(float) * NB * NB * std::max(1, Tr));
      }
      bh->stage.alloc(sizeof(int) * std::max(1, Tr));
      bh->stage.clear(st_);
      bh->sll.alloc(sizeof(unsigned long long) * 2 * NB * std::max(1, T));
      bh->spart.alloc(sizeof(double) * NB * std::max(1, Tr) * std::max(1, (Fd.s + NB - 1) / NB));
      bh->mtile.alloc(sizeof(double) * NB * NB * std::max(1, Tr));
      bh->ntile.alloc(sizeof(double) * NB * NB * std::max(1, Tr));
      // write-once factor storage: the FP64 factor columns beside F, the staged sub-diagonal tiles
      bh->lsep.alloc(sizeof(double) * (size_t)Fd.ld * std::max(1, Fd.r));
      bh->stg.alloc(sizeof(double) * NB * NB * std::max(1, Tr));
      bh->sll.clear(st_);
      BigFront& B = bh->B;
      B.F = Fd.F;
      B.ld = Fd.ld;
      B.r = Fd.r;
      B.s = Fd.s;
      B.T = T;
      B.Tr = Tr;
      B.tb = bh->tb.as<int>();
      B.state = bh->state.as<int>();
      B.nstate = T * (T + 1) / 2;
      B.ticket = d_bigtick_.as<int>() + 2 * big_.size();
      // the abort flag sits in the scalar block, so every scalar read-back also brings it home
      B.abortf = reinterpret_cast<int*>(d_sc_.as<double>() + S_ABORT);
      B.tasks = bh->tasks.as<int4>();
      B.ntask = (int)tasks.size();
      B.Dinv = bh->dinv.as<double>();
      B.F32 = bh->f32.as<float>();
      B.Dinv32 = bh->dinv32.as<float>();
      B.stage = bh->stage.as<int>();
      B.trace = nullptr;
      B.sll = bh->sll.as<unsigned long long>();
      B.spart = bh->spart.as<double>();
      B.Mf = bh->mtile.as<double>();
      B.Nt = bh->ntile.as<double>();
      B.sticket = d_bigstick_.as<long long>() + big_.size();
      B.inst = Fd.inst;
      B.L = bh->lsep.as<double>();
      B.stg = bh->stg.as<double>();
      bh->front = f;
      double r = Fd.r, s = Fd.s;
      L.big_flops += r * r * r / 3.0 + s * r * r + s * s * r;
      L.big.push_back((int)big_.size());
      big_.push_back(std::move(bh));
    }
    L.items_off = (int)solve_items.size();
    L.nsmall = (int)small.size();
    solve_items.insert(solve_items.end(), small.begin(), small.end());
    // extend-add items: children of this level's fronts, in rounds: round k adds the k-th child
    // (with an update matrix) of every front, so no parent entry takes two additions in one launch
    // and each sums its children in a fixed order
    for (int k = 0;; ++k) {
      const int off = (int)ea_items.size();
      int tiles = 0;
      for (int f : L.fronts) {
        int rank = 0;
        for (int c : kids[f]) {
          if (fr_[c].s <= 0) continue;
          if (rank++ != k) continue;
          int T = (fr_[c].s + NB - 1) / NB;
          ea_items.push_back(c);
          ea_tile0.push_back(tiles);
          tiles += T * (T + 1) / 2;
        }
      }
      if (tiles == 0) break;
      L.ea_off.push_back(off);
      L.ea_n.push_back((int)ea_items.size() - off);
      L.ea_tiles.push_back(tiles);
    }
    for (int f : L.fronts) {
      double r = fr_[f].r, s = fr_[f].s;
      L.solve_bytes += (r * (r + 1) / 2 + s * r) * 8.0;
    }
    int kmax = 0;
    for (int f : small) kmax = std::max(kmax, (fr_[f].r + NB - 1) / NB);
    for (int k = 0; k < kmax; ++k) {
      int pb = 0, ub = 0;
      std::vector<StepDesc> pan, upd;
      double pfl = 0.0, ufl = 0.0;
      for (int f : small) {
        int kb = k * NB;
        if (kb >= fr_[f].r) continue;
        int w = std::min(NB, fr_[f].r - kb);
        int T = (fr_[f].ld - kb - w + NB - 1) / NB;
        double rem = (double)(fr_[f].ld - kb - w), ww = (double)w;
        pfl += ww * ww * ww / 3.0 + rem * ww * ww;
        ufl += (rem * rem + rem) * ww;
        pan.push_back(StepDesc{f, kb, w, T, pb});
        pb += 1 + T;
        if (T > 0) {
          upd.push_back(StepDesc{f, kb, w, T, ub});
          ub += T * (T + 1) / 2;
        }
      }
      L.pan_fl.push_back(pfl);
      L.upd_fl.push_back(ufl);
      L.pan_off.push_back((int)steps.size());
      L.pan_n.push_back((int)pan.size());
      L.pan_blk.push_back(pb);
      steps.insert(steps.end(), pan.begin(), pan.end());
      L.upd_off.push_back((int)steps.size());
      L.upd_n.push_back((int)upd.size());
      L.upd_blk.push_back(ub);
      steps.insert(steps.end(), upd.begin(), upd.end());
    }
  }
  d_steps_.upload(steps, st_);
  d_panflag_.alloc(sizeof(int) * std::max<size_t>(steps.size(), 1));
  d_panflag_.clear(st_);
  d_ea_items_.upload(ea_items, st_);
  d_ea_tile0_.upload(ea_tile0, st_);
  d_solve_items_.upload(solve_items, st_);
  hl_.steps = steps;
  hl_.ea_items = ea_items;
  hl_.ea_tile0 = ea_tile0;
  hl_.solve_items = solve_items;
  HIPCHK(hipStreamSynchronize(st_));
  // the large fronts of one level are independent: up to NAUX at a time, each on a stream of its
  // own with a share of the CUs (a chordal clique tree has levels of tens of them)
  size_t maxbig = 0;
  for (const Level& L : levels_) maxbig = std::max(maxbig, L.big.size());
  naux_ = maxbig > 1 ? (int)std::min<size_t>(NAUX, maxbig) : 0;
  for (int k = 0; k < naux_; ++k) {
    hipStream_t a = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
    aux_.push_back(a);
  }
  for (int k = 0; k <= naux_ && naux_ > 0; ++k) {
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    auxev_.push_back(e);
  }
}

// streams 0..k−1 of the auxiliary set start after the work queued on the solver's stream so far
void Solver::fork_aux(int k) {
  HIPCHK(hipEventRecord(auxev_[0], st_));
  for (int i = 0; i < k; ++i) HIPCHK(hipStreamWaitEvent(aux_[i], auxev_[0], 0));
}

// the solver's stream continues after everything queued on streams 0..k−1
void Solver::join_aux(int k) {
  for (int i = 0; i < k; ++i) {
    HIPCHK(hipEventRecord(auxev_[1 + i], aux_[i]));
    HIPCHK(hipStreamWaitEvent(st_, auxev_[1 + i], 0));
  }
}

// position of (i, j), i ≥ j, in the frontal storage; −1 outside the factor's structure
i64 Solver::front_pos(i64 i, i64 j) const {
  const int f = colfront_[j];
  const FrontDesc& F = fr_[f];
  const i64 jl = j - F.res0;
  i64 il;
  if (i < F.res0 + F.r) {
    if (i < j) return -1;
    il = i - F.res0;
  } else {
    const auto& sc = sepcols_[f];
    auto it = std::lower_bound(sc.begin(), sc.end(), (int)i);
    if (it == sc.end() || *it != i) return -1;
    il = F.r + (it - sc.begin());
  }
  return fofs_[f] + il + jl * F.ld;
}

// The caller's H in place of the solver's (Q plus the cone blocks) for the KKT plugin methods:
// its CSR rows take the place of Q's in the products, and the assembly adds δH instead of δQ and
// the cone blocks. use_user_h(false) brings back the solver's own H.
void Solver::use_user_h(bool on) {
  if (on == user_h_) return;
  std::swap(idev_.Q, hsave_);
  std::swap(d_sptiles_[2].p, d_htiles_.p);
  std::swap(d_sptiles_[2].bytes, d_htiles_.bytes);
  std::swap(nsptile_[2], nhtiles_);
  HIPCHK(hipMemcpyAsync(d_inst_.p, &idev_, sizeof(InstDev), hipMemcpyHostToDevice, st_));
  HIPCHK(hipStreamSynchronize(st_));
  user_h_ = on;
}

int Solver::kkt_factor_h(const BSM& H, double delta, double* rho) {
  if (H.nvtx != P_.B.nvtx || H.nout != P_.B.nvtx || H.xcol != P_.B.xcol || H.xrow != P_.B.xcol)
    throw std::invalid_argument("H must be block-sparse over the solver's vertex partition (mb_solver_layout)");
  use_user_h(false);
  std::vector<int> p, c;
  std::vector<double> x;
  bsm_to_csr(H, false, p, c, x);
  // rows of H as Q's are tiled (lanes per row from the mean row length; long rows on their own)
  double avg = n_ > 0 ? (double)c.size() / (double)n_ : 0.0;
  int G = 1;
  while (G < 64 && 2.0 * G <= avg) G *= 2;
  std::vector<Tile> t;
  for (i64 r = 0; r < n_; r += TILE_THREADS / G) t.push_back(Tile{0, (int)r});
  for (i64 r = 0; r < n_; ++r)
    if (p[r + 1] - p[r] > SPMV_LONG * G) t.push_back(Tile{0, (int)(-1 - r)});
  d_Hp_.upload(p, st_);
  d_Hc_.upload(c, st_);
  d_Hv_.upload(x, st_);
  d_htiles_.upload(t, st_);
  nhtiles_ = (int)t.size();
  hsave_ = CSRDev{d_Hp_.as<int>(), d_Hc_.as<int>(), d_Hv_.as<double>(), (int)n_, G};
  // assembly positions of the lower triangle
  std::vector<int64_t> pos;
  std::vector<double> val;
  for (i64 v = 0; v < H.nvtx; ++v) {
    const i64 nc = H.ncols(v), c0 = H.xcol[v];
    for (i64 e = H.xsrc[v]; e < H.xsrc[v + 1]; ++e) {
      const i64 u = H.tgt[e], nr = H.nrows(u), r0 = H.xrow[u];
      for (i64 jl = 0; jl < nc; ++jl)
        for (i64 il = 0; il < nr; ++il) {
          const i64 i = r0 + il, j = c0 + jl;
          const double h = H.val[H.xblk[e] + jl * nr + il];
          if (i < j || h == 0.0) continue;
          const i64 q = front_pos(i, j);
          if (q < 0) throw std::invalid_argument("H has an entry outside the factor's structure (the pattern of Q, BᵀB "
                                                 "and the diagonal blocks)");
          pos.push_back(q);
          val.push_back(h);
        }
    }
  }
  d_hpos_.upload(pos, st_);
  d_hval_.upload(val, st_);
  nhent_ = (i64)pos.size();
  use_user_h(true);
  double r = 0.0;
  set_scalar(S_ALPHA, 1.0 / delta);
  const bool ok = factor(delta, r);
  if (rho) *rho = r;
  return ok ? 1 : 0;
}

double Solver::front_flops() const {
  double t = 0.0;
  for (const FrontDesc& F : fr_) {
    double r = F.r, s = F.s;
    t += r * r * r / 3.0 + s * r * r + s * s * r;
  }
  return t;
}

// ------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------

// The solver whose own launches the calling thread is issuing for a gang (the leader of a tick),
// so that they launch instead of going to the gang again. Per thread: the member's own thread
// goes on handing in ops meanwhile.
static thread_local const Solver* g_issuing = nullptr;
void Solver::set_issuing(const Solver* s) { g_issuing = s; }
bool Solver::ganged() const { return gang_ != nullptr && g_issuing != this; }

void Solver::gang_submit(GOp* op) { gang_->submit(this, op); }

void Solver::sync_scalars() {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_READBACK;
    op.blocking = true;
    op.solo = [this] { raw_sync_scalars(); };
    gang_submit(&op);
    return;
  }
  raw_sync_scalars();
}

void Solver::raw_sync_scalars() {
  HIPCHK(hipMemcpyAsync(hsc_, d_sc_.p, NSCALAR * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  prof_flush();
}

void Solver::set_scalar(int slot, double v) {
  hsc_[slot] = v;
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SET2;
    op.a[0] = slot;
    op.a[1] = -1;
    op.v[0] = v;
    op.solo = [this, slot, v] { launch_set_scalar(st_, d_sc_.as<double>(), slot, v); };
    gang_submit(&op);
    return;
  }
  // the value travels as a kernel argument: no host buffer outlives the call
  launch_set_scalar(st_, d_sc_.as<double>(), slot, v);
}

void Solver::set_scalar2(int s1, double v1, int s2, double v2) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SET2;
    op.a[0] = s1;
    op.a[1] = s2;
    op.v[0] = v1;
    op.v[1] = v2;
    op.solo = [this, s1, v1, s2, v2] { launch_set_scalar2(st_, d_sc_.as<double>(), s1, v1, s2, v2); };
    gang_submit(&op);
    return;
  }
  launch_set_scalar2(st_, d_sc_.as<double>(), s1, v1, s2, v2);
}

void Solver::smu(double nu, double mu, double mut) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SMU;
    op.v[0] = nu;
    op.v[1] = mu;
    op.v[2] = mut;
    op.solo = [this, nu, mu, mut] { launch_smu(st_, d_sc_.as<double>(), nu, mu, mut); };
    gang_submit(&op);
    return;
  }
  launch_smu(st_, d_sc_.as<double>(), nu, mu, mut);
}

void Solver::cg_scalars(int which) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_CGSC;
    op.a[0] = which;
    op.solo = [this, which] { launch_cg_scalars(st_, d_inst_.as<InstDev>(), 1, which); };
    gang_submit(&op);
    return;
  }
  launch_cg_scalars(st_, d_inst_.as<InstDev>(), 1, which);
}

void Solver::axpby(int space, int kx, int ky, int kz, Coef a, Coef b) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_AXPBY;
    op.a[0] = space, op.a[1] = kx, op.a[2] = ky, op.a[3] = kz;
    op.c[0] = a, op.c[1] = b;
    op.solo = [=] { axpby(space, kx, ky, kz, a, b); };
    gang_submit(&op);
    return;
  }
  const InstDev* I = d_inst_.as<InstDev>();
  if (space == 0) launch_axpby(st_, I, d_ntiles_.as<Tile>(), ntile_n_, 0, kx, ky, kz, a, b);
  else launch_axpby(st_, I, d_mtiles_.as<Tile>(), ntile_m_, 1, kx, ky, kz, a, b);
}

void Solver::axpbypcz(int space, int kx, int ky, int kw, int kz, Coef a, Coef b, Coef c) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_AXPBYPCZ;
    op.a[0] = space, op.a[1] = kx, op.a[2] = ky, op.a[3] = kw, op.a[4] = kz;
    op.c[0] = a, op.c[1] = b, op.c[2] = c;
    op.solo = [=] { axpbypcz(space, kx, ky, kw, kz, a, b, c); };
    gang_submit(&op);
    return;
  }
  const InstDev* I = d_inst_.as<InstDev>();
  if (space == 0) launch_axpbypcz(st_, I, d_ntiles_.as<Tile>(), ntile_n_, 0, kx, ky, kw, kz, a, b, c);
  else launch_axpbypcz(st_, I, d_mtiles_.as<Tile>(), ntile_m_, 1, kx, ky, kw, kz, a, b, c);
}

void Solver::dot(int space, int kx, int ky, int mode, int slot, int op, int kw0, int kw1) {
  const InstDev* I = d_inst_.as<InstDev>();
  if (space == 0)
    launch_dot(st_, I, d_ntiles_.as<Tile>(), ntile_n_, d_ntofs_.as<int>(), 1, 0, kx, ky, mode, kw0, kw1,
               d_partial_.as<double>(), d_dotcnt_.as<int>(), slot, op);
  else
    launch_dot(st_, I, d_mtiles_.as<Tile>(), ntile_m_, d_mtofs_.as<int>(), 1, 1, kx, ky, mode, kw0, kw1,
               d_partial_.as<double>(), d_dotcnt_.as<int>(), slot, op);
}

// several dot products in one launch while the vectors are short enough for one workgroup each
void Solver::dots(const DotBatch& b) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_DOTS;
    op.db = b;
    op.solo_only = std::max(n_, m_) > DOTS_LEN_MAX;
    op.solo = [this, b] { dots(b); };
    gang_submit(&op);
    return;
  }
  if (std::max(n_, m_) <= DOTS_LEN_MAX) {
    launch_dots(st_, d_inst_.as<InstDev>(), 0, b);
    return;
  }
  for (int k = 0; k < b.n; ++k) {
    const DotSpec& d = b.s[k];
    dot(d.space, d.kx, d.ky, d.mode, d.slot, d.op, d.kw0, d.kw1);
  }
}

void Solver::lin(const LinSet& N, const LinSet& M) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_LIN;
    op.N = N;
    op.M = M;
    op.solo = [this, N, M] { lin(N, M); };
    gang_submit(&op);
    return;
  }
  launch_lin(st_, d_inst_.as<InstDev>(), d_ntiles_.as<Tile>(), ntile_n_, d_mtiles_.as<Tile>(), ntile_m_, N, M);
}

void Solver::lin2(const Lin& on, const Lin& om) { lin(LinSet().add(on), LinSet().add(om)); }

// n-row and m-row sparse products in one launch (either may be absent). The n rows take the row
// tiling (lanes per row) of Q or Bᵀ, whichever of the products present has more lanes per row.
void Solver::spmv_nm(int ntm, const SpmvN* N, const SpmvM* M) {
  static const SpmvN n0{-1, -1, -1, -1, -1, -1, K(0.0), K(0.0), K(0.0), K(0.0), K(0.0)};
  static const SpmvM m0{-1, -1, -1, -1, K(0.0), K(0.0), K(0.0), 0};
  if (N && N->kxq >= 0 && N->kxt >= 0) ntm = idev_.Bt.lanes > idev_.Q.lanes ? 1 : 2;
  else if (N && N->kxq >= 0) ntm = 2;
  else if (N && N->kxt >= 0) ntm = 1;
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SPMVNM;
    op.a[0] = ntm;
    op.a[1] = N ? 1 : 0;
    op.a[2] = M ? 1 : 0;
    op.sn = N ? *N : n0;
    op.sm = M ? *M : m0;
    const SpmvN nn = op.sn;
    const SpmvM mm = op.sm;
    const bool hn = N != nullptr, hm = M != nullptr;
    op.solo = [=] { launch_spmv_nm(st_, d_inst_.as<InstDev>(), ntm, d_sptiles_[ntm].as<Tile>(), hn ? nsptile_[ntm] : 0,
                                   d_sptiles_[0].as<Tile>(), hm ? nsptile_[0] : 0, nn, mm); };
    gang_submit(&op);
    return;
  }
  launch_spmv_nm(st_, d_inst_.as<InstDev>(), ntm, d_sptiles_[ntm].as<Tile>(), N ? nsptile_[ntm] : 0,
                 d_sptiles_[0].as<Tile>(), M ? nsptile_[0] : 0, N ? *N : n0, M ? *M : m0);
}

void Solver::spmv(int which, int kx, int ky, Coef a, Coef b) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SPMV;
    op.a[0] = which, op.a[1] = kx, op.a[2] = ky;
    op.c[0] = a, op.c[1] = b;
    op.solo = [=] { spmv(which, kx, ky, a, b); };
    gang_submit(&op);
    return;
  }
  const InstDev* I = d_inst_.as<InstDev>();
  launch_spmv(st_, I, d_sptiles_[which].as<Tile>(), nsptile_[which], which, kx, ky, a, b);
}

// scale! (cone.jl:146-247) with the cone blocks' part of ‖H‖² (into S_TMP1)
void Solver::cone_scale_norm() {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SCALE;
    op.solo_only = nbpsd_ > 0;
    op.solo = [this] { cone_scale_norm(); };
    gang_submit(&op);
    return;
  }
  launch_cone_scale(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, d_cone_ofs_.as<int>(), 1,
                    d_coneout_.as<double>(), d_psdlist_.as<int>(), npsd_, bpsd(), d_soclist_.as<int>(), nsoc_);
  launch_cone_hnorm2(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, d_cone_ofs_.as<int>(), 1,
                     d_coneout_.as<double>() + 2 * ncone_, S_TMP1, d_denselist_.as<int>(), ndense_, dense_nmax_,
                     d_dpart_.as<double>(), bpsd(), d_bpart_.as<double>());
}

// initpredictor! with μt > 0: the right-hand side r = −d + σμ(−f′(p)) per cone into kind kr
void Solver::corr0_op(int kr) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_CORR0;
    op.a[0] = kr;
    op.solo_only = nbpsd_ > 0;
    op.solo = [this, kr] { corr0_op(kr); };
    gang_submit(&op);
    return;
  }
  launch_cone_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, kr);
  launch_psd_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), d_psdlist_.as<int>(), npsd_, kr);
  launch_bpsd_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), bpsd(), kr);
}

// initcorrector! per cone into kind kr from the affine direction (kdp, kdd)
void Solver::corr_op(int kr, int kdp, int kdd) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_CORR;
    op.a[0] = kr, op.a[1] = kdp, op.a[2] = kdd;
    op.solo_only = nbpsd_ > 0;
    op.solo = [this, kr, kdp, kdd] { corr_op(kr, kdp, kdd); };
    gang_submit(&op);
    return;
  }
  launch_cone_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, kr, kdp, kdd);
  launch_psd_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), d_psdlist_.as<int>(), npsd_, kr, kdp, kdd);
  launch_bpsd_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), bpsd(), kr, kdp, kdd);
}

void Solver::zerofree(int kx) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_ZEROFREE;
    op.a[0] = kx;
    op.solo = [this, kx] { zerofree(kx); };
    gang_submit(&op);
    return;
  }
  launch_zerofree(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, kx);
}

void Solver::lsolve(int kv) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_LSOLVE;
    op.park = true;
    op.a[0] = kv;
    op.a[1] = fac_fp32_ ? 1 : 0;
    const bool fp32 = fac_fp32_;
    op.solo = [this, kv, fp32] { raw_lsolve(kv, fp32); };
    gang_submit(&op);
    return;
  }
  raw_lsolve(kv, fac_fp32_);
}

void Solver::usolve(int kv) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_USOLVE;
    op.park = true;
    op.a[0] = kv;
    op.a[1] = fac_fp32_ ? 1 : 0;
    const bool fp32 = fac_fp32_;
    op.solo = [this, kv, fp32] { raw_usolve(kv, fp32); };
    gang_submit(&op);
    return;
  }
  raw_usolve(kv, fac_fp32_);
}

void Solver::raw_lsolve(int kv, bool fp32) {
  for (const Level& L : levels_) {
    int ps = prof_begin(PROF_FSOLVE);
    launch_fsolve(st_, d_fr_.as<FrontDesc>(), d_solve_items_.as<int>() + L.items_off, L.nsmall,
                  d_childlist_.as<int>(), d_inst_.as<InstDev>(), kv);
    const int k = L.big.size() > 1 ? std::min<int>(naux_, (int)L.big.size()) : 0;
    if (k) fork_aux(k);
    for (size_t i = 0; i < L.big.size(); ++i) {
      BigHost& bh = *big_[L.big[i]];
      epoch_ = next_epoch();
      bh.sbase += launch_big_fsolve(k ? aux_[i % k] : st_, bh.B, d_fr_.as<FrontDesc>(), bh.front,
                                    d_childlist_.as<int>(), d_inst_.as<InstDev>(), kv, epoch_, bh.sbase, fp32);
    }
    if (k) join_aux(k);
    prof_end(ps, PROF_FSOLVE, 0.0, L.solve_bytes);
  }
}

void Solver::raw_usolve(int kv, bool fp32) {
  for (int l = (int)levels_.size() - 1; l >= 0; --l) {
    const Level& L = levels_[l];
    int ps = prof_begin(PROF_BSOLVE);
    launch_bsolve(st_, d_fr_.as<FrontDesc>(), d_solve_items_.as<int>() + L.items_off, L.nsmall,
                  d_inst_.as<InstDev>(), kv);
    const int k = L.big.size() > 1 ? std::min<int>(naux_, (int)L.big.size()) : 0;
    if (k) fork_aux(k);
    for (size_t i = 0; i < L.big.size(); ++i) {
      BigHost& bh = *big_[L.big[i]];
      epoch_ = next_epoch();
      bh.sbase += launch_big_bsolve(k ? aux_[i % k] : st_, bh.B, d_fr_.as<FrontDesc>(), bh.front,
                                    d_inst_.as<InstDev>(), kv, epoch_, bh.sbase, fp32);
    }
    if (k) join_aux(k);
    prof_end(ps, PROF_BSOLVE, 0.0, L.solve_bytes);
  }
}

// a persistent kernel that hit its wait limit raised its abort flag: fail loudly. The flag is
// part of the scalar block, so this reads the copy of the last sync_scalars, which every caller
// has made after the kernels it checks.
// The device raises the flag as an int in the slot's low word, which read as a double is a
// subnormal: the test reads the word's bits, so a caller running with denormals-are-zero still
// sees it.
void Solver::check_abort() {
  int32_t raised = 0;
  std::memcpy(&raised, &hsc_[S_ABORT], sizeof(raised));
  if (raised != 0) throw std::runtime_error("front kernel exceeded its wait limit");
}

void Solver::run_factor(bool fp32) {
  const FrontDesc* F = d_fr_.as<FrontDesc>();
  const StepDesc* S = d_steps_.as<StepDesc>();
  const InstDev* I = d_inst_.as<InstDev>();
  fac_epoch_ = next_epoch();
  int pf = prof_begin(PROF_FACTOR);
  for (const Level& L : levels_) {
    int pe = prof_begin(PROF_EXTEND);
    for (size_t k = 0; k < L.ea_off.size(); ++k)
      launch_extend_add(st_, F, d_ea_items_.as<int>() + L.ea_off[k], d_ea_tile0_.as<int>() + L.ea_off[k], L.ea_n[k],
                        L.ea_tiles[k], I);
    prof_end(pe, PROF_EXTEND, 0.0, 0.0);
    for (size_t k = 0; k < L.pan_off.size(); ++k) {
      int pp = prof_begin(PROF_PANEL);
      launch_panel(st_, F, S + L.pan_off[k], L.pan_n[k], L.pan_blk[k], I, d_panflag_.as<int>() + L.pan_off[k],
                   fac_epoch_, reinterpret_cast<int*>(d_sc_.as<double>() + S_ABORT));
      prof_end(pp, PROF_PANEL, L.pan_fl[k], 0.0);
      if (L.upd_blk[k] == 0) continue;
      int pu = prof_begin(PROF_UPDATE);
      launch_update(st_, F, S + L.upd_off[k], L.upd_n[k], L.upd_blk[k], I);
      prof_end(pu, PROF_UPDATE, L.upd_fl[k], 0.0);
    }
    if (L.big.size() == 1) {
      int pd = prof_begin(PROF_DAG);
      launch_dag_factor(st_, big_[L.big[0]]->B, I, dag_grid_, fac_epoch_, fp32);
      prof_end(pd, PROF_DAG, L.big_flops, 0.0);
      launch_big_mtiles(st_, big_[L.big[0]]->B, I, fp32);
    } else if (!L.big.empty()) {
      const int k = std::min<int>(naux_, (int)L.big.size());
      const int grid = std::max(16, dag_grid_ / k);
      int pd = prof_begin(PROF_DAG);
      fork_aux(k);
      for (size_t i = 0; i < L.big.size(); ++i) {
        const BigFront& B = big_[L.big[i]]->B;
        launch_dag_factor(aux_[i % k], B, I, grid, fac_epoch_, fp32);
        launch_big_mtiles(aux_[i % k], B, I, fp32);
      }
      join_aux(k);
      prof_end(pd, PROF_DAG, L.big_flops, 0.0);
    }
  }
  prof_end(pf, PROF_FACTOR, sym_.exact_flops, 0.0);
}

int Solver::prof_begin(int kind) {
  if (!(prof_mask_ & (1 << kind))) return -1;
  int slot = prof_next_++;
  while ((int)evpool_.size() < 2 * prof_next_) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    evpool_.push_back(e);
  }
  HIPCHK(hipEventRecord(evpool_[2 * slot], st_));
  return slot;
}

void Solver::prof_end(int slot, int kind, double flops, double bytes) {
  if (slot < 0) return;
  HIPCHK(hipEventRecord(evpool_[2 * slot + 1], st_));
  pending_.push_back(ProfRec{slot, kind, flops, bytes});
}

void Solver::prof_flush() {
  if (pending_.empty()) return;
  HIPCHK(hipStreamSynchronize(st_));
  for (const ProfRec& r : pending_) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, evpool_[2 * r.ev], evpool_[2 * r.ev + 1]));
    prof_ms_[r.kind] += ms;
    prof_n_[r.kind] += 1;
    prof_fl_[r.kind] += r.flops;
    prof_by_[r.kind] += r.bytes;
  }
  pending_.clear();
  prof_next_ = 0;
}

void Solver::set_trace(bool on) {
  for (auto& bh : big_) {
    if (on && bh->trace.p == nullptr) {
      bh->trace.alloc(sizeof(unsigned long long) * 8 * std::max(1, bh->B.ntask));
      HIPCHK(hipMemsetAsync(bh->trace.p, 0, bh->trace.bytes, st_));
    }
    bh->B.trace = on ? bh->trace.as<unsigned long long>() : nullptr;
  }
}

int Solver::get_trace(int b, long long* out, int len) {
  if (b < 0 || b >= (int)big_.size()) return 0;
  BigHost& bh = *big_[b];
  int n = 8 * bh.B.ntask;
  if (bh.trace.p == nullptr || out == nullptr) return n;
  HIPCHK(hipStreamSynchronize(st_));
  HIPCHK(hipMemcpyAsync(out, bh.trace.p, sizeof(long long) * std::min(n, len), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return n;
}

void Solver::prof_get(double* out) const {
  for (int k = 0; k < PROF_NKIND; ++k) {
    out[4 * k] = prof_ms_[k];
    out[4 * k + 1] = prof_n_[k];
    out[4 * k + 2] = prof_fl_[k];
    out[4 * k + 3] = prof_by_[k];
  }
}

// F ← BᵀB + δH (the solver's H: Q and the cone blocks; or the caller's H)
void Solver::raw_assemble() {
  const InstDev* I = d_inst_.as<InstDev>();
  const bool part = fill_free_ && static_ok_ && !user_h_;  // only the entries that change with δ
  if (!part) launch_fill(st_, d_F_.as<double>(), (size_t)fsize_, 0.0);
  if (user_h_) {  // BᵀB + δ (the caller's H)
    launch_assemble_union(st_, I, 0, d_upos_.as<int64_t>(), d_ul_.as<double>(), d_uq_.as<double>(), nunion_, 0);
    launch_scatter_add(st_, I, 0, d_hpos_.as<int64_t>(), d_hval_.as<double>(), nhent_);
    static_ok_ = false;  // the front holds the caller's H now
    return;
  }
  launch_assemble_union(st_, I, 0, d_upos_.as<int64_t>(), d_ul_.as<double>(), d_uq_.as<double>(),
                        part ? nunion_dyn_ : nunion_, 1);
  if (fill_free_) static_ok_ = true;
  launch_bpsd_assemble(st_, I, d_cones_.as<ConeDesc>(), bpsd(), d_cone_front_.as<int>(), d_fr_.as<FrontDesc>());
  launch_assemble_cones(st_, I, d_cones_.as<ConeDesc>(), d_cone_front_.as<int>(), d_fr_.as<FrontDesc>(), ncone_,
                        cone_nmax_);
}

// assembly and factorization, timed by events when the solver runs alone
void Solver::raw_factor_run(bool fp32) {
  raw_assemble();
  if (!gang_) HIPCHK(hipEventRecord(ev0_, st_));
  run_factor(fp32);
  if (!gang_) HIPCHK(hipEventRecord(ev1_, st_));
}

void Solver::factor_run(bool fp32) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_FACTOR;
    op.park = true;
    op.a[0] = fp32 ? 1 : 0;
    op.solo_only = user_h_ || nbpsd_ > 0 || (fill_free_ && !static_ok_);
    op.solo = [this, fp32] { raw_factor_run(fp32); };
    gang_submit(&op);
    return;
  }
  raw_factor_run(fp32);
}

// initkkt! (uzawa.jl:73-115): F Fᵀ = δH + BᵀB, on failure retry with ρ = eps ‖δH + BᵀB‖ I
bool Solver::factor(double delta, double& rho) {
  set_scalar2(S_DELTA, delta, S_FAIL, 0.0);
  // mixed precision: the large fronts in FP32 while μ is above the switch point; an FP32
  // breakdown (non-positive pivot) is redone in FP64 before any diagonal shift
  fac_fp32_ = set_.precision == 1 && !mixed_off_ && !big_.empty() && cur_mu_ > set_.mixed_mu;
  factor_run(fac_fp32_);
  sync_scalars();
  check_abort();
  if (fac_fp32_ && hsc_[S_FAIL] != 0.0) {
    fac_fp32_ = false;
    set_scalar(S_FAIL, 0.0);
    factor_run(false);
    sync_scalars();
    check_abort();
  }
  if (fac_fp32_) ++fac_fp32_calls_;
  if (!gang_) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev0_, ev1_));
    fac_ms_ = ms;
    fac_ms_tot_ += ms;
  }
  fac_calls_ += 1;
  rho = 0.0;
  if (hsc_[S_FAIL] == 0.0) return true;
  // the diagonal shift: ‖F‖ of the assembled matrix, then F + ρI factored again (the solver's own
  // launches, also inside a gang)
  double s = 0.0;
  auto fnorm = [this, &s] {
    raw_assemble();
    launch_front_norm2(st_, d_fr_.as<FrontDesc>(), (int)fr_.size(), d_partial_.as<double>());
    std::vector<double> part(fr_.size());
    HIPCHK(hipMemcpyAsync(part.data(), d_partial_.p, part.size() * 8, hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    for (double x : part) s += x;
  };
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SOLO;
    op.blocking = true;
    op.solo = fnorm;
    gang_submit(&op);
  } else {
    fnorm();
  }
  rho = EPS * std::sqrt(s);
  set_scalar(S_RHO, rho);
  set_scalar(S_FAIL, 0.0);
  auto refactor = [this] {
    launch_add_diag(st_, d_fr_.as<FrontDesc>(), (int)fr_.size(), d_inst_.as<InstDev>());
    run_factor(false);
  };
  fac_fp32_ = false;
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_SOLO;
    op.blocking = true;
    op.solo = refactor;
    gang_submit(&op);
  } else {
    refactor();
  }
  sync_scalars();
  check_abort();
  return hsc_[S_FAIL] == 0.0;
}

int Solver::kkt_factor(double delta, double* rho) {
  use_user_h(false);
  double r = 0.0;
  set_scalar(S_ALPHA, 1.0 / delta);  // the Uzawa step α = 1/δ of a following kkt_solve
  bool ok = factor(delta, r);
  if (rho) *rho = r;
  return ok ? 1 : 0;
}

// ------------------------------------------------------------------------------------------
// cone fan-outs through the C ABI (the plugin-level boundary: scale!, initpredictor!,
// initcorrector!, maxsteps of cone.jl:146-553), on host vectors in internal coordinates
// ------------------------------------------------------------------------------------------

bool Solver::cone_scale(const double* p, const double* d, double* spsd) {
  use_user_h(false);
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_P * n_, p, n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_D * n_, d, n_ * 8, hipMemcpyHostToDevice, st_));
  launch_cone_scale(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, d_cone_ofs_.as<int>(), 1,
                    d_coneout_.as<double>(), d_psdlist_.as<int>(), npsd_, bpsd(), d_soclist_.as<int>(), nsoc_);
  sync_scalars();
  if (spsd) *spsd = hsc_[S_SPSD];
  return hsc_[S_FLAGOK] != 0.0;
}

// the blocks in cone order; a large PSD cone's block W ⊗ₛ W is formed on the device for the copy
void Solver::cone_blocks(double* out) {
  std::vector<double> D(d_D_.bytes / sizeof(double));
  HIPCHK(hipMemcpyAsync(D.data(), d_D_.p, d_D_.bytes, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  size_t o = 0, od = 0;
  for (int v = 0; v < ncone_; ++v) {
    const ConeDesc& C = cones_[v];
    const size_t nc = (size_t)C.n;
    if (C.kind == CONE_FREE) continue;
    if (C.kind == CONE_POS) {
      std::copy(D.begin() + od, D.begin() + od + nc, out + o);
      o += nc;
      od += nc;
    } else if (C.D) {
      std::copy(D.begin() + od, D.begin() + od + nc * nc, out + o);
      o += nc * nc;
      od += nc * nc;
    } else {
      DevBuf blk, one;
      blk.alloc(sizeof(double) * nc * nc);
      one.upload(std::vector<int>{v}, st_);
      launch_bpsd_block(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), one.as<int>(), C.n, blk.as<double>());
      HIPCHK(hipMemcpyAsync(out + o, blk.p, sizeof(double) * nc * nc, hipMemcpyDeviceToHost, st_));
      HIPCHK(hipStreamSynchronize(st_));
      o += nc * nc;
    }
  }
}

size_t Solver::cone_blocks_size() const {
  size_t t = 0;
  for (const ConeDesc& C : cones_)
    t += C.kind == CONE_POS ? (size_t)C.n : (C.kind == CONE_FREE ? 0 : (size_t)C.n * C.n);
  return t;
}

void Solver::cone_corr0(double smu, double* r) {
  set_scalar(S_SMU, smu);
  launch_cone_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, NV_FR);
  launch_psd_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), d_psdlist_.as<int>(), npsd_, NV_FR);
  launch_bpsd_corr0(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), bpsd(), NV_FR);
  HIPCHK(hipMemcpyAsync(r, idev_.nv + (size_t)NV_FR * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
}

void Solver::cone_corr(const double* dp, const double* dd, double smu, double* r) {
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_DPA * n_, dp, n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_DDA * n_, dd, n_ * 8, hipMemcpyHostToDevice, st_));
  set_scalar(S_SMU, smu);
  launch_cone_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, NV_FR, NV_DPA, NV_DDA);
  launch_psd_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), d_psdlist_.as<int>(), npsd_, NV_FR, NV_DPA,
                  NV_DDA);
  launch_bpsd_corr(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), bpsd(), NV_FR, NV_DPA, NV_DDA);
  HIPCHK(hipMemcpyAsync(r, idev_.nv + (size_t)NV_FR * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
}

Solver::KKTOut Solver::kkt_solve(double* x, double* y, const double* f, const double* g, bool warm, double ftol,
                                 double gtol, int irmin, bool stops) {
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_FR * n_, f, n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.mv + (size_t)MV_DG * m_, g, m_ * 8, hipMemcpyHostToDevice, st_));
  if (warm) {
    HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_DP * n_, x, n_ * 8, hipMemcpyHostToDevice, st_));
    HIPCHK(hipMemcpyAsync(idev_.mv + (size_t)MV_DY * m_, y, m_ * 8, hipMemcpyHostToDevice, st_));
  }
  KKTOut o = solvekkt(NV_DP, MV_DY, NV_FR, MV_DG, warm, ftol, gtol, irmin, stops);
  HIPCHK(hipMemcpyAsync(x, idev_.nv + (size_t)NV_DP * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(y, idev_.mv + (size_t)MV_DY * m_, m_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return o;
}

double Solver::cone_maxsteps(const double* dp, const double* dd, double step_frac) {
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_DP * n_, dp, n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_DD * n_, dd, n_ * 8, hipMemcpyHostToDevice, st_));
  return maxsteps(NV_DP, NV_DD, step_frac);
}

// the step into S_TAU on the device (no read-back)
void Solver::launch_maxsteps(int kdp, int kdd, double step_frac) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_MAXSTEP;
    op.a[0] = kdp, op.a[1] = kdd;
    op.c[0] = K(step_frac);
    op.solo_only = nbpsd_ > 0;
    op.solo = [this, kdp, kdd, step_frac] { launch_maxsteps(kdp, kdd, step_frac); };
    gang_submit(&op);
    return;
  }
  launch_cone_maxstep(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), ncone_, d_cone_ofs_.as<int>(), 1, kdp,
                      kdd, step_frac, d_coneout_.as<double>(), d_psdlist_.as<int>(), npsd_, bpsd());
}

double Solver::maxsteps(int kdp, int kdd, double step_frac) {
  launch_maxsteps(kdp, kdd, step_frac);
  sync_scalars();
  return hsc_[S_TAU];
}

// ------------------------------------------------------------------------------------------
// starting point (solver/ipm.jl:321-429)
// ------------------------------------------------------------------------------------------

static void host_identity(double* x, const Cone& c, i64 n) {
  if (c.kind == CONE_POS) {
    for (i64 i = 0; i < n; ++i) x[i] = 1.0;
  } else if (c.kind == CONE_FREE) {
    for (i64 i = 0; i < n; ++i) x[i] = 0.0;
  } else if (c.kind == CONE_SOC) {
    for (i64 i = 0; i < n; ++i) x[i] = 0.0;
    x[0] = std::sqrt(2.0);
  } else if (c.kind == CONE_PSD) {  // svec(I) (sdp.jl:92-111)
    i64 d = cone_degree(c, n), k = 0;
    for (i64 i = 0; i < n; ++i) x[i] = 0.0;
    for (i64 j = 0; j < d; ++j) {
      x[k] = 1.0;
      k += d - j;
    }
  } else {  // exp.jl:15-25, pow.jl:27-33
    tdc::identity(c.kind, c.param, x);
  }
}

// X⁻¹ in svec form through the Cholesky factor of X (sdp.jl:666-678); false if X is not
// positive definite
static bool psd_shadow(double* s, const double* x, i64 n, i64 d) {
  const double r2 = std::sqrt(2.0);
  std::vector<double> L(d * d, 0.0), Z(d * d, 0.0);
  for (i64 j = 0, k = 0; j < d; ++j)
    for (i64 i = j; i < d; ++i, ++k) L[i + j * d] = i == j ? x[k] : x[k] / r2;
  for (i64 j = 0; j < d; ++j) {
    double p = L[j + j * d];
    for (i64 k = 0; k < j; ++k) p -= L[j + k * d] * L[j + k * d];
    if (!(p > 0.0)) return false;
    L[j + j * d] = std::sqrt(p);
    for (i64 i = j + 1; i < d; ++i) {
      double a = L[i + j * d];
      for (i64 k = 0; k < j; ++k) a -= L[i + k * d] * L[j + k * d];
      L[i + j * d] = a / L[j + j * d];
    }
  }
  for (i64 c = 0; c < d; ++c) {  // Z = L⁻ᵀ L⁻¹ e_c
    std::vector<double> y(d, 0.0);
    for (i64 i = 0; i < d; ++i) {
      double a = i == c ? 1.0 : 0.0;
      for (i64 k = 0; k < i; ++k) a -= L[i + k * d] * y[k];
      y[i] = a / L[i + i * d];
    }
    for (i64 i = d - 1; i >= 0; --i) {
      double a = y[i];
      for (i64 k = i + 1; k < d; ++k) a -= L[k + i * d] * Z[k + c * d];
      Z[i + c * d] = a / L[i + i * d];
    }
  }
  for (i64 j = 0, k = 0; j < d; ++j)
    for (i64 i = j; i < d; ++i, ++k) s[k] = i == j ? Z[i + j * d] : r2 * Z[i + j * d];
  (void)n;
  return true;
}

// s = −f'(x) (dualshadow!); for the exponential and power cones, whose barriers are not
// self-concordant-dual to themselves, primal = true solves −f'(s) = x instead (primalshadow!)
static bool host_shadow(double* s, const double* x, const Cone& c, i64 n, bool primal) {
  if (c.kind == CONE_EXP || c.kind == CONE_POW) {
    double cache[tdc::CACHE] = {0.0};
    if (primal) tdc::primal_shadow(c.kind, c.param, s, x, cache);
    else tdc::dual_shadow(c.kind, c.param, s, x, cache);
    return true;
  }
  if (c.kind == CONE_POS) {
    for (i64 i = 0; i < n; ++i) s[i] = 1.0 / x[i];
    return true;
  }
  if (c.kind == CONE_FREE) {
    for (i64 i = 0; i < n; ++i) s[i] = 0.0;
    return true;
  }
  if (c.kind == CONE_PSD) return psd_shadow(s, x, n, cone_degree(c, n));
  // SOC: s = 2 Jx / det(x) with the compensated determinant (soc.jl:488-504)
  double xdet = x[0] * x[0];
  double corr = std::fma(x[0], x[0], -xdet);
  for (i64 i = 1; i < n; ++i) {
    double p = x[i] * x[i], e = std::fma(x[i], x[i], -p);
    double sum = xdet - p, bb = sum - xdet;
    double e2 = (xdet - (sum - bb)) + (-p - bb);
    xdet = sum;
    corr += -e + e2;
  }
  xdet = xdet + corr;
  bool ok = xdet > 0 && x[0] > 0;
  if (ok) {
    double gm = 2.0 / xdet;
    s[0] = gm * x[0];
    for (i64 i = 1; i < n; ++i) s[i] = -gm * x[i];
  }
  return ok;
}

static void bsm_gemv(const BSM& A, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(A.nrow(), 0.0);
  for (i64 v = 0; v < A.nvtx; ++v) {
    i64 nc = A.ncols(v);
    for (i64 e = A.xsrc[v]; e < A.xsrc[v + 1]; ++e) {
      i64 u = A.tgt[e], nr = A.nrows(u);
      for (i64 jl = 0; jl < nc; ++jl)
        for (i64 il = 0; il < nr; ++il) y[A.xrow[u] + il] += A.val[A.xblk[e] + jl * nr + il] * x[A.xcol[v] + jl];
    }
  }
}

void Solver::reinit(const double* p0, const double* d0, const double* y0) {
  std::vector<double> p(n_), d(n_), y(m_, 0.0);
  const i64 N = P_.B.nvtx;
  if (!p0 && !d0) {
    for (i64 v = 0; v < N; ++v) {
      host_identity(&p[P_.B.xcol[v]], P_.K[v], P_.B.ncols(v));
      host_identity(&d[P_.B.xcol[v]], P_.K[v], P_.B.ncols(v));
    }
    std::vector<double> z, w;
    bsm_gemv(P_.B, p, z);
    bsm_gemv(P_.Q, p, w);
    double np = norm2(p), nz = norm2(z), nw = norm2(w);
    double sp = nz > EPS * np ? std::max(1.0, norm2(P_.g) / nz) : 1.0;
    double sd = np > EPS ? std::max(1.0, (norm2(P_.f) + sp * nw) / np) : 1.0;
    for (auto& x : p) x *= sp;
    for (auto& x : d) x *= sd;
    if (y0)
      for (i64 i = 0; i < m_; ++i) y[i] = y0[P_.P1[i]] / yscl_[i];
  } else {
    if (p0)
      for (i64 j = 0; j < n_; ++j) p[j] = p0[P_.P2[j]];
    if (d0)
      for (i64 j = 0; j < n_; ++j) d[j] = d0[P_.P2[j]];
    for (i64 v = 0; v < N; ++v) {
      i64 a = P_.B.xcol[v], nc = P_.B.ncols(v);
      if (!d0) host_shadow(&d[a], &p[a], P_.K[v], nc, false);
      else if (!p0) host_shadow(&p[a], &d[a], P_.K[v], nc, true);
    }
    if (y0)
      for (i64 i = 0; i < m_; ++i) y[i] = y0[P_.P1[i]];
    for (i64 j = 0; j < n_; ++j) {
      p[j] /= pscl_[j];
      d[j] *= pscl_[j];
    }
    for (i64 i = 0; i < m_; ++i) y[i] /= yscl_[i];
  }
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_P * n_, p.data(), n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_D * n_, d.data(), n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.mv + (size_t)MV_Y * m_, y.data(), m_ * 8, hipMemcpyHostToDevice, st_));
  launch_td_initcache(st_, d_inst_.as<InstDev>(), d_cones_.as<ConeDesc>(), d_tdlist_.as<int>(), ntd_);  // initcache!
  HIPCHK(hipStreamSynchronize(st_));
  hist_.clear();
  hist_fp32_.clear();
  step_pending_ = false;
  nH_ = norm2(P_.Q.val);
  status_ = CONTINUE;
  mixed_off_ = false;
}

// ------------------------------------------------------------------------------------------
// KKT: Uzawa refinement + dual CG (uzawa.jl:319-507, cg.jl:69-167)
// ------------------------------------------------------------------------------------------

// whether cg below takes a step, from the same read-back values
bool Solver::cg_runs(double atol, int itmax, bool stops) const {
  const bool stop = stops && hsc_[S_R9] / (1.0 + sg_) <= set_.feas_tol;
  return itmax > 0 && std::sqrt(hsc_[S_R8]) > atol && !stop;
}

int Solver::cg(int kx, int ky, int ku, double atol, int itmax, bool stops, int& status) {
  // cg.jl:69-167 with p = MV_CP, w = NV_CW. On entry hsc_[S_R8] holds ‖u‖² and hsc_[S_R9] its
  // scaled stop norm (read back with the last Uzawa residual)
  int niter = 0;
  double nu2 = hsc_[S_R8], nu = std::sqrt(nu2), np2 = nu2;
  double chist0 = nu;
  std::vector<double>& ch = cghist_;
  ch.assign(1, chist0);
  status = 1;  // CG_SOLVED
  auto stopf = [&](double scaled) { return stops && scaled / (1.0 + sg_) <= set_.feas_tol; };
  if (nu > atol && !stopf(hsc_[S_R9])) {
    status = 3;  // CG_ITMAX
    if (itmax > 0) {
      // y ← 0, p ← u (y is read only after a step, so it is cleared only when one is taken)
      lin(LinSet(), LinSet()
                        .add(Lin{ku, ku, ku, MV_CP, K(1.0), K(0.0), K(0.0)})
                        .add(Lin{ku, ku, ku, ky, K(0.0), K(0.0), K(0.0)}));
      status = 0;
      set_scalar2(S_CG_NU2, nu2, S_CG_NP2, np2);
      while (status == 0) {
        niter += 1;
        spmv(1, MV_CP, NV_CW, K(1.0), K(0.0));  // w = Bᵀ p
        lsolve(NV_CW);
        dots(DotBatch().add(0, NV_CW, NV_CW, 0, S_CG_EP2, 0));
        sync_scalars();
        double ep2 = hsc_[S_CG_EP2];
        if (ep2 <= EPS * np2) {
          status = 2;  // CG_NUMERICAL_FAILURE
        } else {
          usolve(NV_CW);
          cg_scalars(0);  // α = ν²/ε²
          // x += α w, y += α p
          lin2(Lin{NV_CW, kx, kx, kx, KS(1.0, S_CG_ALPHA), K(1.0), K(0.0)},
               Lin{MV_CP, ky, ky, ky, KS(1.0, S_CG_ALPHA), K(1.0), K(0.0)});
          spmv(0, NV_CW, ku, KS(-1.0, S_CG_ALPHA), K(1.0));  // u −= α B w
          DotBatch nb;
          nb.add(1, ku, ku, 0, S_NORM0, 0);
          if (stops) nb.add(1, ku, MV_YSCL, 1, S_R9, 1);
          dots(nb);
          sync_scalars();
          double pnu2 = nu2;
          nu2 = hsc_[S_NORM0];
          nu = std::sqrt(nu2);
          ch.push_back(nu);
          if (nu <= atol || 1.0 + nu <= 1.0 || stopf(hsc_[S_R9])) {
            status = 1;
            set_scalar(S_CG_NU2, nu2);
          } else if (niter >= itmax) {
            status = 3;
          } else {
            double beta = nu2 / pnu2;
            cg_scalars(1);  // β, ν², np² on device
            axpby(1, ku, MV_CP, MV_CP, K(1.0), KS(1.0, S_CG_BETA));
            np2 = nu2 + beta * beta * np2;
          }
        }
      }
    }
  }
  return niter;
}

Solver::KKTOut Solver::solvekkt(int kx, int ky, int kf, int kg, bool warm, double ftol, double gtol, int irmin,
                                bool stops) {
  const double stall = set_.refine_stall_tol;
  const int irmax = set_.refine_max_iter, cgmax = set_.newton_max_iter;
  int nir = 0;
  std::vector<double> hist;
  // the norms of a residual: ‖KF‖ (and its scaled stop), and ‖KG‖² (and its scaled stop) for the
  // dual CG that may follow, all in one read-back
  auto norms = [&]() {
    DotBatch b;
    b.add(0, NV_KF, NV_KF, 0, S_R10, 1);
    if (stops) b.add(0, NV_KF, NV_PSCL, 1, S_R11, 1);
    b.add(1, MV_KG, MV_KG, 0, S_R8, 0);
    if (stops) b.add(1, MV_KG, MV_YSCL, 1, S_R9, 1);
    dots(b);
    sync_scalars();
  };
  auto residual = [&](bool read) {
    // KF = f − Q x + Bᵀ y (one pass over the rows), then − (cone blocks) x ; KG = g − B x
    const SpmvN rn{kx, ky, kf, -1, -1, NV_KF, K(-1.0), K(1.0), K(1.0), K(0.0), K(0.0)};
    const SpmvM rm{kx, kg, MV_KG, -1, K(-1.0), K(1.0), K(0.0), 0};
    spmv_nm(2, &rn, &rm);
    if (read && user_h_) {  // the caller's H holds the cone blocks too
      norms();
      return;
    }
    if (user_h_) return;
    cone_matvec_all(kx, NV_KF, K(-1.0));
    if (read) norms();
  };
  // with irmin ≥ 1 and irmax ≥ 1 the starting residual decides nothing (and its norm enters no
  // later test), so it is not read back
  const bool read0 = irmin == 0 || irmax == 0;
  // a cold start that refines at least once starts from x = y = 0, whose residual is (f, g): the
  // first refinement step then reads f and g directly and sets x and y instead of adding to them
  const bool cold1 = !warm && !read0;
  if (warm) {
    residual(read0);
  } else if (!cold1) {
    lin2(Lin{kf, kf, kf, NV_KF, K(1.0), K(0.0), K(0.0)}, Lin{kg, kg, kg, MV_KG, K(1.0), K(0.0), K(0.0)});
    lin2(Lin{kx, kx, kx, kx, K(0.0), K(0.0), K(0.0)}, Lin{ky, ky, ky, ky, K(0.0), K(0.0), K(0.0)});
    if (read0) norms();
  }
  auto fstop = [&]() { return stops && hsc_[S_R11] / (1.0 + sf_) <= set_.feas_tol; };
  double fres = read0 ? hsc_[S_R10] : NaN, fprv = fres;
  hist.push_back(fres);
  int irstat = KKT_CONTINUE;
  if (read0) {
    if (nir >= irmin && (fres <= ftol || fstop())) irstat = KKT_SOLVED;
    else if (nir >= irmax) irstat = KKT_IRMAX;
  }
  while (irstat == KKT_CONTINUE) {
    nir += 1;
    const bool first = cold1 && nir == 1;  // δf = f, δg = g, x = y = 0
    const SpmvN dx{-1, first ? kg : MV_KG, first ? kf : NV_KF, -1, -1, NV_KX, K(0.0), K(1.0), KS(1.0, S_DELTA),
                   K(0.0), K(0.0)};
    spmv_nm(1, &dx, nullptr);  // δx = δ δf + Bᵀ δg
    lsolve(NV_KX);
    usolve(NV_KX);
    // δg −= B δx, then y += α δg ; x += δx
    const SpmvN xu{-1, -1, NV_KX, kx, -1, kx, K(0.0), K(0.0), K(1.0), first ? K(0.0) : K(1.0), K(0.0)};
    const SpmvM gu{NV_KX, first ? kg : MV_KG, MV_KG, ky, K(-1.0), K(1.0), KS(1.0, S_ALPHA), first ? 1 : 0};
    spmv_nm(1, &xu, &gu);
    residual(true);
    fres = hsc_[S_R10];
    hist.push_back(fres);
    if (nir >= irmin && (fres <= ftol || fstop())) irstat = KKT_SOLVED;
    else if (nir > 1 && fres > (1 - stall) * fprv) irstat = KKT_STAGNATED;
    else if (nir >= irmax) irstat = KKT_IRMAX;
    else irstat = KKT_CONTINUE;
    fprv = fres;
  }
  // the dual CG on the last residual: y −= α δg, CG, y += α δy + α δg. When CG takes no step
  // (δy = 0, δg unchanged) that is y itself, and nothing is launched.
  const bool cgruns = cg_runs(gtol, cgmax, stops);
  if (cgruns) axpby(1, MV_KG, ky, ky, KS(-1.0, S_ALPHA), K(1.0));
  int cgstat = 1;
  int ncg = cg(kx, MV_KY, MV_KG, gtol, cgmax, stops, cgstat);
  check_abort();
  if (cgruns) axpbypcz(1, MV_KY, MV_KG, ky, ky, KS(1.0, S_ALPHA), KS(1.0, S_ALPHA), K(1.0));
  KKTOut o;
  o.ncg = ncg;
  o.nir = nir;
  if (cgstat == 2) o.status = KKT_NUMERICAL_FAILURE;
  else if (irstat == KKT_STAGNATED) o.status = KKT_STAGNATED;
  else if (irstat == KKT_IRMAX || cgstat == 3) o.status = KKT_ITMAX;
  else o.status = KKT_SOLVED;
  // kktwindow (uzawa.jl:199-227)
  o.dmin = o.dmax = NaN;
  double delta = hsc_[S_DELTA];
  if (nir >= 1) {
    double logd = std::log(delta), logfres = std::log(hist[1]), loggtol = std::log(gtol),
           logftol = std::log(ftol);
    int i = 0;
    double logdmax = -std::numeric_limits<double>::infinity();
    double logdmin = logd + logfres - logftol;
    while (logdmin > logdmax && i < ncg + 1) {
      i += 1;
      double loggres = std::log(cghist_[i - 1]);
      logdmax = logd + (loggtol - loggres) / i;
    }
    o.dmin = std::exp(logdmin);
    o.dmax = std::exp(logdmax);
  }
  return o;
}

// ------------------------------------------------------------------------------------------
// differentiation (src/IPM/src/ipm/differential.jl): frule! and rrule! on the solver's last
// iterate, its scaling H and a fresh factor, every vector on the device
// ------------------------------------------------------------------------------------------

// dualshadow! (cone.jl:556-566) of vector kind kx into ky, every cone kind
void Solver::dshadow(int kx, int ky) {
  const InstDev* I = d_inst_.as<InstDev>();
  launch_cone_dshadow(st_, I, d_cones_.as<ConeDesc>(), ncone_, kx, ky);
  launch_psd_dshadow(st_, I, d_cones_.as<ConeDesc>(), d_psdlist_.as<int>(), npsd_, kx, ky);
  launch_bpsd_dshadow(st_, I, d_cones_.as<ConeDesc>(), bpsd(), kx, ky);
}

// y += a (H − Q) x: the cone blocks of H
void Solver::cone_matvec_all(int kx, int ky, Coef a) {
  if (ganged()) {
    GOp& op = *new GOp;
    op.kind = GK_MATVEC;
    op.a[0] = kx, op.a[1] = ky;
    op.c[0] = a;
    op.solo_only = nbpsd_ > 0;
    op.solo = [this, kx, ky, a] { cone_matvec_all(kx, ky, a); };
    gang_submit(&op);
    return;
  }
  const InstDev* I = d_inst_.as<InstDev>();
  launch_cone_matvec(st_, I, d_cones_.as<ConeDesc>(), ncone_, kx, ky, a);
  launch_cone_matvec_dense(st_, I, d_cones_.as<ConeDesc>(), d_denselist_.as<int>(), ndense_, dense_nmax_, kx, ky, a);
  launch_bpsd_matvec(st_, I, d_cones_.as<ConeDesc>(), bpsd(), kx, ky, a);
}

// the factor of δH + BᵀB at getaug's δ and the forcing tolerances of differential.jl:5-16
void Solver::diff_setup(double& ftol, double& gtol) {
  if (hist_.empty()) throw std::logic_error("differentiation needs a solved (stepped) solver");
  use_user_h(false);
  const double delta = getaug();
  set_scalar(S_ALPHA, 1.0 / delta);
  double rho = 0.0;
  if (!factor(delta, rho)) throw std::runtime_error("initkkt! failed");
  const double mu0 = hist_.front().mu;
  const double tol = (mu0 == 0.0 || !std::isfinite(mu0)) ? FORCING_FRAC * set_.gap_tol
                                                          : FORCING_FRAC * hist_.back().mu / mu0;
  ftol = tol * (1 + nf_);
  gtol = tol * (1 + ng_);
}

// frule! (differential.jl:5-75): the tangent (Δp, Δy, Δd) of the solution for seeds (Δf, Δg, Δμ)
void Solver::frule(const double* df, const double* dg, double dmu, double* dp, double* dy, double* dd) {
  double ftol, gtol;
  diff_setup(ftol, gtol);
  std::vector<double> fx(n_), gx(m_), px(n_), yx(m_), dx(n_);
  for (i64 j = 0; j < n_; ++j) fx[j] = df[P_.P2[j]] * pscl_[j];
  for (i64 i = 0; i < m_; ++i) gx[i] = dg[P_.P1[i]] * yscl_[i];
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_FR * n_, fx.data(), n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.mv + (size_t)MV_DG * m_, gx.data(), m_ * 8, hipMemcpyHostToDevice, st_));
  dshadow(NV_P, NV_T1);  // d*(p)
  // Δd = Δμ d* (into T2), f-side right-hand side Δf + Δμ d*
  lin(LinSet()
          .add(Lin{NV_T1, NV_T1, NV_T1, NV_T2, K(dmu), K(0.0), K(0.0)})
          .add(Lin{NV_T2, NV_FR, NV_FR, NV_FR, K(1.0), K(1.0), K(0.0)}),
      LinSet());
  solvekkt(NV_DP, MV_DY, NV_FR, MV_DG, false, ftol, gtol, 1, false);
  cone_matvec_all(NV_DP, NV_T2, K(-1.0));  // Δd = Δμ d* − (H − Q) Δp
  HIPCHK(hipMemcpyAsync(px.data(), idev_.nv + (size_t)NV_DP * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(yx.data(), idev_.mv + (size_t)MV_DY * m_, m_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(dx.data(), idev_.nv + (size_t)NV_T2 * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (i64 j = 0; j < n_; ++j) {
    dp[P_.P2[j]] = px[j] * pscl_[j];
    dd[P_.P2[j]] = dx[j] / pscl_[j];
  }
  for (i64 i = 0; i < m_; ++i) dy[P_.P1[i]] = yx[i] * yscl_[i];
}

// rrule! (differential.jl:252-334): the cotangent (Δf, Δg, dμ) of the seeds (Δp, Δy, Δd)
double Solver::rrule(const double* dp, const double* dy, const double* dd, double* df, double* dg) {
  double ftol, gtol;
  diff_setup(ftol, gtol);
  std::vector<double> px(n_), dx(n_), yx(m_), fx(n_), gx(m_);
  for (i64 j = 0; j < n_; ++j) {
    px[j] = dp[P_.P2[j]] * pscl_[j];
    dx[j] = dd[P_.P2[j]] / pscl_[j];
  }
  for (i64 i = 0; i < m_; ++i) yx[i] = -dy[P_.P1[i]] * yscl_[i];
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_FR * n_, px.data(), n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.nv + (size_t)NV_T2 * n_, dx.data(), n_ * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipMemcpyAsync(idev_.mv + (size_t)MV_DG * m_, yx.data(), m_ * 8, hipMemcpyHostToDevice, st_));
  cone_matvec_all(NV_T2, NV_FR, K(-1.0));  // Δp −= (H − Q) Δd
  solvekkt(NV_DP, MV_DY, NV_FR, MV_DG, false, ftol, gtol, 1, false);
  lin(LinSet().add(Lin{NV_DP, NV_T2, NV_T2, NV_T2, K(1.0), K(1.0), K(0.0)}), LinSet());  // λd = Δd + λp
  dshadow(NV_P, NV_T1);
  dots(DotBatch().add(0, NV_T2, NV_T1, 0, S_TMP3, 0));
  sync_scalars();
  const double dmu = hsc_[S_TMP3];
  HIPCHK(hipMemcpyAsync(fx.data(), idev_.nv + (size_t)NV_DP * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(gx.data(), idev_.mv + (size_t)MV_DY * m_, m_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (i64 j = 0; j < n_; ++j) df[P_.P2[j]] = fx[j] * pscl_[j];
  for (i64 i = 0; i < m_; ++i) dg[P_.P1[i]] = -gx[i] * yscl_[i];
  return dmu;
}

// ------------------------------------------------------------------------------------------
// step! (solver/ipm.jl:577-774)
// ------------------------------------------------------------------------------------------

double Solver::getaug() {
  // solver.jl:40-52, history.jl:9-25
  if (nB_ == 0.0 || nH_ == 0.0) return 1.0;
  if (hist_.empty()) return set_.aug_tol * nB_ * nB_ / nH_;
  const HistRow& h = hist_.back();
  double delta = h.delta;
  // mixed precision: after an FP32-factor iteration a second refinement pass reflects the FP32
  // rounding floor rather than too large a δ, so only the window applies (the window then keeps
  // δ large enough for the FP32 factorization to stay positive definite)
  const bool fp32 = !hist_fp32_.empty() && hist_fp32_.back();
  if (h.ppass > 1 && !fp32) {
    if (std::isfinite(h.dmin)) delta = std::max(delta, h.dmin);
    delta *= 10;
  } else if (std::isfinite(h.dmin) && std::isfinite(h.dmax) && h.dmin > 0 && h.dmax > 0) {
    delta = std::sqrt(h.dmin) * std::sqrt(h.dmax);
  }
  return delta;
}

bool Solver::isstalled() const {
  const int window = 5;
  double tol = set_.stall_tol;
  int n = (int)hist_.size();
  if (n <= window) return false;
  for (int i = n - window; i < n; ++i) {
    const HistRow &h = hist_[i], &hp = hist_[i - 1];
    if (h.mu < (1 - tol) * hp.mu || h.pres < (1 - tol) * hp.pres || h.dres < (1 - tol) * hp.dres) return false;
  }
  return true;
}

bool Solver::isopt(double mu, double mus, double pobj, double dobj, double pres, double dres, double fac) const {
  double feas = fac * set_.feas_tol, gap = fac * set_.gap_tol;
  if (!(pres < feas && dres < feas)) return false;
  double tol = gap * std::max(1.0, std::min(std::fabs(pobj), std::fabs(dobj)));
  if (mu_t_ > 0 && nu_ > 0) return nu_ * (mu - mu_t_ + mu_t_ * std::log1p(mus * mu_t_ - 1.0)) <= tol;
  return pobj - dobj < tol;
}

// statistics of the current iterate for the termination and certificate tests
void Solver::compute_stats() {
  // R0 gy, R1 ‖y‖, R2 ‖Qp‖, R3 ‖d‖, R4 ‖Qp − d − Bᵀy‖, R5 fp, R6 ‖p‖, R7 ‖Bp‖
  // T2 = Qp − d − Bᵀy ; T0 = B p
  const SpmvN sn{-1, MV_Y, NV_QP, NV_D, -1, NV_T2, K(0.0), K(-1.0), K(1.0), K(-1.0), K(0.0)};
  const SpmvM sm{NV_P, -1, MV_T0, -1, K(1.0), K(0.0), K(0.0), 0};
  spmv_nm(1, &sn, &sm);
  DotBatch b;
  b.add(1, MV_G, MV_Y, 0, S_R0, 0).add(1, MV_Y, MV_Y, 0, S_R1, 1).add(0, NV_QP, NV_QP, 0, S_R2, 1);
  b.add(0, NV_D, NV_D, 0, S_R3, 1).add(0, NV_T2, NV_T2, 0, S_R4, 1).add(0, NV_F, NV_P, 0, S_R5, 0);
  b.add(0, NV_P, NV_P, 0, S_R6, 1).add(1, MV_T0, MV_T0, 0, S_R7, 1);
  dots(b);
  sync_scalars();
}

// solver/ipm.jl:502-517
bool Solver::isprimalinfeasible(double rtol, double atol) {
  double gy = hsc_[S_R0], ny = hsc_[S_R1];
  bool flag = gy > atol * ny * (1 + ng_);
  if (flag) flag = std::max(hsc_[S_R2], hsc_[S_R4]) < rtol * gy * (1 + hsc_[S_R3] / ny);
  return flag;
}

// solver/ipm.jl:528-540
bool Solver::isdualinfeasible(double rtol, double atol) {
  double fp = hsc_[S_R5], np = hsc_[S_R6];
  bool flag = fp > atol * np * (1 + nf_);
  if (flag) flag = std::max(hsc_[S_R7], hsc_[S_R2]) < rtol * std::fabs(fp);
  return flag;
}

int Solver::nearstatus(int status, double mu, double mus, double pobj, double dobj, double pres, double dres) {
  double f = set_.near_factor;
  if (isopt(mu, mus, pobj, dobj, pres, dres, f)) return NEAR_OPTIMAL;
  if (isprimalinfeasible(f * set_.infeas_rel, set_.infeas_abs)) return NEAR_PRIMAL_INFEASIBLE;
  if (isdualinfeasible(f * set_.infeas_rel, set_.infeas_abs)) return NEAR_DUAL_INFEASIBLE;
  return status;
}

// The last update's step length, from a read-back made after it: into its history row, and a
// collapsing step with an FP32 factor (its direction was not accurate enough) ends mixed precision
void Solver::settle_step() {
  if (!step_pending_) return;
  step_pending_ = false;
  const double stp = hsc_[S_TAU];
  if (!hist_.empty()) hist_.back().step = stp;
  if (step_fp32_ && stp < MIXED_MIN_STEP) mixed_off_ = true;
}

const std::vector<HistRow>& Solver::history() {
  if (step_pending_) {
    sync_scalars();
    settle_step();
  }
  return hist_;
}

int Solver::step() {
  use_user_h(false);
  int status = CONTINUE;
  HistRow row{NaN, NaN, NaN, NaN, 0.0, 0, 0, KKT_SOLVED, 0, 0, KKT_SOLVED, NaN, NaN, NaN, 0.0, NaN, NaN};
  const double mut = mu_t_;
  double delta = getaug();
  row.delta = delta;
  set_scalar2(S_DELTA, delta, S_ALPHA, 1.0 / delta);
  // pᵀd, scale!, residuals!
  cone_scale_norm();
  {
    // Qp ; Δf = d + f − Qp + Bᵀy ; Δg = g − Bp
    SpmvN fn{NV_P, MV_Y, NV_D, NV_F, -1, NV_DF, K(-1.0), K(1.0), K(1.0), K(1.0), K(0.0)};
    fn.kyq = NV_QP;
    const SpmvM gm{NV_P, MV_G, MV_DG, -1, K(-1.0), K(1.0), K(0.0), 0};
    spmv_nm(2, &fn, &gm);
  }
  {
    // the residual norms and objectives, and the statistics of the termination and certificate
    // tests (compute_stats) of the same iterate, in one launch and one read-back
    DotBatch b;
    b.add(0, NV_P, NV_D, 0, S_TMP0, 0).add(1, MV_DG, MV_YSCL, 1, S_NORM0, 1).add(0, NV_DF, NV_PSCL, 1, S_NORM1, 1);
    b.add(0, NV_P, NV_QP, 0, S_NORM2, 0).add(0, NV_F, NV_P, 0, S_NORM3, 0).add(1, MV_G, MV_Y, 0, S_NORM4, 0);
    b.add(1, MV_G, MV_Y, 0, S_R0, 0).add(1, MV_Y, MV_Y, 0, S_R1, 1).add(0, NV_QP, NV_QP, 0, S_R2, 1);
    b.add(0, NV_D, NV_D, 0, S_R3, 1).add(0, NV_F, NV_DF, 4, S_R4, 1).add(0, NV_F, NV_P, 0, S_R5, 0);
    b.add(0, NV_P, NV_P, 0, S_R6, 1).add(1, MV_G, MV_DG, 4, S_R7, 1);
    dots(b);
  }
  sync_scalars();
  settle_step();
  double pd = hsc_[S_TMP0];
  bool flag = hsc_[S_FLAGOK] != 0.0;
  double spsd = hsc_[S_SPSD];
  nH_ = std::sqrt(qoff2_ + hsc_[S_TMP1]);
  if (!flag) {
    status = NUMERICAL_FAILURE;
  } else {
    double nu1 = std::max(nu_, 1);
    double mu = pd / nu1, mus = spsd / nu1, chi = mu * mus - 1;
    cur_mu_ = mu;
    double pres = hsc_[S_NORM0] / (1 + sg_), dres = hsc_[S_NORM1] / (1 + sf_);
    double pQp = hsc_[S_NORM2];
    double pobj = pQp / 2 - hsc_[S_NORM3];
    double dobj = hsc_[S_NORM4] - pQp / 2;
    row.mu = mu;
    row.chi = chi;
    row.pres = pres;
    row.dres = dres;
    row.pobj = pobj;
    row.dobj = dobj;
    // S_R0..S_R7 came with the read-back above: ‖f − Δf‖ = ‖Qp − d − Bᵀy‖, ‖g − Δg‖ = ‖Bp‖
    if (isopt(mu, mus, pobj, dobj, pres, dres, 1.0)) {
      status = OPTIMAL;
    } else if (isprimalinfeasible(set_.infeas_rel, set_.infeas_abs)) {
      status = PRIMAL_INFEASIBLE;
    } else if (isdualinfeasible(set_.infeas_rel, set_.infeas_abs)) {
      status = DUAL_INFEASIBLE;
    } else if ((int)hist_.size() >= set_.max_iter) {
      status = nearstatus(ITERATION_LIMIT, mu, mus, pobj, dobj, pres, dres);
    } else if (isstalled()) {
      status = nearstatus(STALLED, mu, mus, pobj, dobj, pres, dres);
    } else if (nu_ == 0) {
      double rho;
      bool ok = factor(delta, rho);
      row.rho = rho;
      if (!ok) {
        status = nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres);
      } else {
        // solveexact! (solver/ipm.jl:284-315)
        axpby(0, NV_DF, NV_DF, NV_FR, K(1.0), K(0.0));
        KKTOut o = solvekkt(NV_DPA, MV_DYA, NV_FR, MV_DG, false, EPS, EPS, 0, true);
        row.piter = o.ncg;
        row.ppass = o.nir;
        row.pstat = o.status;
        row.dmin = o.dmin;
        row.dmax = o.dmax;
        axpby(0, NV_DPA, NV_P, NV_P, K(1.0), K(1.0));
        axpby(1, MV_DYA, MV_Y, MV_Y, K(1.0), K(1.0));
      }
    } else if (!(mu > 0)) {
      status = nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres);
    } else {
      double rho;
      bool ok = factor(delta, rho);
      row.rho = rho;
      if (!ok) {
        status = nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres);
      } else {
        // attempt 0 with the factor of this iteration; with mixed precision a KKT solve that the
        // FP32 factor cannot refine to tolerance is repeated once on an FP64 factor, which then
        // stays for the rest of the solve (μ only decreases)
        for (int attempt = 0;; ++attempt) {
          double mu1 = hist_.empty() ? mu : hist_[0].mu;
          double tol;
          if (mut > 0) tol = FORCING_FRAC * std::min(std::fabs(mu - mut) / std::fabs(mu1 - mut), 1.0);
          else tol = FORCING_FRAC * mu / mu1;
          double ftol = tol * (1 + nf_), gtol = tol * (1 + ng_);
          // predictor (solver/ipm.jl:111-163)
          if (mut > 0) {
            set_scalar(S_SMU, mut);
            corr0_op(NV_FR);
            axpby(0, NV_DF, NV_FR, NV_FR, K(1.0), K(1.0));
          } else {
            lin(LinSet().add(Lin{NV_DF, NV_D, NV_D, NV_FR, K(1.0), K(-1.0), K(0.0)}), LinSet());  // Δf − d
          }
          KKTOut po = solvekkt(NV_DPA, MV_DYA, NV_FR, MV_DG, false, ftol, gtol, 1, false);
          // Δda = Q Δpa − Bᵀ Δya − Δf
          const SpmvN da{NV_DPA, MV_DYA, NV_DF, -1, -1, NV_DDA, K(1.0), K(-1.0), K(-1.0), K(0.0), K(0.0)};
          spmv_nm(2, &da, nullptr);
          zerofree(NV_DDA);
          // corrector (solver/ipm.jl:181-272)
          // τa stays on the device: the dot reads it there, and one read-back brings μa home
          launch_maxsteps(NV_DPA, NV_DDA, 1.0);
          dots(DotBatch().add(0, NV_P, NV_D, 3, S_TMP2, 0, NV_DPA, NV_DDA));
          smu((double)nu_, mu, mut);  // σμ into S_SMU
          corr_op(NV_FR, NV_DPA, NV_DDA);
          // corrector right-hand side Δf + fr, warm start (Δpa, Δya)
          lin(LinSet()
                  .add(Lin{NV_DF, NV_FR, NV_FR, NV_FR, K(1.0), K(1.0), K(0.0)})
                  .add(Lin{NV_DPA, NV_DPA, NV_DPA, NV_DP, K(1.0), K(0.0), K(0.0)}),
              LinSet().add(Lin{MV_DYA, MV_DYA, MV_DYA, MV_DY, K(1.0), K(0.0), K(0.0)}));
          KKTOut co = solvekkt(NV_DP, MV_DY, NV_FR, MV_DG, true, ftol, gtol, 1, false);
          const SpmvN dd{NV_DP, MV_DY, NV_DF, -1, -1, NV_DD, K(1.0), K(-1.0), K(-1.0), K(0.0), K(0.0)};
          spmv_nm(2, &dd, nullptr);
          zerofree(NV_DD);
          row.piter = po.ncg;
          row.ppass = po.nir;
          row.pstat = po.status;
          row.dmin = po.dmin;
          row.dmax = po.dmax;
          row.citer = co.ncg;
          row.cpass = co.nir;
          row.cstat = co.status;
          if (attempt == 0 && fac_fp32_ && (po.status != KKT_SOLVED || co.status != KKT_SOLVED)) {
            mixed_off_ = true;
            ok = factor(delta, rho);
            row.rho = rho;
            if (ok) continue;
          }
          // once the FP32 factor needs a second refinement pass (predictor or corrector), FP64
          // for the rest of the solve
          if (fac_fp32_ && (po.nir >= 2 || co.nir >= 2)) mixed_off_ = true;
          if (po.status != KKT_SOLVED || co.status != KKT_SOLVED) {
            compute_stats();
            status = nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres);
          } else {
            // the step length stays on the device (S_TAU) for the update; the host reads it with
            // the next read-back (settle_step)
            launch_maxsteps(NV_DP, NV_DD, set_.step_frac);
            lin(LinSet()
                    .add(Lin{NV_DP, NV_P, NV_P, NV_P, KS(1.0, S_TAU), K(1.0), K(0.0)})
                    .add(Lin{NV_DD, NV_D, NV_D, NV_D, KS(1.0, S_TAU), K(1.0), K(0.0)}),
                LinSet().add(Lin{MV_DY, MV_Y, MV_Y, MV_Y, KS(1.0, S_TAU), K(1.0), K(0.0)}));
            step_pending_ = true;
            step_fp32_ = fac_fp32_;
          }
          break;
        }
      }
    }
  }
  hist_.push_back(row);
  hist_fp32_.push_back(fac_fp32_ ? 1 : 0);
  status_ = status;
  return status;
}

int Solver::solve() {
  int status = CONTINUE;
  while (status == CONTINUE) status = step();
  return status;
}

int Solver::solve_in(Gang& g) { return g.run(this); }

// result (solver/ipm.jl:30-74)
void Solver::result(double* p, double* d, double* y, double* info) {
  use_user_h(false);
  std::vector<double> hp(n_), hd(n_), hy(m_);
  HIPCHK(hipMemcpyAsync(hp.data(), idev_.nv + (size_t)NV_P * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(hd.data(), idev_.nv + (size_t)NV_D * n_, n_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(hy.data(), idev_.mv + (size_t)MV_Y * m_, m_ * 8, hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  double mu = NaN, pres = NaN, dres = NaN, pobj = NaN, dobj = NaN;
  int s = status_;
  if (!(s == PRIMAL_INFEASIBLE || s == NEAR_PRIMAL_INFEASIBLE || s == DUAL_INFEASIBLE ||
        s == NEAR_DUAL_INFEASIBLE || hist_.empty())) {
    spmv(2, NV_P, NV_QP, K(1.0), K(0.0));
    spmv(1, MV_Y, NV_T0, K(1.0), K(0.0));
    axpbypcz(0, NV_D, NV_F, NV_QP, NV_DF, K(1.0), K(1.0), K(-1.0));
    axpby(0, NV_T0, NV_DF, NV_DF, K(1.0), K(1.0));
    axpby(1, MV_G, MV_G, MV_DG, K(1.0), K(0.0));
    spmv(0, NV_P, MV_DG, K(-1.0), K(1.0));
    DotBatch b;
    b.add(1, MV_DG, MV_YSCL, 1, S_NORM0, 1).add(0, NV_DF, NV_PSCL, 1, S_NORM1, 1).add(0, NV_P, NV_QP, 0, S_NORM2, 0);
    b.add(0, NV_F, NV_P, 0, S_NORM3, 0).add(1, MV_G, MV_Y, 0, S_NORM4, 0).add(0, NV_P, NV_D, 0, S_TMP0, 0);
    dots(b);
    sync_scalars();
    double pQp = hsc_[S_NORM2];
    pobj = pQp / 2 - hsc_[S_NORM3];
    dobj = hsc_[S_NORM4] - pQp / 2;
    mu = nu_ == 0 ? NaN : hsc_[S_TMP0] / nu_;
    pres = hsc_[S_NORM0] / (1 + sg_);
    dres = hsc_[S_NORM1] / (1 + sf_);
  }
  for (i64 j = 0; j < n_; ++j) {
    if (p) p[P_.P2[j]] = hp[j] * pscl_[j];
    if (d) d[P_.P2[j]] = hd[j] / pscl_[j];
  }
  if (y)
    for (i64 i = 0; i < m_; ++i) y[P_.P1[i]] = hy[i] * yscl_[i];
  if (info) {
    int nsolve = 0;
    for (const HistRow& h : hist_) nsolve += h.piter + h.ppass + h.citer + h.cpass;
    info[0] = mu;
    info[1] = pres;
    info[2] = dres;
    info[3] = pobj;
    info[4] = dobj;
    info[5] = (double)hist_.size();
    info[6] = (double)nsolve;
    info[7] = (double)status_;
  }
}

void Solver::dense_factor(double* out) {
  std::vector<double> F(fsize_);
  HIPCHK(hipMemcpyAsync(F.data(), d_F_.p, fsize_ * 8, hipMemcpyDeviceToHost, st_));
  // a large front's FP64 factor columns live beside its front
  for (auto& bh : big_)
    if (bh->B.L)
      HIPCHK(hipMemcpyAsync(F.data() + fofs_[bh->front], bh->B.L, sizeof(double) * (size_t)bh->B.ld * bh->B.r,
                            hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  std::memset(out, 0, sizeof(double) * n_ * n_);
  size_t ofs = 0;
  for (size_t f = 0; f < fr_.size(); ++f) {
    const FrontDesc& Fd = fr_[f];
    std::vector<int> rows;
    for (int i = 0; i < Fd.r; ++i) rows.push_back(Fd.res0 + i);
    std::vector<int> sep(Fd.s);
    if (Fd.s) {
      HIPCHK(hipMemcpyAsync(sep.data(), Fd.sepcols, Fd.s * 4, hipMemcpyDeviceToHost, st_));
      HIPCHK(hipStreamSynchronize(st_));
    }
    rows.insert(rows.end(), sep.begin(), sep.end());
    for (int j = 0; j < Fd.r; ++j)
      for (int i = j; i < Fd.ld; ++i) out[(size_t)rows[i] + (size_t)rows[j] * n_] = F[ofs + i + (size_t)j * Fd.ld];
    ofs += (size_t)Fd.ld * Fd.ld;
  }
}

}  // namespace mb
