// Persistent kernels for large frontal matrices (ld ≥ BIG_LD).
//
// Factorization: a left-looking tile Cholesky executed by a persistent grid. Task (i, j), i ≥ j,
// owns the lower tile (i, j): it holds the tile in FP64 MFMA accumulators and subtracts
// L_ik L_jkᵀ for every factor column k < min(j, Tr) as soon as both tiles are final. Then:
//   diagonal (j < Tr): first forms L_{j,j−1} = C_{j,j−1} L_{j−1,j−1}⁻ᵀ itself from the staged
//     sub-diagonal tile and subtracts it, then factors the tile, inverts the factor and publishes;
//   sub-diagonal (i = j+1 < Tr): stages its accumulated tile for the next diagonal task;
//   other factor tiles: multiply by the inverted diagonal block L_jj⁻¹ (MFMA) and publish;
//   update-matrix tiles (j ≥ Tr): store.
// Tasks are claimed from an atomic ticket in column order, diagonal first, so a task waits only
// on tasks with smaller tickets, which running workgroups already hold: the grid always makes
// progress. Final tiles publish the factorization epoch in their flag. Waits are bounded: a
// workgroup that spins past the limit raises the abort flag and every workgroup leaves.
//
// Triangular solves: one workgroup per 64-row tile, ticket-ordered; tile b waits for the
// solution blocks it needs and accumulates their contributions while later blocks are still
// being solved. The tile it needs last and its inverted diagonal block are staged in LDS before
// it starts waiting, so each step of the dependency chain is two 64×64 products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>

#include "device.h"
#include "kernels.h"
#include "kernels_big.h"
#include "tile64.h"

namespace mb {

constexpr int LROW = NB + 2;   // row stride of a row-major tile in LDS (16-byte aligned rows)
constexpr long SPIN_LIMIT = 1L << 24;

// Cross-workgroup data moves with the device-coherent cache policy (sc1): flags and single
// values through agent-scope relaxed atomics, tiles through buffer loads/stores carrying the
// same policy, so no L2 writeback or invalidate is needed. A flag is stored only after the
// producer's coherent stores completed (s_waitcnt), and consumers read the data with coherent
// loads after observing the flag.
__device__ __forceinline__ int ld_flag(const int* p) {
  return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_flag(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double cload(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// thread 0 waits until *p >= v; returns false on abort / spin limit (then sets abort)
__device__ __forceinline__ bool wait_geq(const int* p, int v, int* abortf) {
  __shared__ int ok;
  if (threadIdx.x == 0) {
    long spins = 0;
    int good = 1;
    while (ld_flag(p) < v) {
      if (ld_flag(abortf) != 0) {
        good = 0;
        break;
      }
      if (++spins > SPIN_LIMIT) {
        st_flag(abortf, 1);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    ok = good;
  }
  __syncthreads();
  bool r = ok != 0;
  __syncthreads();  // nobody may overwrite ok before every thread has read it
  return r;
}

// every thread's coherent stores have completed, then thread 0 publishes
__device__ __forceinline__ void publish(int* p, int v) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) st_flag(p, v);
}

__device__ __forceinline__ int tri_index(int i, int j) { return i * (i + 1) / 2 + j; }

// ------------------------------------------------------------------------------------------
// the left-looking tile factorization kernel
// ------------------------------------------------------------------------------------------

// T: element type of the factor (FP64, or FP32 for the mixed-precision form). The assembled
// front is always FP64 (B.F) and the update-matrix tiles go back to it in FP64 for the parent's
// extend-add; the factor tiles, the staged sub-diagonal tiles and the inverted diagonal blocks
// live in B.F (FP64) or B.F32 / B.Dinv32 (FP32).
template <class T>
__device__ __forceinline__ T* factor_of(const BigFront& B);
template <>
__device__ __forceinline__ double* factor_of<double>(const BigFront& B) { return B.L ? B.L : B.F; }
template <>
__device__ __forceinline__ float* factor_of<float>(const BigFront& B) { return B.F32; }
template <class T>
__device__ __forceinline__ T* dinv_of(const BigFront& B);
template <>
__device__ __forceinline__ double* dinv_of<double>(const BigFront& B) { return B.Dinv; }
template <>
__device__ __forceinline__ float* dinv_of<float>(const BigFront& B) { return B.Dinv32; }

// Makes sure that columns < need of factor rows i and j are final (their tiles' flags hold the
// epoch), advancing `ready` (uniform). Wave 0 reads the flags of up to 64 columns from `ready` on
// at once, so a task whose inputs are final already pays one round trip for all of them rather
// than one per column; a column not final yet is waited for (bounded) and the poll repeats.
// False on abort or at the wait limit.
__device__ __forceinline__ bool cols_final(const BigFront& B, int i, int j, int kmax, int need, int epoch,
                                           int& ready) {
  __shared__ int ready_s;
  while (ready < need) {
    if (threadIdx.x < 64) {
      const int l = ready + (int)threadIdx.x;
      bool ok = true;
      if (l < kmax) {
        ok = ld_flag(&B.state[tri_index(i, l)]) >= epoch;
        if (i != j) ok = ok && ld_flag(&B.state[tri_index(j, l)]) >= epoch;
      }
      const unsigned long long bad = __ballot(!ok);
      if (threadIdx.x == 0) ready_s = ready + (bad ? __ffsll((long long)bad) - 1 : 64);
    }
    __syncthreads();
    const int r = min(ready_s, kmax);
    __syncthreads();
    if (r >= need) {
      ready = r;
      break;
    }
    ready = r;
    // column `ready` is not final: wait for it, then poll on
    if (!wait_geq(&B.state[tri_index(i, ready)], epoch, B.abortf) ||
        (i != j && !wait_geq(&B.state[tri_index(j, ready)], epoch, B.abortf)))
      return false;
  }
  return true;
}

// One task of the tile DAG (tasks[t] of front B), by the whole workgroup. False when a wait
// ended on the abort flag or at its limit. SEP: write-once factor storage (B.L / B.F32 and B.stg);
// otherwise the factor overwrites F in place and the staged tile sits at its final place.
//
// Tiles and inverted diagonal blocks that other workgroups finalize in this launch are read
// device-coherent (CONS) in both forms. Cached reads would be wrong even for data written once:
// when ld is not a multiple of a cache line's elements, one line holds the last rows of tile
// (i, k) and the first rows of tile (i+1, k) of a column, so a line fetched for a final tile can
// carry rows of its neighbour from before the neighbour was final, and a later cached read of
// the neighbour would hit that copy.
template <class T, bool SEP>
__device__ __forceinline__ bool dag_task(const BigFront& B, const InstDev& I, int epoch, int t, T* SA, T* SB, T* W) {
  constexpr Coh CONS = Coh::Dev;
  double* F = B.F;
  T* Lf = factor_of<T>(B);
  T* Di = dinv_of<T>(B);
  T* Stg = reinterpret_cast<T*>(B.stg);
  const int LD = B.ld;
  stamp(B.trace, t, 0);
  const int4 task = B.tasks[t];
  const int type = task.x, i = task.y, j = task.z, kmax = task.w;
  const int r0i = B.tb[i], hi = B.tb[i + 1] - r0i;
  const int r0j = B.tb[j], hj = B.tb[j + 1] - r0j;
  // the assembled tile (written by earlier launches; only this task writes it here)
  load_tile<Coh::None>(SA, F, LD, r0i, r0j, hi, hj);
  __syncthreads();
  Acc<T> C;
  acc_from_lds(C, SA);
  __syncthreads();
  // C −= Σ_k L_ik L_jkᵀ: the tiles of column k+1 are in flight while the MFMAs of column k run.
  // One LDS tile pair per workgroup, so two workgroups share a CU and one's staging and barriers
  // overlap the other's products.
  bool alive = true;
  TileRegs<T> RA, RB;
  // columns < ready hold final tiles in rows i and j (uniform across the workgroup)
  int ready = 0;
  if (kmax > 0) {
    if (!cols_final(B, i, j, kmax, 1, epoch, ready)) return false;
    const int c00 = B.tb[0], w0 = B.tb[1] - c00;
    fetch_tile<CONS>(RA, Lf, LD, r0i, c00, hi, w0);
    if (i != j) fetch_tile<CONS>(RB, Lf, LD, r0j, c00, hj, w0);
  }
  for (int k = 0; k < kmax; ++k) {
    put_tile(SA, RA);
    if (i != j) put_tile(SB, RB);
    __syncthreads();
    if (k + 1 < kmax) {
      if (k + 1 >= ready && !cols_final(B, i, j, kmax, k + 2, epoch, ready)) {
        alive = false;
        break;
      }
      const int c0n = B.tb[k + 1], wn = B.tb[k + 2] - c0n;
      fetch_tile<CONS>(RA, Lf, LD, r0i, c0n, hi, wn);
      if (i != j) fetch_tile<CONS>(RB, Lf, LD, r0j, c0n, hj, wn);
    }
    mfma_acc<true>(C, SA, i != j ? SB : SA, B.tb[k + 1] - B.tb[k]);
    __syncthreads();  // the pair is overwritten next
  }
  if (!alive) return false;
  stamp(B.trace, t, 1);
  if (type == 0) {
    if (j > 0) {
      // L_{j,j−1} = C_{j,j−1} L_{j−1,j−1}⁻ᵀ from the staged tile (fetched before the inverse is
      // ready), then C_jj −= L_{j,j−1} L_{j,j−1}ᵀ
      const int c0p = B.tb[j - 1], wp = B.tb[j] - c0p;
      if (!wait_geq(&B.stage[j - 1], epoch, B.abortf)) return false;
      if (SEP) load_tile<Coh::Dev>(SA, Stg + (size_t)(j - 1) * NB * NB, NB, 0, 0, hj, wp);
      else load_tile<Coh::Dev>(SA, Lf, LD, r0j, c0p, hj, wp);
      if (!wait_geq(&B.state[tri_index(j - 1, j - 1)], epoch, B.abortf)) return false;
      load_tile<CONS>(SB, Di + (size_t)(j - 1) * NB * NB, NB, 0, 0, NB, NB);
      __syncthreads();
      Acc<T> X;
      acc_zero(X);
      mfma_acc<false>(X, SA, SB, NB);
      __syncthreads();
      acc_to_lds(X, SA);
      __syncthreads();
      store_tile<Coh::Dev>(Lf, LD, r0j, c0p, hj, wp, SA, false);
      mfma_acc<true>(C, SA, SA, wp);
      publish(&B.state[tri_index(j, j - 1)], epoch);  // the stores drained during the update
    }
    stamp(B.trace, t, 2);
    // factor and invert; L_jj⁻¹ → Dinv and publish, then L_jj → F (read only by later launches)
    acc_to_lds(C, SA);
    __syncthreads();
    bool ok = chol_inv64(SA, SB, W, hj, B.trace, t);
    stamp(B.trace, t, 6);
    if (!ok && threadIdx.x == 0) atomicAdd(&I.sc[S_FAIL], 1.0);
    store_tile<Coh::Dev>(Di + (size_t)j * NB * NB, NB, 0, 0, NB, NB, SB, false);
    publish(&B.state[tri_index(j, j)], epoch);
    stamp(B.trace, t, 7);
    store_tile<Coh::Dev>(Lf, LD, r0j, r0j, hj, hj, SA, true);
  } else if (type == 3) {
    // sub-diagonal tile: hand the accumulated tile to the next diagonal task
    acc_to_lds(C, SA);
    __syncthreads();
    stamp(B.trace, t, 2);
    if (SEP) store_tile<Coh::Dev>(Stg + (size_t)j * NB * NB, NB, 0, 0, hi, hj, SA, false);
    else store_tile<Coh::Dev>(Lf, LD, r0i, r0j, hi, hj, SA, false);
    publish(&B.stage[j], epoch);
  } else if (type == 1) {
    // L_ij = C_ij L_jj⁻ᵀ
    acc_to_lds(C, SA);
    if (!wait_geq(&B.state[tri_index(j, j)], epoch, B.abortf)) return false;
    stamp(B.trace, t, 2);
    load_tile<CONS>(SB, Di + (size_t)j * NB * NB, NB, 0, 0, NB, NB);
    __syncthreads();
    Acc<T> X;
    acc_zero(X);
    mfma_acc<false>(X, SA, SB, NB);
    __syncthreads();
    acc_to_lds(X, SA);
    __syncthreads();
    store_tile<Coh::Dev>(Lf, LD, r0i, r0j, hi, hj, SA, false);
    publish(&B.state[tri_index(i, j)], epoch);
  } else {
    // update-matrix tile: final once every factor column has been subtracted (FP64 for the
    // parent's extend-add)
    acc_to_lds(C, SA);
    __syncthreads();
    stamp(B.trace, t, 2);
    store_tile<Coh::Dev>(F, LD, r0i, r0j, hi, hj, SA, i == j);
  }
  stamp(B.trace, t, 3);
  return true;
}

// LDS of one factorization workgroup: a tile pair and the diagonal factor's scratch (about
// 68 KiB in FP64), so DAG_PER_CU = 2 workgroups fit a CU's 160 KiB
template <class T>
struct DagLds {
  T W[3 * NB];
  T A[NB * LROW];
  T B[NB * LROW];
};

template <class T, bool SEP>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(DAG_PER_CU, DAG_PER_CU)))
k_dag_factor(BigFront B, const InstDev* insts, int epoch) {
  __shared__ __align__(16) DagLds<T> S;
  __shared__ int task_s;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  for (;;) {
    if (threadIdx.x == 0) {
      int t = atomicAdd(B.ticket, 1);
      if (t < B.ntask && ld_flag(B.abortf) != 0) t = B.ntask;
      task_s = t;
    }
    __syncthreads();
    const int t = task_s;
    __syncthreads();
    if (t >= B.ntask) break;
    if (!dag_task<T, SEP>(B, I, epoch, t, S.A, S.B, S.W)) break;
  }
}

// The large fronts of several instances (a gang) in one persistent launch: ticket g is task
// order[g] = (front, task) of the list; each front's tasks appear in their own topological
// order, so every wait is on a smaller ticket, as in the single-front kernel. A front whose
// instance is not taking part, or whose abort flag is raised, has its remaining tasks skipped;
// the other fronts go on. The ticket goes back to 0 in k_big_mtiles_multi.
template <class T>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(DAG_PER_CU, DAG_PER_CU)))
k_dag_factor_multi(const BigFront* Bs, const int2* order, DagQueues Q, int* ticket, const InstDev* insts,
                   int epoch) {
  __shared__ __align__(16) DagLds<T> S;
  __shared__ int task_s;
  // queue q is served by the workgroups ≡ q (mod nq): with nq = 8 those of one XCD (the dispatcher
  // deals workgroup i to XCD i mod 8), so a front's tiles are re-read from that XCD's L2
  const int q = (int)blockIdx.x % Q.n;
  const int qn = Q.off[q + 1] - Q.off[q];
  for (;;) {
    if (threadIdx.x == 0) task_s = atomicAdd(ticket + q, 1);
    __syncthreads();
    const int g = __builtin_amdgcn_readfirstlane(task_s);
    __syncthreads();
    if (g >= qn) break;
    const int2 o = order[Q.off[q] + g];
    const BigFront& B = Bs[o.x];
    const InstDev& I = insts[B.inst];
    if (I.sc[S_ACTIVE] == 0.0 || ld_flag(B.abortf) != 0) continue;
    (void)dag_task<T, true>(B, I, epoch, o.y, S.A, S.B, S.W);
  }
}

// ------------------------------------------------------------------------------------------
// pipelined triangular solves of one big front
// ------------------------------------------------------------------------------------------

// gather children's update vectors into v[res] and the front's own uvec (one block)
__global__ void k_big_gather(const FrontDesc* fronts, int f, const int* childlist, const InstDev* insts, int kv) {
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  double* u = Fd.uvec;
  int r = Fd.r;
  for (int k = threadIdx.x; k < Fd.s; k += blockDim.x) u[k] = 0.0;
  __syncthreads();
  for (int c = 0; c < Fd.nchild; ++c) {
    const FrontDesc& C = fronts[childlist[Fd.child0 + c]];
    for (int k = threadIdx.x; k < C.s; k += blockDim.x) {
      int p = C.relidx[k];
      if (p < r) v[p] -= C.uvec[k];
      else u[p - r] += C.uvec[k];
    }
    __syncthreads();
  }
}


// the gathers of every front of a gang launch, one workgroup each
__global__ void k_big_gather_multi(const int* bfront, const FrontDesc* fronts, const int* childlist,
                                   const InstDev* insts, int kv) {
  const FrontDesc& Fd = fronts[bfront[blockIdx.x]];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  double* u = Fd.uvec;
  int r = Fd.r;
  for (int k = threadIdx.x; k < Fd.s; k += blockDim.x) u[k] = 0.0;
  __syncthreads();
  for (int c = 0; c < Fd.nchild; ++c) {
    const FrontDesc& C = fronts[childlist[Fd.child0 + c]];
    for (int k = threadIdx.x; k < C.s; k += blockDim.x) {
      int p = C.relidx[k];
      if (p < r) v[p] -= C.uvec[k];
      else u[p - r] += C.uvec[k];
    }
    __syncthreads();
  }
}

// Solution blocks pass between the solve workgroups as (value, epoch) words: entry r of block b
// is two 64-bit words, each one 32-bit half of x_b[r] beside the launch's epoch, stored and loaded
// whole (single-copy atomic at device scope). A reader that sees the epoch in both words of an
// entry holds its value: no flag, and no wait for the producer's stores to complete. Entries past
// a block's order carry zeros. Every launch has a new epoch, so older contents never match.
__device__ __forceinline__ void ll_put(unsigned long long* L, int blk, int r, double x, int epoch) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned long long e = (unsigned long long)(unsigned)epoch << 32;
  unsigned long long* p = L + ((size_t)blk * NB + r) * 2;
  __hip_atomic_store(p, (u & 0xffffffffull) | e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, (u >> 32) | e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one wave reads block blk (lane r: entry r); true (wave-uniform) when every entry is complete
__device__ __forceinline__ bool ll_get(const unsigned long long* L, int blk, int epoch, double& x) {
  const int lane = threadIdx.x & 63;
  unsigned long long* p = const_cast<unsigned long long*>(L) + ((size_t)blk * NB + lane) * 2;
  const unsigned long long lo = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long hi = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const bool ok = (unsigned)(lo >> 32) == (unsigned)epoch && (unsigned)(hi >> 32) == (unsigned)epoch;
  x = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
  return __ballot(ok) == ~0ull;
}

// one wave waits for block blk; false on abort or at the spin limit (which raises the abort flag)
__device__ __forceinline__ bool ll_wait(const unsigned long long* L, int blk, int epoch, int* abortf, double& x) {
  long spins = 0;
  for (;;) {
    if (ll_get(L, blk, epoch, x)) return true;
    if (ld_flag(abortf) != 0) return false;
    if (++spins > SPIN_LIMIT) {
      if ((threadIdx.x & 63) == 0) st_flag(abortf, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// The workgroup takes blocks j, j+step, … (at most n ≤ 4): wave w reads block j + w·step into
// xs[w·64, w·64+64), wave 0 waiting for its block. Returns how many leading blocks are complete
// (≥ 1), or 0 on abort.
__device__ __forceinline__ int ll_take(const unsigned long long* L, int j, int step, int n, int epoch, int* abortf,
                                       double* xs) {
  __shared__ int okw[NT / 64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  bool ok = false;
  double x = 0.0;
  if (w == 0) ok = ll_wait(L, j, epoch, abortf, x);
  else if (w < n) ok = ll_get(L, j + step * w, epoch, x);
  if (ok) xs[w * NB + lane] = x;
  if (lane == 0) okw[w] = ok;
  __syncthreads();
  int c = 0;
  while (c < n && okw[c]) ++c;
  __syncthreads();  // okw is rewritten by the next call
  return c;
}

// The last step of a chain, by wave 0 alone: lane r holds w[r] and row r of the 64×64 step tile
// S (column-major FP64); it waits for block blk, then x[r] = w[r] − Σ_c S[r][c] x_blk[c] (four
// partial sums in a fixed order). xb: 64 doubles of LDS for the broadcast.
__device__ __forceinline__ bool chain_step(const double* S, const unsigned long long* L, int blk, int epoch,
                                           int* abortf, double* xb, double& x) {
  const int lane = threadIdx.x & 63;
  double srow[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) srow[c] = S[lane + c * NB];
  double xp;
  if (!ll_wait(L, blk, epoch, abortf, xp)) return false;
  xb[lane] = xp;
  wave_lds_sync();
  double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < NB; ++c) a[c & 3] += srow[c] * xb[c];
  x -= (a[0] + a[1]) + (a[2] + a[3]);
  return true;
}

// A 64×64 factor tile held for a matrix-vector product, 16 elements per thread. Row layout
// (forward solve): thread (r, q) = (t & 63, t >> 6) holds element (r, q + 4i). Column layout
// (backward solve): thread (c, q) holds element (q + 4i, c). Entries past the tile are zero.
struct GemvRegs {
  double v[16];
};

template <class TS>
__device__ __forceinline__ void gemv_rows(GemvRegs& R, const TS* F, int LD, int r0, int c0, int h, int w) {
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int c = q + 4 * i;
    R.v[i] = (r < h && c < w) ? (double)F[(size_t)(r0 + r) + (size_t)(c0 + c) * LD] : 0.0;
  }
}

// column layout through an LDS tile S (the global reads stay coalesced); S is free again on return
template <class TS>
__device__ __forceinline__ void gemv_cols(GemvRegs& R, double* S, const TS* F, int LD, int r0, int c0, int h, int w) {
  const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
  load_tile<Coh::None>(S, F, LD, r0, c0, h, w);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; ++i) R.v[i] = S[(q + 4 * i) + c * LDT];
  __syncthreads();
}

// Σ over this thread's 16 elements of R times x[q + 4i]
__device__ __forceinline__ double gemv_dot(const GemvRegs& R, const double* x) {
  const int q = threadIdx.x >> 6;
  double a = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) a += R.v[i] * x[q + 4 * i];
  return a;
}

// the 64×64 tile at (r0, c0) of F (h × w, zero padded) into LDS as FP64, transposed if TR
template <bool TR, class TS>
__device__ __forceinline__ void lds_tile64(double* S, const TS* F, int LD, int r0, int c0, int h, int w) {
  double x[NB * NB / NT];  // all loads in flight before the first LDS write
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    const int t = threadIdx.x + u * NT, i = t % NB, k = t / NB;
    x[u] = (i < h && k < w) ? (double)F[(size_t)(r0 + i) + (size_t)(c0 + k) * LD] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    const int t = threadIdx.x + u * NT, i = t % NB, k = t / NB;
    if (TR) S[k + i * LDT] = x[u];
    else S[i + k * LDT] = x[u];
  }
}

// After a factorization: for each factor block b, M_b = L_bb⁻¹ L_{b,b−1} (b ≥ 1) and
// N_bᵀ with N_b = L_{b+1,b} L_bb⁻¹ (b+1 < Tr), FP64, 64×64 column-major. With them each step of
// the solves' chains is one 64×64 product after the previous block arrives:
// x_b = w_b − M_b x_{b−1} (forward), x_b = w_b − N_bᵀ x_{b+1} (backward), where w_b is formed
// from everything else while the workgroup waits. One workgroup per block.
template <class TF>
__device__ __forceinline__ void mtiles_body(const BigFront& B, int blk, double* SX, double* SY) {
  // workgroups [0, Tr) form M_b, workgroups [Tr, 2 Tr) form N_b
  const bool mpart = blk < B.Tr;
  const int b = mpart ? blk : blk - B.Tr;
  if (b >= B.Tr) return;
  const TF* F = factor_of<TF>(B);
  const TF* Di = dinv_of<TF>(B) + (size_t)b * NB * NB;
  const int LD = B.ld, c0 = B.tb[b], w = B.tb[b + 1] - c0;
  if (mpart && b >= 1) {
    // M_b = X_b L with L = L_{b,b−1}: mfma_acc forms A += X Yᵀ, so Y = Lᵀ
    const int cp = B.tb[b - 1], wp = c0 - cp;
    lds_tile64<false>(SX, Di, NB, 0, 0, NB, NB);
    lds_tile64<true>(SY, F, LD, c0, cp, w, wp);
    __syncthreads();
    Acc<double> A;
    acc_zero(A);
    mfma_acc<false>(A, SX, SY, NB);
    __syncthreads();
    acc_to_lds(A, SX);
    __syncthreads();
    double* M = B.Mf + (size_t)b * NB * NB;
    for (int t = threadIdx.x; t < NB * NB; t += NT) M[t] = SX[(t % NB) + (t / NB) * LDT];
  }
  if (!mpart && b + 1 < B.Tr) {
    // N_b = L' X_b with L' = L_{b+1,b}: Y = X_bᵀ
    const int rn = B.tb[b + 1], hn = B.tb[b + 2] - rn;
    lds_tile64<false>(SX, F, LD, rn, c0, hn, w);
    lds_tile64<true>(SY, Di, NB, 0, 0, NB, NB);
    __syncthreads();
    Acc<double> A;
    acc_zero(A);
    mfma_acc<false>(A, SX, SY, NB);
    __syncthreads();
    acc_to_lds(A, SX);
    __syncthreads();
    double* Nt = B.Nt + (size_t)b * NB * NB;
    for (int t = threadIdx.x; t < NB * NB; t += NT) Nt[t] = SX[(t / NB) + (t % NB) * LDT];
  }
}

template <class TF>
__global__ void __launch_bounds__(NT) k_big_mtiles(BigFront B, const InstDev* insts) {
  __shared__ __align__(16) double SX[NB * LDT];
  __shared__ __align__(16) double SY[NB * LDT];
  // the factorization that just ended took its last ticket: the next one starts from 0
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.ticket = 0;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  mtiles_body<TF>(B, (int)blockIdx.x, SX, SY);
}

// the chain-step tiles of every front of a gang launch (grid: 2·max Tr × fronts), and the gang's
// factorization ticket back to 0
template <class TF>
__global__ void __launch_bounds__(NT) k_big_mtiles_multi(const BigFront* Bs, int* ticket, const InstDev* insts) {
  __shared__ __align__(16) double SX[NB * LDT];
  __shared__ __align__(16) double SY[NB * LDT];
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < DAG_QUEUES) ticket[threadIdx.x] = 0;
  const BigFront& B = Bs[blockIdx.y];
  if ((int)blockIdx.x >= 2 * B.Tr) return;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  mtiles_body<TF>(B, (int)blockIdx.x, SX, SY);
}

// forward: tile b (ticket order) accumulates L[b, j] x_j for the complete j < b−1 while the
// earlier tiles solve; the factor tiles stream through registers one block ahead of the solution,
// and every block's contribution is summed in block order whatever the batches (reproducible
// results). A factor tile then forms w_b = L_bb⁻¹ (v_b − Σ) before block b−1 arrives, and wave 0
// finishes x_b = w_b − M_b x_{b−1} after it. A separator tile adds its Σ over all factor blocks
// into the front's update vector. TF: element type of the stored factor; the arithmetic is FP64.
template <class TF>
__device__ __forceinline__ void fsolve_body(const BigFront& B, const FrontDesc& Fd, const InstDev& I, int kv,
                                            int epoch, int b, double* Sinv, double* xs, double (*part)[NB]) {
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  const TF* F = factor_of<TF>(B);
  const int LD = B.ld;
  const int row = threadIdx.x & 63, q = threadIdx.x >> 6;
  const bool diag = b < B.Tr;
  const int jacc = diag ? max(b - 1, 0) : B.Tr;  // blocks summed before the last step
  const int r0 = B.tb[b], h = B.tb[b + 1] - r0;
  if (diag) load_tile<Coh::None>(Sinv, dinv_of<TF>(B) + (size_t)b * NB * NB, NB, 0, 0, NB, NB);
  double acc = 0.0;
  GemvRegs Rn;
  if (jacc > 0) gemv_rows(Rn, F, LD, r0, B.tb[0], h, B.tb[1] - B.tb[0]);
  for (int j = 0; j < jacc;) {
    const int n = ll_take(B.sll, j, 1, min(4, jacc - j), epoch, B.abortf, xs);
    if (n == 0) return;
    for (int k = j; k < j + n; ++k) {
      const GemvRegs Rc = Rn;
      if (k + 1 < jacc) gemv_rows(Rn, F, LD, r0, B.tb[k + 1], h, B.tb[k + 2] - B.tb[k + 1]);
      acc += gemv_dot(Rc, xs + (k - j) * NB);
    }
    __syncthreads();
    j += n;
  }
  part[q][row] = acc;
  __syncthreads();
  if (!diag) {
    if (threadIdx.x < h) Fd.uvec[r0 - Fd.r + row] += part[0][row] + part[1][row] + part[2][row] + part[3][row];
    return;
  }
  // w_b = L_bb⁻¹ (v_b − Σ_{j<b−1} L_bj x_j), before block b−1 arrives
  if (threadIdx.x < NB)
    xs[row] = row < h ? v[r0 + row] - (part[0][row] + part[1][row] + part[2][row] + part[3][row]) : 0.0;
  __syncthreads();
  double y = 0.0;
#pragma unroll
  for (int c = q; c < NB; c += 4) y += Sinv[row + c * LDT] * xs[c];
  part[q][row] = y;
  __syncthreads();
  if (threadIdx.x >= NB) return;  // wave 0 finishes
  double x = part[0][row] + part[1][row] + part[2][row] + part[3][row];
  if (b > 0 && !chain_step(B.Mf + (size_t)b * NB * NB, B.sll, b - 1, epoch, B.abortf, xs, x)) return;
  if (row >= h) x = 0.0;
  ll_put(B.sll, b, row, x, epoch);
  if (row < h) v[r0 + row] = x;
}

template <class TF>
__global__ void __launch_bounds__(NT) k_big_fsolve(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                   int kv, int epoch, long long base) {
  __shared__ double Sinv[NB * LDT];
  __shared__ double xs[4 * NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  // every workgroup consumes exactly one ticket (the host advances the base by the grid size)
  if (threadIdx.x == 0) tk = (long long)atomicAdd((unsigned long long*)B.sticket, 1ull) - base;
  __syncthreads();
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int b = (int)tk;
  if (b < 0 || b >= B.T) return;
  fsolve_body<TF>(B, Fd, I, kv, epoch, b, Sinv, xs, part);
}

// The forward solves of the large fronts of a gang launch: ticket g is block order[g].y of front
// order[g].x (each front's blocks in increasing order, so a block's inputs are claimed before it);
// bfront: each listed front's index in the front table. One ticket per workgroup.
template <class TF>
__global__ void __launch_bounds__(NT) k_big_fsolve_multi(const BigFront* Bs, const int* bfront, const int2* order,
                                                         int ntot, unsigned long long* ticket, const FrontDesc* fronts,
                                                         const InstDev* insts, int kv, int epoch, long long base) {
  __shared__ double Sinv[NB * LDT];
  __shared__ double xs[4 * NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  if (threadIdx.x == 0) tk = (long long)atomicAdd(ticket, 1ull) - base;
  __syncthreads();
  const int g = (int)tk;
  if (g < 0 || g >= ntot) return;
  const int2 o = order[g];
  const BigFront& B = Bs[o.x];
  const FrontDesc& Fd = fronts[bfront[o.x]];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  fsolve_body<TF>(B, Fd, I, kv, epoch, o.y, Sinv, xs, part);
}

// L[sep, b]ᵀ x_sep for one factor column block b and one 64-row chunk of the separator rows
// (grid: Tr × ⌈s/64⌉): the tile goes through LDS so that the global reads run down its columns.
// Separator solution values are final before the backward solve of a front starts.
template <class TF>
__device__ __forceinline__ void sepgemv_body(const BigFront& B, const FrontDesc& Fd, const InstDev& I, int kv, int b,
                                             int chunk, double* S, double* xsep, double (*part)[NB]) {
  const int rs0 = Fd.r + chunk * NB;
  const int h = min(NB, B.ld - rs0), c0 = B.tb[b], w = B.tb[b + 1] - c0;
  const double* vall = I.nv + (size_t)kv * I.n;
  load_tile<Coh::None>(S, factor_of<TF>(B), B.ld, rs0, c0, h, w);
  if (threadIdx.x < NB) xsep[threadIdx.x] = threadIdx.x < h ? vall[Fd.sepcols[rs0 - Fd.r + threadIdx.x]] : 0.0;
  __syncthreads();
  const int col = threadIdx.x & 63, q = threadIdx.x >> 6;
  double a = 0.0;
#pragma unroll
  for (int rr = q; rr < NB; rr += 4) a += S[rr + col * LDT] * xsep[rr];
  part[q][col] = a;
  __syncthreads();
  if (threadIdx.x < NB)
    B.spart[((size_t)chunk * B.Tr + b) * NB + col] = (part[0][col] + part[1][col]) + (part[2][col] + part[3][col]);
}

template <class TF>
__global__ void __launch_bounds__(NT) k_big_sepgemv(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                    int kv) {
  __shared__ double S[NB * LDT];
  __shared__ double xsep[NB];
  __shared__ double part[4][NB];
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  sepgemv_body<TF>(B, Fd, I, kv, (int)blockIdx.x, (int)blockIdx.y, S, xsep, part);
}

// the separator products of every front of a gang launch (grid: max Tr × max chunks × fronts)
template <class TF>
__global__ void __launch_bounds__(NT) k_big_sepgemv_multi(const BigFront* Bs, const int* bfront, const FrontDesc* fronts,
                                                          const InstDev* insts, int kv) {
  __shared__ double S[NB * LDT];
  __shared__ double xsep[NB];
  __shared__ double part[4][NB];
  const BigFront& B = Bs[blockIdx.z];
  const FrontDesc& Fd = fronts[bfront[blockIdx.z]];
  if ((int)blockIdx.x >= B.Tr || (int)blockIdx.y * NB >= B.s) return;
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  sepgemv_body<TF>(B, Fd, I, kv, (int)blockIdx.x, (int)blockIdx.y, S, xsep, part);
}

// backward: column block b = Tr−1−ticket accumulates L[j, b]ᵀ x_j for the complete j > b+1:
// thread (r, q) sums L[j, b][r][q + 4i] x_j[r] over the blocks (in descending order whatever the
// batches), one transpose through LDS then sums over r. With L[sep, b]ᵀ x_sep (the chunk partials
// of k_big_sepgemv) it forms w_b = L_bbᵀ⁻¹ (v_b − Σ) before block b+1 arrives, and wave 0
// finishes x_b = w_b − N_bᵀ x_{b+1} after it.
template <class TF>
__device__ __forceinline__ void bsolve_body(const BigFront& B, const FrontDesc& Fd, const InstDev& I, int kv,
                                            int epoch, int tkt, double* S, double* Sinv, double* xs,
                                            double (*part)[NB]) {
  const int b = B.Tr - 1 - tkt;
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  const TF* F = factor_of<TF>(B);
  const int LD = B.ld;
  const int c0 = B.tb[b], w = B.tb[b + 1] - c0;
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  load_tile<Coh::None>(Sinv, dinv_of<TF>(B) + (size_t)b * NB * NB, NB, 0, 0, NB, NB);
  // L[sep, b]ᵀ x_sep: the chunk partials of k_big_sepgemv, chunks c ≡ q (mod 4) by wave q in
  // chunk order, read while the later blocks solve
  double sep = 0.0;
  {
    const int nch = (LD - Fd.r + NB - 1) / NB;
#pragma unroll 4
    for (int c = q; c < nch; c += 4) sep += B.spart[((size_t)c * B.Tr + b) * NB + lane];
  }
  const int jlo = b + 2;  // blocks jlo..Tr−1 are summed before the last step
  double av[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) av[i] = 0.0;
  GemvRegs Rn;
  if (B.Tr - 1 >= jlo) gemv_rows(Rn, F, LD, B.tb[B.Tr - 1], c0, B.tb[B.Tr] - B.tb[B.Tr - 1], w);
  for (int j = B.Tr - 1; j >= jlo;) {
    const int n = ll_take(B.sll, j, -1, min(4, j - jlo + 1), epoch, B.abortf, xs);
    if (n == 0) return;
    for (int k = j; k > j - n; --k) {
      const GemvRegs Rc = Rn;
      if (k - 1 >= jlo) gemv_rows(Rn, F, LD, B.tb[k - 1], c0, B.tb[k] - B.tb[k - 1], w);
      const double xr = xs[(j - k) * NB + lane];
#pragma unroll
      for (int i = 0; i < 16; ++i) av[i] += Rc.v[i] * xr;
    }
    __syncthreads();
    j -= n;
  }
  // column sums: S[r][c] = thread (r, ·)'s partial for column c, then thread (c, q) sums rows q + 4m
#pragma unroll
  for (int i = 0; i < 16; ++i) S[lane + (q + 4 * i) * LDT] = av[i];
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 16; ++m) acc += S[(q + 4 * m) + lane * LDT];
  acc += sep;
  // w_b = L_bbᵀ⁻¹ (v_b − Σ), before block b+1 arrives
  part[q][lane] = acc;
  __syncthreads();
  if (threadIdx.x < NB)
    xs[lane] = lane < w ? v[c0 + lane] - (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]) : 0.0;
  __syncthreads();
  double y = 0.0;
#pragma unroll
  for (int rr = q; rr < NB; rr += 4) y += Sinv[rr + lane * LDT] * xs[rr];
  part[q][lane] = y;
  __syncthreads();
  if (threadIdx.x >= NB) return;  // wave 0 finishes
  double x = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
  if (b + 1 < B.Tr && !chain_step(B.Nt + (size_t)b * NB * NB, B.sll, b + 1, epoch, B.abortf, xs, x)) return;
  if (lane >= w) x = 0.0;
  ll_put(B.sll, b, lane, x, epoch);
  if (lane < w) v[c0 + lane] = x;
}

template <class TF>
__global__ void __launch_bounds__(NT) k_big_bsolve(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                   int kv, int epoch, long long base) {
  __shared__ double S[NB * LDT];
  __shared__ double Sinv[NB * LDT];
  __shared__ double xs[4 * NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (threadIdx.x == 0) tk = (long long)atomicAdd((unsigned long long*)B.sticket, 1ull) - base;
  __syncthreads();
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int tkt = (int)tk;
  if (tkt < 0 || tkt >= B.Tr) return;
  bsolve_body<TF>(B, Fd, I, kv, epoch, tkt, S, Sinv, xs, part);
}

// The backward solves of the large fronts of a gang launch: ticket g is the order[g].y-th block
// from the last of front order[g].x
template <class TF>
__global__ void __launch_bounds__(NT) k_big_bsolve_multi(const BigFront* Bs, const int* bfront, const int2* order,
                                                         int ntot, unsigned long long* ticket, const FrontDesc* fronts,
                                                         const InstDev* insts, int kv, int epoch, long long base) {
  __shared__ double S[NB * LDT];
  __shared__ double Sinv[NB * LDT];
  __shared__ double xs[4 * NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  if (threadIdx.x == 0) tk = (long long)atomicAdd(ticket, 1ull) - base;
  __syncthreads();
  const int g = (int)tk;
  if (g < 0 || g >= ntot) return;
  const int2 o = order[g];
  const BigFront& B = Bs[o.x];
  const FrontDesc& Fd = fronts[bfront[o.x]];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  bsolve_body<TF>(B, Fd, I, kv, epoch, o.y, S, Sinv, xs, part);
}

// ------------------------------------------------------------------------------------------

// The ticket starts at 0 (zeroed with the front's buffers, then put back by k_big_mtiles after
// each factorization, which launch_big_mtiles must follow); the abort flag stays raised once set.
void launch_dag_factor(hipStream_t st, const BigFront& B, const InstDev* I, int grid, int epoch, bool fp32) {
  const bool sep = B.stg != nullptr;
  if (fp32 && !sep) throw std::invalid_argument("FP32 factorization without staging tiles");
  grid *= DAG_PER_CU;
  if (fp32) k_dag_factor<float, true><<<grid, NT, 0, st>>>(B, I, epoch);
  else if (sep && B.L) k_dag_factor<double, true><<<grid, NT, 0, st>>>(B, I, epoch);
  else k_dag_factor<double, false><<<grid, NT, 0, st>>>(B, I, epoch);
}

void launch_big_mtiles(hipStream_t st, const BigFront& B, const InstDev* I, bool fp32) {
  const int mgrid = std::max(1, 2 * B.Tr);
  if (fp32) k_big_mtiles<float><<<mgrid, NT, 0, st>>>(B, I);
  else k_big_mtiles<double><<<mgrid, NT, 0, st>>>(B, I);
}

int launch_big_fsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const int* childlist,
                      const InstDev* I, int kv, int epoch, long long base, bool fp32) {
  const int grid = B.T;
  k_big_gather<<<1, 256, 0, st>>>(F, f, childlist, I, kv);
  if (fp32) k_big_fsolve<float><<<grid, NT, 0, st>>>(B, F, f, I, kv, epoch, base);
  else k_big_fsolve<double><<<grid, NT, 0, st>>>(B, F, f, I, kv, epoch, base);
  return grid;
}

#define HIPCHK_K(x)                                                                        \
  do {                                                                                     \
    if ((x) != hipSuccess) throw std::runtime_error("HIP error at " #x);                  \
  } while (0)

void launch_dag_factor_multi(hipStream_t st, const BigFront* Bs, int nf, const int2* order, const DagQueues& Q,
                             int* ticket, const InstDev* I, int grid, int epoch, bool fp32, int max2tr,
                             hipEvent_t before, hipEvent_t after) {
  if (before) HIPCHK_K(hipEventRecord(before, st));
  if (Q.n < 1 || Q.n > DAG_QUEUES || Q.off[Q.n] <= 0) return;
  grid = std::max(grid * DAG_PER_CU, Q.n);
  if (fp32) k_dag_factor_multi<float><<<grid, NT, 0, st>>>(Bs, order, Q, ticket, I, epoch);
  else k_dag_factor_multi<double><<<grid, NT, 0, st>>>(Bs, order, Q, ticket, I, epoch);
  if (after) HIPCHK_K(hipEventRecord(after, st));
  const dim3 mg(std::max(1, max2tr), nf);
  if (fp32) k_big_mtiles_multi<float><<<mg, NT, 0, st>>>(Bs, ticket, I);
  else k_big_mtiles_multi<double><<<mg, NT, 0, st>>>(Bs, ticket, I);
}

void launch_big_fsolve_multi(hipStream_t st, const BigFront* Bs, const int* bfront, int nf, const int2* order,
                             int ntot, unsigned long long* ticket, const FrontDesc* F, const int* childlist,
                             const InstDev* I, int kv, int epoch, long long base, bool fp32) {
  if (ntot <= 0) return;
  k_big_gather_multi<<<nf, 256, 0, st>>>(bfront, F, childlist, I, kv);
  if (fp32) k_big_fsolve_multi<float><<<ntot, NT, 0, st>>>(Bs, bfront, order, ntot, ticket, F, I, kv, epoch, base);
  else k_big_fsolve_multi<double><<<ntot, NT, 0, st>>>(Bs, bfront, order, ntot, ticket, F, I, kv, epoch, base);
}

void launch_big_bsolve_multi(hipStream_t st, const BigFront* Bs, const int* bfront, int nf, const int2* order,
                             int ntot, unsigned long long* ticket, const FrontDesc* F, const InstDev* I, int kv,
                             int epoch, long long base, bool fp32, int maxtr, int maxch) {
  if (ntot <= 0) return;
  if (maxch > 0) {
    const dim3 sg(std::max(1, maxtr), maxch, nf);
    if (fp32) k_big_sepgemv_multi<float><<<sg, NT, 0, st>>>(Bs, bfront, F, I, kv);
    else k_big_sepgemv_multi<double><<<sg, NT, 0, st>>>(Bs, bfront, F, I, kv);
  }
  if (fp32) k_big_bsolve_multi<float><<<ntot, NT, 0, st>>>(Bs, bfront, order, ntot, ticket, F, I, kv, epoch, base);
  else k_big_bsolve_multi<double><<<ntot, NT, 0, st>>>(Bs, bfront, order, ntot, ticket, F, I, kv, epoch, base);
}

int launch_big_bsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const InstDev* I, int kv,
                      int epoch, long long base, bool fp32) {
  const int grid = B.Tr;
  if (B.s > 0 && B.Tr > 0) {
    const dim3 sg(B.Tr, (B.s + NB - 1) / NB);
    if (fp32) k_big_sepgemv<float><<<sg, NT, 0, st>>>(B, F, f, I, kv);
    else k_big_sepgemv<double><<<sg, NT, 0, st>>>(B, F, f, I, kv);
  }
  if (fp32) k_big_bsolve<float><<<grid, NT, 0, st>>>(B, F, f, I, kv, epoch, base);
  else k_big_bsolve<double><<<grid, NT, 0, st>>>(B, F, f, I, kv, epoch, base);
  return grid;
}

}  // namespace mb
