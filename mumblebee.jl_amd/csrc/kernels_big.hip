// Persistent kernels for large frontal matrices (ld ≥ BIG_LD).
//
// Factorization: a tile DAG executed by a persistent grid. Tasks are claimed in a fixed
// topological order from an atomic ticket counter: POTRF(k), TRSM(i,k), UPDATE(i,j,k). A task
// waits only on tasks with smaller tickets, which are already held by running workgroups, so the
// grid always makes progress. Each lower tile (i, j) carries a progress counter: the number of
// updates applied, then k+1 once it is final in column k. Waits are bounded: a workgroup that
// spins past the limit raises the abort flag and every workgroup leaves.
//
// Triangular solves: one workgroup per 64-row tile, ticket-ordered; tile b waits for the
// solution blocks it needs (published with release/acquire flags) and accumulates their
// contributions while later blocks are still being solved.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.h"
#include "kernels.h"
#include "kernels_big.h"

namespace mb {

#define LDT (NB + 1)
constexpr long SPIN_LIMIT = 1L << 24;

// Cross-workgroup data moves through agent-scope relaxed atomics: on this part they are
// coherent loads/stores at the device coherence point (sc1), so no L2 writeback or
// invalidate is needed. A flag is stored only after the producer's coherent stores completed
// (s_waitcnt), and consumers read the data with coherent loads after observing the flag.
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
// tile tasks (256 threads)
// ------------------------------------------------------------------------------------------

// load an h×w tile at (r0, c0) of F (leading dimension LD) into LDS (stride LDT), zero padded
__device__ __forceinline__ void load_tile(double* S, const double* F, int LD, int r0, int c0, int h, int w) {
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int i = t % NB, j = t / NB;
    S[i + j * LDT] = (i < h && j < w) ? cload(&F[(size_t)(r0 + i) + (size_t)(c0 + j) * LD]) : 0.0;
  }
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cholesky of the w×w diagonal tile at (c0, c0) of F, in place, by wave 0: lane i holds row i
// in registers (identity padded past w); column j's pivot and row j are broadcast with
// v_readlane, fully unrolled so every register index is static. Returns the pivot status.
__device__ __forceinline__ bool potrf64_wave0(double* F, int LD, int c0, int w) {
  __shared__ int okflag;
  if (threadIdx.x < 64) {
    const int i = threadIdx.x;
    double row[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
      row[k] = (i < w && k < w && k <= i) ? cload(&F[(size_t)(c0 + i) + (size_t)(c0 + k) * LD]) : (i == k ? 1.0 : 0.0);
    int bad = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double s = row[j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= row[k] * readlane_d(row[k], j);
      double d = readlane_d(s, j);
      if (!(d > 0.0) || !isfinite(d)) {
        bad = 1;
        d = 1.0;
      }
      double l = sqrt(d);
      double inv = 1.0 / l;
      row[j] = (i == j) ? l : ((i > j) ? s * inv : row[j]);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (i < w && k < w && k <= i) cstore(&F[(size_t)(c0 + i) + (size_t)(c0 + k) * LD], row[k]);
    if (i == 0) okflag = !bad;
  }
  __syncthreads();
  return okflag != 0;
}

constexpr int LROW = NB + 2;  // row stride of a transposed factor tile in LDS (16-byte aligned rows)

// X = A L⁻ᵀ for the h×w tile at (ri, ck): L (w×w, factored, identity padded) staged row-major
// in LDS and X staged column-major in LDS by all threads (coalesced); wave 0 then holds row i of
// X in registers (lane i), fully unrolled, and the tile goes back through LDS.
__device__ __forceinline__ void trsm64(double* F, int LD, int ri, int h, int ck, int w, double* Lt, double* Xs,
                                       double* Linv) {
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int a = t % NB, b = t / NB;  // L[a][b], column-major read
    Lt[a * LROW + b] = (a < w && b < w && b <= a) ? cload(&F[(size_t)(ck + a) + (size_t)(ck + b) * LD]) : (a == b ? 1.0 : 0.0);
  }
  load_tile(Xs, F, LD, ri, ck, h, w);
  __syncthreads();
  if (threadIdx.x < NB) Linv[threadIdx.x] = 1.0 / Lt[threadIdx.x * LROW + threadIdx.x];
  __syncthreads();
  if (threadIdx.x < 64) {
    const int i = threadIdx.x;
    double x[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) x[k] = Xs[i + k * LDT];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double s = x[j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= x[k] * Lt[j * LROW + k];
      x[j] = s * Linv[j];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) Xs[i + k * LDT] = x[k];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int a = t % NB, b = t / NB;
    if (a < h && b < w) cstore(&F[(size_t)(ri + a) + (size_t)(ck + b) * LD], Xs[a + b * LDT]);
  }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// C (hi×hj at (ri, rj) in F) -= A Bᵀ with A (hi×w), B (hj×w) in LDS; lower part only if diag
__device__ __forceinline__ void mfma_update(double* F, int LD, const double* As, const double* Bs, int ri, int rj, int hi,
                            int hj, int w, bool diag) {
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int qi = (wave >> 1) * 32, qj = (wave & 1) * 32;
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  int li = lane & 15, lk = lane >> 4;
  int kmax = (w + 3) & ~3;
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    double av[2], bv[2];
    for (int a = 0; a < 2; ++a) av[a] = As[(qi + a * 16 + li) + (k0 + lk) * LDT];
    for (int b = 0; b < 2; ++b) bv[b] = Bs[(qj + b * 16 + li) + (k0 + lk) * LDT];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int t = 0; t < 4; ++t) {
        int i = qi + a * 16 + (lane >> 4) + 4 * t;
        int j = qj + b * 16 + (lane & 15);
        if (i < hi && j < hj && (!diag || i >= j)) {
          double* q = &F[(size_t)(ri + i) + (size_t)(rj + j) * LD];
          cstore(q, cload(q) - acc[a][b][t]);
        }
      }
}

// ------------------------------------------------------------------------------------------
// the DAG factorization kernel
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_dag_factor(BigFront B, const InstDev* insts) {
  __shared__ double S0[NB * LROW];  // also the row-major factor tile of TRSM
  __shared__ double S1[NB * LDT];
  __shared__ double Lv[NB];
  __shared__ int task_s;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* F = B.F;
  const int LD = B.ld;
  for (;;) {
    if (threadIdx.x == 0) {
      int t = atomicAdd(B.ticket, 1);
      if (t < B.ntask && ld_flag(B.abortf) != 0) t = B.ntask;
      task_s = t;
    }
    __syncthreads();
    int t = task_s;
    __syncthreads();
    if (t >= B.ntask) break;
    int4 task = B.tasks[t];
    int type = task.x, i = task.y, j = task.z, k = task.w;
    int r0i = B.tb[i], hi = B.tb[i + 1] - r0i;
    int c0k = B.tb[k], wk = B.tb[k + 1] - c0k;
    if (type == 0) {
      // POTRF(k): tile (k, k) has all k updates
      if (!wait_geq(&B.state[tri_index(k, k)], k, B.abortf)) break;
      bool ok = potrf64_wave0(F, LD, c0k, wk);
      if (!ok && threadIdx.x == 0) atomicAdd(&I.sc[S_FAIL], 1.0);
      publish(&B.state[tri_index(k, k)], k + 1);
    } else if (type == 1) {
      // TRSM(i, k)
      if (!wait_geq(&B.state[tri_index(k, k)], k + 1, B.abortf)) break;
      if (!wait_geq(&B.state[tri_index(i, k)], k, B.abortf)) break;
      trsm64(F, LD, r0i, hi, c0k, wk, S0, S1, Lv);
      publish(&B.state[tri_index(i, k)], k + 1);
    } else {
      // UPDATE(i, j, k): C_ij -= L_ik L_jkᵀ
      int r0j = B.tb[j], hj = B.tb[j + 1] - r0j;
      if (!wait_geq(&B.state[tri_index(i, k)], k + 1, B.abortf)) break;
      if (j != i && !wait_geq(&B.state[tri_index(j, k)], k + 1, B.abortf)) break;
      if (!wait_geq(&B.state[tri_index(i, j)], k, B.abortf)) break;
      load_tile(S0, F, LD, r0i, c0k, hi, wk);
      load_tile(S1, F, LD, r0j, c0k, hj, wk);
      __syncthreads();
      mfma_update(F, LD, S0, S1, r0i, r0j, hi, hj, wk, i == j);
      publish(&B.state[tri_index(i, j)], k + 1);
    }
  }
}

__global__ void k_fill_int(int* x, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
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

// forward: tile b (ticket order) accumulates L[b, j] x_j for published j < min(b, Tr), then
// solves its diagonal block (res tiles) or adds into uvec (sep tiles)
__global__ void __launch_bounds__(256) k_big_fsolve(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                    int kv, int epoch, long long base) {
  __shared__ double xs[NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  // every workgroup consumes exactly one ticket (the host advances the base by the grid size)
  if (threadIdx.x == 0) tk = (long long)atomicAdd((unsigned long long*)B.sticket, 1ull) - base;
  __syncthreads();
  if (I.sc[S_ACTIVE] == 0.0) return;
  int b = (int)tk;
  if (b < 0 || b >= B.T) return;
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  const double* F = B.F;
  const int LD = B.ld;
  int r0 = B.tb[b], h = B.tb[b + 1] - r0;
  int row = threadIdx.x & 63, q = threadIdx.x >> 6;
  double acc = 0.0;
  int jmax = b < B.Tr ? b : B.Tr;
  for (int j = 0; j < jmax; ++j) {
    if (!wait_geq(&B.sflag[j], epoch, B.abortf)) return;
    int c0 = B.tb[j], w = B.tb[j + 1] - c0;
    if (threadIdx.x < w) xs[threadIdx.x] = cload(&v[c0 + threadIdx.x]);
    __syncthreads();
    if (row < h)
      for (int c = q; c < w; c += 4) acc += F[(size_t)(r0 + row) + (size_t)(c0 + c) * LD] * xs[c];
    __syncthreads();
  }
  part[q][row] = acc;
  __syncthreads();
  if (b < B.Tr) {
    if (threadIdx.x < 64) {
      int lane = threadIdx.x;
      double t = 0.0;
      if (lane < h) t = v[r0 + lane] - (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
      double Lr[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        Lr[j] = (lane < h && j <= lane) ? F[(size_t)(r0 + lane) + (size_t)(r0 + j) * LD] : 0.0;
      double xi = t;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < h) {
          double xj = __shfl(xi / Lr[j], j, 64);
          if (lane == j) xi = xj;
          if (lane > j) xi -= Lr[j] * xj;
        }
      }
      if (lane < h) cstore(&v[r0 + lane], xi);
    }
    publish(&B.sflag[b], epoch);
  } else {
    if (threadIdx.x < h) {
      int lane = threadIdx.x;
      Fd.uvec[r0 - Fd.r + lane] += part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
    }
  }
}

// backward: column tile b (ticket order descending) accumulates L[j, b]ᵀ x_j for j > b and
// L21[:, b]ᵀ x_sep, then solves L_bbᵀ x_b
__global__ void __launch_bounds__(256) k_big_bsolve(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                    int kv, int epoch, long long base) {
  __shared__ double S[NB * LDT];
  __shared__ double xs[NB];
  __shared__ double part[4][NB];
  __shared__ long long tk;
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (threadIdx.x == 0) tk = (long long)atomicAdd((unsigned long long*)B.sticket, 1ull) - base;
  __syncthreads();
  if (I.sc[S_ACTIVE] == 0.0) return;
  int tkt = (int)tk;
  if (tkt < 0 || tkt >= B.Tr) return;
  int b = B.Tr - 1 - tkt;
  double* vall = I.nv + (size_t)kv * I.n;
  double* v = vall + Fd.res0;
  const double* F = B.F;
  const int LD = B.ld;
  int c0 = B.tb[b], w = B.tb[b + 1] - c0;
  int col = threadIdx.x & 63, q = threadIdx.x >> 6;
  double acc = 0.0;
  // separator part (rows ≥ r; x_sep is final), 64-row chunks through LDS so that the global
  // reads run down the columns
  for (int rs0 = Fd.r; rs0 < LD; rs0 += NB) {
    int h = min(NB, LD - rs0);
    load_tile(S, F, LD, rs0, c0, h, w);
    if (threadIdx.x < h) xs[threadIdx.x] = vall[Fd.sepcols[rs0 - Fd.r + threadIdx.x]];
    __syncthreads();
    if (col < w)
      for (int rr = q; rr < h; rr += 4) acc += S[rr + col * LDT] * xs[rr];
    __syncthreads();
  }
  for (int j = B.Tr - 1; j > b; --j) {
    if (!wait_geq(&B.sflag[j], epoch, B.abortf)) return;
    int r0 = B.tb[j], h = B.tb[j + 1] - r0;
    load_tile(S, F, LD, r0, c0, h, w);
    if (threadIdx.x < h) xs[threadIdx.x] = cload(&v[r0 + threadIdx.x]);
    __syncthreads();
    if (col < w)
      for (int rr = q; rr < h; rr += 4) acc += S[rr + col * LDT] * xs[rr];
    __syncthreads();
  }
  part[q][col] = acc;
  load_tile(S, F, LD, c0, c0, w, w);
  __syncthreads();
  if (threadIdx.x < 64) {
    int lane = threadIdx.x;
    double xi = 0.0;
    if (lane < w) xi = v[c0 + lane] - (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
    // solve L_bbᵀ x = t bottom-up: lane i holds column i of the block (from LDS)
    for (int j = w - 1; j >= 0; --j) {
      double dj = S[j + j * LDT];
      double xj = __shfl(xi / dj, j, 64);
      if (lane == j) xi = xj;
      if (lane < j) xi -= S[j + lane * LDT] * xj;
    }
    if (lane < w) cstore(&v[c0 + lane], xi);
  }
  publish(&B.sflag[b], epoch);
}

// ------------------------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------------------------

void launch_dag_factor(hipStream_t st, const BigFront& B, const InstDev* I, int grid) {
  k_fill_int<<<(B.nstate + 255) / 256, 256, 0, st>>>(B.state, B.nstate, 0);
  k_fill_int<<<1, 64, 0, st>>>(B.ticket, 2, 0);  // ticket and abort flag
  k_dag_factor<<<grid, 256, 0, st>>>(B, I);
}

void launch_big_fsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const int* childlist,
                       const InstDev* I, int kv, int epoch, long long base) {
  k_big_gather<<<1, 256, 0, st>>>(F, f, childlist, I, kv);
  k_big_fsolve<<<B.T, 256, 0, st>>>(B, F, f, I, kv, epoch, base);
}

void launch_big_bsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const InstDev* I, int kv,
                       int epoch, long long base) {
  k_big_bsolve<<<B.Tr, 256, 0, st>>>(B, F, f, I, kv, epoch, base);
}

}  // namespace mb
