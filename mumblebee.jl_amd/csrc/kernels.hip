// HIP kernels for the MI355X conic IPM solve path (gfx950, 64-wide wavefronts).
//
//  * vector algebra over per-instance tile lists, deterministic two-pass reductions;
//  * CSR sparse matrix-vector products for B, Bᵀ and Q;
//  * cone kernels (nonnegative, free, second-order): NT scaling blocks, correctors, max steps;
//  * assembly of δH + BᵀB into dense frontal matrices;
//  * batched blocked right-looking partial Cholesky of frontal matrices (FP64 MFMA trailing
//    update), extend-add of child update matrices;
//  * per-front forward / backward triangular solves.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdint>

#include "device.h"
#include "kernels.h"
#include "tdc.h"
#include "tile64.h"

namespace mb {

#define TILE TILE_THREADS

__device__ __forceinline__ double coefval(const Coef& c, const double* sc) {
  if (c.slot < 0) return c.c;
  double v = sc[c.slot];
  return c.inv ? c.c / v : c.c * v;
}

__device__ __forceinline__ double* nvec(const InstDev& I, int kind) { return I.nv + (size_t)kind * I.n; }
__device__ __forceinline__ double* mvec(const InstDev& I, int kind) { return I.mv + (size_t)kind * I.m; }

__device__ __forceinline__ double* vecof(const InstDev& I, int space, int kind) {
  return space == 0 ? nvec(I, kind) : mvec(I, kind);
}

__device__ __forceinline__ int spacelen(const InstDev& I, int space) { return space == 0 ? I.n : I.m; }

// ------------------------------------------------------------------------------------------
// vector algebra
// ------------------------------------------------------------------------------------------

// z = a x + b y   (x, y, z of one space; kinds may alias)
__global__ void k_axpby(const InstDev* insts, const Tile* tiles, int space, int kx, int ky, int kz,
                        Coef a, Coef b) {
  Tile t = tiles[blockIdx.x];
  const InstDev& I = insts[t.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  int i = t.start + threadIdx.x;
  int n = spacelen(I, space);
  if (i >= n) return;
  double av = coefval(a, I.sc), bv = coefval(b, I.sc);
  const double* x = vecof(I, space, kx);
  const double* y = vecof(I, space, ky);
  double* z = vecof(I, space, kz);
  double yv = (bv == 0.0) ? 0.0 : bv * y[i];
  z[i] = av * x[i] + yv;
}

// z = a x + b y + c w
__global__ void k_axpbypcz(const InstDev* insts, const Tile* tiles, int space, int kx, int ky, int kw,
                           int kz, Coef a, Coef b, Coef c) {
  Tile t = tiles[blockIdx.x];
  const InstDev& I = insts[t.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  int i = t.start + threadIdx.x;
  int n = spacelen(I, space);
  if (i >= n) return;
  double av = coefval(a, I.sc), bv = coefval(b, I.sc), cv = coefval(c, I.sc);
  const double* x = vecof(I, space, kx);
  const double* y = vecof(I, space, ky);
  const double* w = vecof(I, space, kw);
  double* z = vecof(I, space, kz);
  double acc = av * x[i];
  if (bv != 0.0) acc += bv * y[i];
  if (cv != 0.0) acc += cv * w[i];
  z[i] = acc;
}


__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  return s;
}

// partial[tile] = Σ x⊙y (mode 0), Σ (x/s)² with s = kind ky (mode 1), Σ x (mode 2),
// Σ (x + τ dx)(y + τ dy) with τ = S_TAU (mode 3: x=P, y=D, extra kinds in kw0/kw1), Σ (x − y)²
// (mode 4)
__global__ void k_dot_partial(const InstDev* insts, const Tile* tiles, int space, int kx, int ky,
                              int mode, int kw0, int kw1, double* partial) {
  __shared__ double red[TILE / 64];
  Tile t = tiles[blockIdx.x];
  const InstDev& I = insts[t.inst];
  int i = t.start + threadIdx.x;
  double v = 0.0;
  if (I.sc[S_ACTIVE] != 0.0 && i < spacelen(I, space)) {
    const double* x = vecof(I, space, kx);
    if (mode == 0) {
      v = x[i] * vecof(I, space, ky)[i];
    } else if (mode == 1) {
      double q = x[i] / vecof(I, space, ky)[i];
      v = q * q;
    } else if (mode == 2) {
      v = x[i];
    } else if (mode == 4) {
      const double q = x[i] - vecof(I, space, ky)[i];
      v = q * q;
    } else {
      double tau = I.sc[S_TAU];
      double a = x[i] + tau * vecof(I, space, kw0)[i];
      double b = vecof(I, space, ky)[i] + tau * vecof(I, space, kw1)[i];
      v = a * b;
    }
  }
  double s = block_sum(v, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// The same partial sums, and in the last block of an instance to finish, sc[slot] = op(Σ partial)
// over the instance's tiles in k_reduce_slot's order (identical result, one launch instead of
// two). counter[inst] counts the instance's finished blocks; the last block puts it back to 0.
__global__ void k_dot(const InstDev* insts, const Tile* tiles, const int* tile_ofs, int space, int kx, int ky,
                      int mode, int kw0, int kw1, double* partial, int* counter, int slot, int op) {
  __shared__ double red[TILE / 64];
  __shared__ int last;
  const Tile t = tiles[blockIdx.x];
  const InstDev& I = insts[t.inst];
  const int i = t.start + threadIdx.x;
  double v = 0.0;
  if (I.sc[S_ACTIVE] != 0.0 && i < spacelen(I, space)) {
    const double* x = vecof(I, space, kx);
    if (mode == 0) {
      v = x[i] * vecof(I, space, ky)[i];
    } else if (mode == 1) {
      const double q = x[i] / vecof(I, space, ky)[i];
      v = q * q;
    } else if (mode == 2) {
      v = x[i];
    } else if (mode == 4) {
      const double q = x[i] - vecof(I, space, ky)[i];
      v = q * q;
    } else {
      const double tau = I.sc[S_TAU];
      const double a = x[i] + tau * vecof(I, space, kw0)[i];
      const double b = vecof(I, space, ky)[i] + tau * vecof(I, space, kw1)[i];
      v = a * b;
    }
  }
  const double s = block_sum(v, red);
  const int a = tile_ofs[t.inst], b = tile_ofs[t.inst + 1];
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partial[blockIdx.x], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();  // the partial is visible device-wide before the count
    last = atomicAdd(&counter[t.inst], 1) == b - a - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double w = 0.0;
  for (int k = a + (int)threadIdx.x; k < b; k += blockDim.x)
    w += __hip_atomic_load(&partial[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double r = block_sum(w, red);
  if (threadIdx.x == 0) {
    if (I.sc[S_ACTIVE] != 0.0) I.sc[slot] = op == 1 ? sqrt(r) : r;
    __hip_atomic_store(&counter[t.inst], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The dot products of a batch, one workgroup each over the whole vector, summed in a fixed order
// (thread-strided, then the wave tree, then the waves in order).
constexpr int DOTS_THREADS = 1024;
__global__ void __launch_bounds__(DOTS_THREADS) k_dots(const InstDev* insts, int inst, DotBatch B) {
  __shared__ double red[DOTS_THREADS / 64];
  const InstDev& I = insts[inst + (int)blockIdx.y];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const DotSpec d = B.s[blockIdx.x];
  const int len = spacelen(I, d.space);
  const double* x = vecof(I, d.space, d.kx);
  const double* y = vecof(I, d.space, d.ky);
  double v = 0.0;
  if (d.mode == 0) {
    for (int i = threadIdx.x; i < len; i += DOTS_THREADS) v += x[i] * y[i];
  } else if (d.mode == 1) {
    for (int i = threadIdx.x; i < len; i += DOTS_THREADS) {
      const double q = x[i] / y[i];
      v += q * q;
    }
  } else if (d.mode == 2) {
    for (int i = threadIdx.x; i < len; i += DOTS_THREADS) v += x[i];
  } else if (d.mode == 4) {
    for (int i = threadIdx.x; i < len; i += DOTS_THREADS) {
      const double q = x[i] - y[i];
      v += q * q;
    }
  } else {
    const double tau = I.sc[S_TAU];
    const double* dx = vecof(I, d.space, d.kw0);
    const double* dy = vecof(I, d.space, d.kw1);
    for (int i = threadIdx.x; i < len; i += DOTS_THREADS) v += (x[i] + tau * dx[i]) * (y[i] + tau * dy[i]);
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < DOTS_THREADS / 64; ++k) s += red[k];
    I.sc[d.slot] = d.op == 1 ? sqrt(s) : s;
  }
}

// z = a x + b y + c w for each op of the set, in order: blocks [0, nnt) on the n-space tiles with
// N, the rest on the m-space tiles with M
__global__ void k_lin(const InstDev* insts, const Tile* ntiles, int nnt, const Tile* mtiles, LinSet N, LinSet M) {
  const bool isn = (int)blockIdx.x < nnt;
  const Tile t = isn ? ntiles[blockIdx.x] : mtiles[blockIdx.x - nnt];
  const InstDev& I = insts[t.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int space = isn ? 0 : 1;
  const int i = t.start + threadIdx.x;
  if (i >= spacelen(I, space)) return;
  const int nop = isn ? N.n : M.n;
  for (int k = 0; k < nop; ++k) {
    const Lin L = isn ? N.op[k] : M.op[k];
    const double av = coefval(L.a, I.sc), bv = coefval(L.b, I.sc), cv = coefval(L.c, I.sc);
    double acc = av * vecof(I, space, L.kx)[i];
    if (bv != 0.0) acc += bv * vecof(I, space, L.ky)[i];
    if (cv != 0.0) acc += cv * vecof(I, space, L.kw)[i];
    vecof(I, space, L.kz)[i] = acc;
  }
}

// sc[slot] = op(Σ partial over the instance's tiles), one block per instance, fixed order
__global__ void k_reduce_slot(const InstDev* insts, const int* tile_ofs, const double* partial, int slot,
                              int op) {
  __shared__ double red[TILE / 64];
  int inst = blockIdx.x;
  const InstDev& I = insts[inst];
  int a = tile_ofs[inst], b = tile_ofs[inst + 1];
  double v = 0.0;
  for (int k = a + (int)threadIdx.x; k < b; k += blockDim.x) v += partial[k];
  double s = block_sum(v, red);
  if (threadIdx.x == 0 && I.sc[S_ACTIVE] != 0.0) I.sc[slot] = op == 1 ? sqrt(s) : s;
}

// ------------------------------------------------------------------------------------------
// sparse matrix-vector products: y = a A x + b y with A ∈ {B (m rows), Bᵀ (n rows), Q (n rows)}
// ------------------------------------------------------------------------------------------

// Each tile is 256/G consecutive rows of one instance's matrix; G = A.lanes lanes share a row,
// summing a lane-strided slice of it, and a fixed xor tree combines the G partial sums (the
// result does not depend on timing). Every lane runs the same trip count of the row loop so
// the shuffles see whole groups.
// row i of A against x by the whole workgroup (thread-strided, then the fixed block tree); the
// sum is in thread 0
__device__ __forceinline__ double csr_row_block(const CSRDev& A, int i, const double* x, double* red) {
  double acc = 0.0;
  for (int k = A.ptr[i] + (int)threadIdx.x; k < A.ptr[i + 1]; k += (int)blockDim.x) acc += A.val[k] * x[A.col[k]];
  return block_sum(acc, red);
}

// whether row i of A is left to a workgroup of its own (see SPMV_LONG)
__device__ __forceinline__ bool csr_long(const CSRDev& A, int i) {
  return A.ptr[i + 1] - A.ptr[i] > SPMV_LONG * A.lanes;
}

__global__ void __launch_bounds__(TILE) k_spmv(const InstDev* insts, const Tile* tiles, int which, int kx, int ky,
                                               Coef a, Coef b) {
  __shared__ double red[TILE / 64];
  const Tile t = tiles[blockIdx.x];
  const InstDev& I = insts[t.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const CSRDev& A = which == 0 ? I.B : (which == 1 ? I.Bt : I.Q);
  const int xspace = which == 0 ? 0 : (which == 1 ? 1 : 0);
  const int yspace = which == 0 ? 1 : 0;
  const int G = A.lanes;
  const double* x = vecof(I, xspace, kx);
  double* y = vecof(I, yspace, ky);
  if (t.start < 0) {  // one long row
    const int i = -1 - t.start;
    const double acc = csr_row_block(A, i, x, red);
    if (threadIdx.x == 0) {
      const double av = coefval(a, I.sc), bv = coefval(b, I.sc);
      y[i] = av * acc + (bv == 0.0 ? 0.0 : bv * y[i]);
    }
    return;
  }
  const int lane = threadIdx.x & (G - 1);
  const int i = t.start + (int)threadIdx.x / G;
  const bool ok = i < A.nrow && !csr_long(A, i);
  double acc = 0.0;
  const int k0 = ok ? A.ptr[i] : 0, k1 = ok ? A.ptr[i + 1] : 0;
  for (int k = k0 + lane; k < k1; k += G) acc += A.val[k] * x[A.col[k]];
  for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (ok && lane == 0) {
    const double av = coefval(a, I.sc), bv = coefval(b, I.sc);
    y[i] = av * acc + (bv == 0.0 ? 0.0 : bv * y[i]);
  }
}

// row i of A against x by a group of G lanes: a lane-strided slice each, then the fixed xor tree
// (every lane of the group reaches the shuffles)
__device__ __forceinline__ double csr_row(const CSRDev& A, int i, bool ok, int lane, int G, const double* x) {
  double acc = 0.0;
  const int k0 = ok ? A.ptr[i] : 0, k1 = ok ? A.ptr[i + 1] : 0;
  for (int k = k0 + lane; k < k1; k += G) acc += A.val[k] * x[A.col[k]];
  for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  return acc;
}

__device__ __forceinline__ double coefterm(const Coef& c, const double* sc, const double* v, int i) {
  const double cv = coefval(c, sc);
  return cv == 0.0 ? 0.0 : cv * v[i];
}

// blocks [0, nn): n rows in the tiles of matrix ntm; the rest: m rows in the tiles of B. A long-row
// tile (start < 0) sums its row with the whole workgroup; thread 0 then writes.
__global__ void __launch_bounds__(TILE) k_spmv_nm(const InstDev* insts, int ntm, const Tile* ntiles, int nn,
                                                  const Tile* mtiles, SpmvN N, SpmvM M) {
  __shared__ double red[TILE / 64];
  const bool isn = (int)blockIdx.x < nn;
  const Tile t = isn ? ntiles[blockIdx.x] : mtiles[blockIdx.x - nn];
  const InstDev& I = insts[t.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const bool lrow = t.start < 0;
  if (isn) {
    const CSRDev& At = ntm == 2 ? I.Q : I.Bt;
    const int G = At.lanes;
    int lane, i;
    bool ok;
    double sq = 0.0, st = 0.0;
    if (lrow) {
      lane = (int)threadIdx.x;
      i = -1 - t.start;
      ok = true;
      if (N.kxq >= 0) sq = csr_row_block(I.Q, i, nvec(I, N.kxq), red);
      if (N.kxt >= 0) st = csr_row_block(I.Bt, i, mvec(I, N.kxt), red);
    } else {
      lane = threadIdx.x & (G - 1);
      i = t.start + (int)threadIdx.x / G;
      ok = i < At.nrow && !csr_long(At, i);
      if (N.kxq >= 0) sq = csr_row(I.Q, i, ok, lane, G, nvec(I, N.kxq));
      if (N.kxt >= 0) st = csr_row(I.Bt, i, ok, lane, G, mvec(I, N.kxt));
    }
    if (ok && lane == 0) {
      double y = 0.0;
      if (N.kxq >= 0) y += coefval(N.aq, I.sc) * sq;
      if (N.kxt >= 0) y += coefval(N.at, I.sc) * st;
      if (N.kz >= 0) y += coefterm(N.c, I.sc, nvec(I, N.kz), i);
      if (N.kw >= 0) y += coefterm(N.d, I.sc, nvec(I, N.kw), i);
      if (N.kv >= 0) y += coefterm(N.e, I.sc, nvec(I, N.kv), i);
      nvec(I, N.ky)[i] = y;
      if (N.kyq >= 0) nvec(I, N.kyq)[i] = sq;
    }
  } else {
    const int G = I.B.lanes;
    int lane, i;
    bool ok;
    double s;
    if (lrow) {
      lane = (int)threadIdx.x;
      i = -1 - t.start;
      ok = true;
      s = csr_row_block(I.B, i, nvec(I, M.kx), red);
    } else {
      lane = threadIdx.x & (G - 1);
      i = t.start + (int)threadIdx.x / G;
      ok = i < I.B.nrow && !csr_long(I.B, i);
      s = csr_row(I.B, i, ok, lane, G, nvec(I, M.kx));
    }
    if (ok && lane == 0) {
      double y = coefval(M.a, I.sc) * s;
      if (M.kz >= 0) y += coefterm(M.c, I.sc, mvec(I, M.kz), i);
      mvec(I, M.ky)[i] = y;
      if (M.ky2 >= 0) {
        double* y2 = mvec(I, M.ky2) + i;
        const double ey = coefval(M.e, I.sc) * y;
        *y2 = M.y2set ? ey : *y2 + ey;
      }
    }
  }
}

// y[cols] += a D_v x[cols] over all cones (the cone-block part of H x)
__global__ void k_cone_matvec(const InstDev* insts, const ConeDesc* cones, int ncone, int kx, int ky, Coef a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const double* x = nvec(I, kx) + C.col0;
  double* y = nvec(I, ky) + C.col0;
  double av = coefval(a, I.sc);
  if (C.kind == CONE_POS) {  // dense blocks (second-order, PSD): k_cone_matvec_dense
    for (int i = 0; i < C.n; ++i) y[i] += av * C.D[i] * x[i];
  } else if (C.kind == CONE_EXP || C.kind == CONE_POW) {
    const double* D = C.D;
    const double x0 = x[0], x1 = x[1], x2 = x[2];
    for (int i = 0; i < 3; ++i) y[i] += av * (D[i] * x0 + D[i + 3] * x1 + D[i + 6] * x2);
  }
}

// ------------------------------------------------------------------------------------------
// error-free transformations (src/IPM/src/utils.jl:27-53)
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ void twosum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void twoprod(double a, double b, double& p, double& e) {
  p = a * b;
  e = fma(a, b, -p);
}

__device__ double cdot(const double* p, const double* d, int n) {
  double s = 0.0, c = 0.0;
  for (int i = 0; i < n; ++i) {
    double pr, e, e2;
    twoprod(p[i], d[i], pr, e);
    twosum(s, pr, s, e2);
    c += e + e2;
  }
  return s + c;
}

// compensated J-inner product x₁y₁ − Σ xᵢyᵢ as (hi, lo) (soc.jl:48-61)
__device__ void socdot(const double* x, const double* y, int n, double& hi, double& lo) {
  double s, c;
  twoprod(x[0], y[0], s, c);
  for (int i = 1; i < n; ++i) {
    double p, e, e2;
    twoprod(x[i], y[i], p, e);
    twosum(s, -p, s, e2);
    c += -e + e2;
  }
  twosum(s, c, hi, lo);
}

// √H x (inverse=0) or √H y = x (inverse=1), in place (soc.jl:82-110)
__device__ void socroot(double* x, const double* w, double beta, int n, int inverse) {
  double sigma = inverse ? 1.0 : -1.0;
  double alpha = inverse ? beta : 1.0 / beta;
  double w1p1 = w[0] + 1.0;
  double wpdx = w1p1 * x[0];
  for (int i = 1; i < n; ++i) wpdx = fma(sigma * w[i], x[i], wpdx);
  wpdx /= w1p1;
  x[0] = alpha * fma(wpdx, w1p1, -x[0]);
  for (int i = 1; i < n; ++i) x[i] = alpha * fma(sigma * wpdx, w[i], x[i]);
}

// ------------------------------------------------------------------------------------------
// cone kernels: one thread per cone (SOC caches: β, pdet, ddet, pdetlo, ddetlo, w[n])
// ------------------------------------------------------------------------------------------

// D ← cone scaling block; per-cone (flag, spsd) into out[2c], out[2c+1]
__global__ void k_cone_scale(const InstDev* insts, const ConeDesc* cones, int ncone, double* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  if (C.kind == CONE_PSD || C.kind == CONE_SOC) return;  // k_psd_scale, k_soc_scale
  const double* p = nvec(I, NV_P) + C.col0;
  const double* d = nvec(I, NV_D) + C.col0;
  int n = C.n;
  double flag = 1.0, spsd = 0.0;
  if (C.kind == CONE_POS) {
    // pos.jl:75-84
    for (int i = 0; i < n; ++i) {
      C.D[i] = d[i] / p[i];
      spsd += 1.0 / (p[i] * d[i]);
    }
  } else if (C.kind == CONE_EXP || C.kind == CONE_POW) {
    // tdscale! (tdc.jl:185-270)
    flag = tdc::scale(C.kind, C.param, p, d, C.cache, C.D, spsd) ? 1.0 : 0.0;
  }
  out[2 * c] = flag;
  out[2 * c + 1] = spsd;
}

// socscale! (soc.jl:192-263), one wavefront per second-order cone: lane 0 forms the compensated
// determinants, ⟨p, d⟩, β and κ in the reference's order (soc.jl:48-70), the lanes then form w
// and write the n×n block D = η(2 w wᵀ − J) column by column (coalesced along each column).
// SOC caches: β, pdet, ddet, pdetlo, ddetlo, w[n]. out[2c] = flag, out[2c+1] = ⟨p*, d*⟩.
constexpr int SOC_LDS = 4096;  // w staged in LDS up to this order
__global__ void __launch_bounds__(64) k_soc_scale(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                  double* out) {
  __shared__ double ws[SOC_LDS];
  __shared__ double sh[3];
  __shared__ int okf;
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int lane = threadIdx.x, n = C.n;
  const double* p = nvec(I, NV_P) + C.col0;
  const double* d = nvec(I, NV_D) + C.col0;
  double* cache = C.cache;
  double* w = cache + 5;
  if (lane == 0) {
    double pdet, pdetlo, ddet, ddetlo;
    socdot(p, p, n, pdet, pdetlo);
    socdot(d, d, n, ddet, ddetlo);
    cache[1] = pdet;
    cache[2] = ddet;
    cache[3] = pdetlo;
    cache[4] = ddetlo;
    const bool ok = pdet > 0 && p[0] > 0 && ddet > 0 && d[0] > 0;
    double beta = 0.0, spsd = 0.0;
    if (ok) {
      const double pdot = cdot(p, d, n);
      const double spdet = sqrt(pdet), sddet = sqrt(ddet);
      sh[1] = sqrt(2.0 * fma(spdet, sddet, pdot));  // κ
      beta = sqrt(spdet / sddet);
      spsd = 4.0 * pdot / (pdet * ddet);
    }
    cache[0] = beta;
    sh[0] = beta;
    okf = ok ? 1 : 0;
    out[2 * c] = ok ? 1.0 : 0.0;
    out[2 * c + 1] = spsd;
  }
  __syncthreads();
  if (!okf) return;
  const double beta = sh[0], kappa = sh[1];
  for (int i = lane; i < n; i += 64) {
    const double wi = i == 0 ? fma(d[0], beta, p[0] / beta) / kappa : fma(-d[i], beta, p[i] / beta) / kappa;
    w[i] = wi;
    if (i < SOC_LDS) ws[i] = wi;
  }
  __syncthreads();
  const double eta = 1.0 / (beta * beta), eta2 = 2.0 * eta;
  const double w1 = ws[0], nw1 = -eta2 * w1;
  for (int j = 0; j < n; ++j) {
    const double wj = j < SOC_LDS ? ws[j] : w[j];
    double* Dj = C.D + (size_t)j * n;
    for (int i = lane; i < n; i += 64) {
      const double wi = i < SOC_LDS ? ws[i] : w[i];
      double v;
      if (i == 0 && j == 0) v = fma(eta2, w1 * w1, -eta);
      else if (i == 0) v = nw1 * wj;
      else if (j == 0) v = nw1 * wi;
      else if (i == j) v = fma(eta2, wj * wj, eta);
      else v = i > j ? (eta2 * wj) * wi : (eta2 * wi) * wj;
      Dj[i] = v;
    }
  }
}

// dualshadow! (cone.jl:556-566): y = −f′(x) per cone, one thread per cone — nonnegative 1/x
// (pos.jl), free 0 (noc.jl), second-order 2 J x / det(x) with the compensated determinant
// (soc.jl:488-504), exponential / power by the Halley shadow solve seeded from the cache (exp.jl,
// pow.jl); PSD cones: k_psd_dshadow and the large-cone pipeline. y[cols] stays untouched where
// the cone's point is outside the cone (the reference's shadow is then not defined).
__global__ void k_cone_dshadow(const InstDev* insts, const ConeDesc* cones, int ncone, int kx, int ky) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || C.kind == CONE_PSD) return;
  const double* x = nvec(I, kx) + C.col0;
  double* y = nvec(I, ky) + C.col0;
  const int n = C.n;
  if (C.kind == CONE_POS) {
    for (int i = 0; i < n; ++i) y[i] = 1.0 / x[i];
  } else if (C.kind == CONE_FREE) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
  } else if (C.kind == CONE_SOC) {
    double det, lo;
    socdot(x, x, n, det, lo);
    det += lo;
    if (det > 0.0 && x[0] > 0.0) {
      const double gm = 2.0 / det;
      y[0] = gm * x[0];
      for (int i = 1; i < n; ++i) y[i] = -gm * x[i];
    }
  } else {
    tdc::dual_shadow(C.kind, C.param, y, x, C.cache);
  }
}

// ‖Q_vv + D_v‖² per cone into out[c] (the diagonal-block part of ‖H‖)
__global__ void k_cone_hnorm2(const InstDev* insts, const ConeDesc* cones, int ncone, double* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  if (C.kind == CONE_SOC || C.kind == CONE_PSD) return;  // k_cone_hnorm2_dense
  int n = C.n;
  double s = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double h = C.Qd ? C.Qd[i + (size_t)j * n] : 0.0;
      if (C.kind == CONE_POS && i == j) h += C.D[i];
      if (C.kind == CONE_EXP || C.kind == CONE_POW) h += C.D[i + (size_t)j * n];
      s += h * h;
    }
  out[c] = s;
}

// per-instance sum of out[c] into sc[slot] (one block per instance, fixed-order tree)
__global__ void k_cone_sum(const InstDev* insts, const int* cone_ofs, const double* out, int slot) {
  __shared__ double red[TILE / 64];
  int inst = blockIdx.x;
  const InstDev& I = insts[inst];
  double s = 0.0;
  for (int c = cone_ofs[inst] + (int)threadIdx.x; c < cone_ofs[inst + 1]; c += blockDim.x) s += out[c];
  s = block_sum(s, red);
  if (threadIdx.x == 0 && I.sc[S_ACTIVE] != 0.0) I.sc[slot] = s;
}

__device__ __forceinline__ double block_min(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = v;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmin(m, red[k]);
  return m;
}

// per-instance reduction of cone outputs: S_FLAGOK = all(flag), S_SPSD = Σ spsd (mode 0);
// S_TAU = min(step) (mode 1, out[c] per cone)
__global__ void k_cone_reduce(const InstDev* insts, const int* cone_ofs, const double* out, int mode) {
  __shared__ double red[TILE / 64];
  int inst = blockIdx.x;
  const InstDev& I = insts[inst];
  int a = cone_ofs[inst], b = cone_ofs[inst + 1];
  if (mode == 0) {
    double bad = 0.0, s = 0.0;
    for (int c = a + (int)threadIdx.x; c < b; c += blockDim.x) {
      bad += out[2 * c] == 0.0 ? 1.0 : 0.0;
      s += out[2 * c + 1];
    }
    s = block_sum(s, red);
    bad = block_sum(bad, red);
    if (threadIdx.x == 0 && I.sc[S_ACTIVE] != 0.0) {
      I.sc[S_FLAGOK] = bad == 0.0 ? 1.0 : 0.0;
      I.sc[S_SPSD] = s;
    }
  } else {
    double tau = 1.0;
    for (int c = a + (int)threadIdx.x; c < b; c += blockDim.x) tau = fmin(tau, out[c]);
    tau = block_min(tau, red);
    if (threadIdx.x == 0 && I.sc[S_ACTIVE] != 0.0) I.sc[S_TAU] = tau;
  }
}

// predictor right-hand side r = −d + σμ(−f′(p)) per cone (corr0!: pos.jl:131-137,
// noc.jl:303-306, soc.jl:362-379); writes kind kr
__global__ void k_cone_corr0(const InstDev* insts, const ConeDesc* cones, int ncone, int kr) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  if (C.kind == CONE_PSD) return;  // k_psd_corr0
  const double* p = nvec(I, NV_P) + C.col0;
  const double* d = nvec(I, NV_D) + C.col0;
  double* r = nvec(I, kr) + C.col0;
  double smu = I.sc[S_SMU];
  int n = C.n;
  if (C.kind == CONE_POS) {
    for (int i = 0; i < n; ++i) r[i] = smu / p[i] - d[i];
  } else if (C.kind == CONE_FREE) {
    for (int i = 0; i < n; ++i) r[i] = 0.0;
  } else if (C.kind == CONE_SOC) {
    const double* cache = C.cache;
    double beta = cache[0], pdet = cache[1];
    const double* w = cache + 5;
    double ib = 1.0 / beta;
    double eta = ib * ib;
    double wp, wplo;
    socdot(w, p, n, wp, wplo);
    double t = 2.0 * eta * wp;
    double c1 = 2.0 * smu / pdet + eta;
    r[0] = fma(c1, p[0], -t * w[0]);
    for (int i = 1; i < n; ++i) r[i] = fma(t, w[i], -c1 * p[i]);
  } else if (C.kind == CONE_EXP || C.kind == CONE_POW) {
    tdc::corr0(r, d, C.cache, smu);
  }
}

// Mehrotra corrector r = −d + (σμ e − Δp∘Δd)/p (corr!: pos.jl:107-120, noc.jl:285-297,
// soc.jl:282-351). Needs 3n doubles of scratch per SOC cone: scratch + 3*col0.
__global__ void k_cone_corr(const InstDev* insts, const ConeDesc* cones, int ncone, int kr, int kdp, int kdd) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const double* p = nvec(I, NV_P) + C.col0;
  const double* d = nvec(I, NV_D) + C.col0;
  const double* dp0 = nvec(I, kdp) + C.col0;
  const double* dd0 = nvec(I, kdd) + C.col0;
  double* r = nvec(I, kr) + C.col0;
  double smu = I.sc[S_SMU];
  int n = C.n;
  if (C.kind == CONE_POS) {
    for (int i = 0; i < n; ++i) r[i] = (smu - dp0[i] * dd0[i]) / p[i] - d[i];
  } else if (C.kind == CONE_FREE) {
    for (int i = 0; i < n; ++i) r[i] = 0.0;
  } else if (C.kind == CONE_SOC) {
    const double* cache = C.cache;
    double beta = cache[0], pdet = cache[1];
    const double* w = cache + 5;
    double* dp = I.scratch + 3 * (size_t)C.col0;
    double* dd = dp + n;
    double* v = dd + n;
    for (int i = 0; i < n; ++i) {
      dp[i] = dp0[i];
      dd[i] = dd0[i];
    }
    socroot(dp, w, beta, n, 0);
    socroot(dd, w, beta, n, 1);
    double dp1 = dp[0], dd1 = dd[0];
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += dp[i] * dd[i];
    r[0] = acc;
    for (int i = 1; i < n; ++i) r[i] = fma(dd1, dp[i], dp1 * dd[i]);
    for (int i = 0; i < n; ++i) v[i] = p[i];
    socroot(v, w, beta, n, 0);
    double vdet, vdetlo, vr, vrlo;
    socdot(v, v, n, vdet, vdetlo);
    double v1 = v[0];
    socdot(v, r, n, vr, vrlo);
    double r1 = vr / vdet;
    r[0] = r1;
    double iv1 = 1.0 / v1;
    for (int i = 1; i < n; ++i) r[i] = fma(-r1, v[i], r[i]) * iv1;
    for (int i = 0; i < n; ++i) r[i] += v[i];
    socroot(r, w, beta, n, 0);
    double gamma = 2.0 * smu / pdet;
    r[0] = fma(gamma, p[0], -r[0]);
    for (int i = 1; i < n; ++i) r[i] = fma(-gamma, p[i], -r[i]);
  } else if (C.kind == CONE_EXP || C.kind == CONE_POW) {
    tdc::corr(C.kind, C.param, r, p, d, dp0, dd0, C.cache, smu);
  }
}

__device__ double posmaxstep(const double* x, const double* dx, int n) {
  double tau = 1.0;
  for (int i = 0; i < n; ++i)
    if (dx[i] < 0) tau = fmin(tau, -x[i] / dx[i]);
  return tau;
}

// soc.jl:391-470
__device__ double socmaxstep(const double* x, const double* dx, int n, double c, double clo) {
  const double eps = 2.220446049250313e-16;
  double a, alo, b, blo;
  socdot(dx, dx, n, a, alo);
  socdot(x, dx, n, b, blo);
  b *= 2.0;
  blo *= 2.0;
  double tau = 1.0;
  if (fabs(a) > eps * fmax(fabs(b), fabs(c))) {
    double bb, bblo, ac, aclo, s0, s0lo;
    twoprod(b, b, bb, bblo);
    bblo = fma(2.0 * b, blo, bblo);
    twoprod(4.0 * a, c, ac, aclo);
    aclo = fma(4.0 * a, clo, fma(4.0 * alo, c, aclo));
    twosum(bb, -ac, s0, s0lo);
    double disc = s0 + ((bblo - aclo) + s0lo);
    if (disc > -(eps * fmax(bb, fabs(ac)))) {
      double s = sqrt(fmax(disc, 0.0));
      double q = -(b + copysign(s, b)) / 2.0;
      double tau1 = b >= 0 ? q / a : c / q;
      if (tau1 > 0) tau = fmin(tau, tau1);
    }
  } else if (b < 0.0) {
    tau = fmin(tau, -c / b);
  }
  if (dx[0] < 0) tau = fmin(tau, -x[0] / dx[0]);
  return tau;
}

// per-cone max step (cone.jl:455-476, × step_frac except free cones) into out[c]
__global__ void k_cone_maxstep(const InstDev* insts, const ConeDesc* cones, int ncone, int kdp, int kdd,
                               double step_frac, double* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const double* p = nvec(I, NV_P) + C.col0;
  const double* d = nvec(I, NV_D) + C.col0;
  if (C.kind == CONE_PSD) return;  // k_psd_maxstep
  const double* dp = nvec(I, kdp) + C.col0;
  const double* dd = nvec(I, kdd) + C.col0;
  double tp = 1.0, td = 1.0;
  if (C.kind == CONE_POS) {
    tp = posmaxstep(p, dp, C.n);
    td = posmaxstep(d, dd, C.n);
  } else if (C.kind == CONE_SOC) {
    tp = socmaxstep(p, dp, C.n, C.cache[1], C.cache[3]);
    td = socmaxstep(d, dd, C.n, C.cache[2], C.cache[4]);
  } else if (C.kind == CONE_EXP || C.kind == CONE_POW) {
    tdc::maxsteps(C.kind, C.param, p, dp, d, dd, C.cache, tp, td);
  }
  if (C.kind != CONE_FREE) {
    tp *= step_frac;
    td *= step_frac;
  }
  out[c] = fmin(tp, td);
}

// shadow-solve seed of the listed exponential / power cones (initcache!: exp.jl:27-30,
// pow.jl:35-38)
__global__ void k_td_initcache(const InstDev* insts, const ConeDesc* cones, const int* list, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const ConeDesc C = cones[list[t]];
  if (insts[C.inst].sc[S_ACTIVE] == 0.0) return;
  C.cache[tdc::C_SEED] = tdc::seed0(C.kind);
}

void launch_td_initcache(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n) {
  if (n) k_td_initcache<<<(n + 255) / 256, 256, 0, st>>>(I, C, list, n);
}

// zero the Δd entries of free cones (solver/ipm.jl:551-559)
__global__ void k_zerofree(const InstDev* insts, const ConeDesc* cones, int ncone, int kx) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncone) return;
  ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || C.kind != CONE_FREE) return;
  double* x = nvec(I, kx) + C.col0;
  for (int i = 0; i < C.n; ++i) x[i] = 0.0;
}

// ------------------------------------------------------------------------------------------
// assembly: F ← 0; F[pos] = L + δ Q on the union pattern; F[diag blocks] += δ D
// ------------------------------------------------------------------------------------------

// x[0, n) = v with 16-byte stores, one per thread (grid-strided if the grid is smaller; x 16-byte
// aligned)
__global__ void k_fill(double* x, size_t n, double v) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const size_t np = n / 2, stride = (size_t)gridDim.x * blockDim.x;
  d2* x2 = reinterpret_cast<d2*>(x);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) x2[i] = d2{v, v};
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) x[n - 1] = v;
}

__global__ void k_assemble_union(const InstDev* insts, int inst, const int64_t* pos, const double* lval,
                                 const double* qval, int64_t nent, int withq) {
  const InstDev& I = insts[inst];
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nent) return;
  double delta = I.sc[S_DELTA];
  I.F[pos[k]] = withq ? lval[k] + delta * qval[k] : lval[k];
}

__global__ void k_scatter_add(const InstDev* insts, int inst, const int64_t* pos, const double* val, int64_t nent) {
  const InstDev& I = insts[inst];
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nent) return;
  I.F[pos[k]] += I.sc[S_DELTA] * val[k];
}

// cone blocks: one thread per (cone, column); lower triangle of the block lands in the
// front holding the cone's columns
// grid (cone, column group): a dense block's columns j ≡ blockIdx.y (mod gridDim.y), the threads
// along each column (coalesced); a nonnegative cone's diagonal spread over the whole group
__global__ void k_assemble_cones(const InstDev* insts, const ConeDesc* cones, const int* cone_front,
                                 const FrontDesc* fronts, int ncone) {
  const int c = blockIdx.x;
  if (c >= ncone) return;
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  // free cones add nothing; large PSD cones (no stored block) are assembled by k_bp_assemble
  if (I.sc[S_ACTIVE] == 0.0 || C.kind == CONE_FREE || C.D == nullptr) return;
  const FrontDesc& Fd = fronts[cone_front[c]];
  const double delta = I.sc[S_DELTA];
  const int off = C.col0 - Fd.res0;
  if (C.kind == CONE_POS) {
    for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < C.n; j += gridDim.y * blockDim.x)
      Fd.F[(size_t)(off + j) * Fd.ld + off + j] += delta * C.D[j];
    return;
  }
  for (int j = blockIdx.y; j < C.n; j += gridDim.y) {
    double* col = Fd.F + (size_t)(off + j) * Fd.ld + off;
    const double* Dj = C.D + (size_t)j * C.n;
    for (int i = j + threadIdx.x; i < C.n; i += blockDim.x) col[i] += delta * Dj[i];
  }
}

// F[diag] += ρ over res columns of every front of the instance; also ‖tril F‖² partials
__global__ void k_add_diag(const FrontDesc* fronts, int nfront, const InstDev* insts) {
  int f = blockIdx.x;
  if (f >= nfront) return;
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  double rho = I.sc[S_RHO];
  for (int j = threadIdx.x; j < Fd.r; j += blockDim.x) Fd.F[(size_t)j * Fd.ld + j] += rho;
}

// Σ F² over the res columns' lower part of every front (one block per front) → partial[f]
__global__ void k_front_norm2(const FrontDesc* fronts, int nfront, double* partial) {
  __shared__ double red[TILE / 64];
  int f = blockIdx.x;
  const FrontDesc& Fd = fronts[f];
  double v = 0.0;
  size_t tot = (size_t)Fd.ld * Fd.r;
  for (size_t k = threadIdx.x; k < tot; k += blockDim.x) {
    size_t j = k / Fd.ld, i = k % Fd.ld;
    if (i >= j) {
      double x = Fd.F[k];
      v += x * x;
    }
  }
  double s = block_sum(v, red);
  if (threadIdx.x == 0) partial[f] = s;
}

// ------------------------------------------------------------------------------------------
// extend-add: parent F += child's trailing update block, at relidx positions (atomics:
// siblings of one parent run concurrently)
// ------------------------------------------------------------------------------------------

__global__ void k_extend_add(const FrontDesc* fronts, const int* items, const int* item_tile0, int nitem,
                             const InstDev* insts) {
  // binary search the item owning this block
  int b = blockIdx.x;
  int lo = 0, hi = nitem - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (item_tile0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  int child = items[lo];
  int tile = b - item_tile0[lo];
  const FrontDesc& C = fronts[child];
  if (insts[C.inst].sc[S_ACTIVE] == 0.0) return;
  const FrontDesc& P = fronts[C.parent];
  int s = C.s;
  int T = (s + NB - 1) / NB;
  // lower tile pair (ti ≥ tj) from linear index
  int ti = (int)((sqrt(8.0 * tile + 1.0) - 1.0) / 2.0);
  while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
  while (ti * (ti + 1) / 2 > tile) --ti;
  int tj = tile - ti * (ti + 1) / 2;
  if (ti >= T) return;
  const double* U = C.F + (size_t)C.r * C.ld + C.r;
  int lr = threadIdx.x & 63;
  int i = ti * NB + lr;
  if (i >= s) return;
  int pi = C.relidx[i];
  for (int jj = threadIdx.x >> 6; jj < NB; jj += blockDim.x >> 6) {
    int j = tj * NB + jj;
    if (j > i || j >= s) continue;
    int pj = C.relidx[j];
    // one child per parent in a launch: the entry is this thread's alone
    P.F[pi + (size_t)pj * P.ld] += U[i + (size_t)j * C.ld];
  }
}

// ------------------------------------------------------------------------------------------
// blocked partial Cholesky of frontal matrices
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ int find_item(const StepDesc* steps, int nstep, int b) {
  int lo = 0, hi = nstep - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (steps[mid].tile0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Factor the w×w diagonal block held in LDS (column-major, stride NB+1). Returns false on a
// nonpositive or non-finite pivot (the block is then garbage).
__device__ bool lds_potrf(double* A, int w) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int ld = NB + 1;
  for (int j = 0; j < w; ++j) {
    double piv = A[j + j * ld];
    if (!(piv > 0.0) || !isfinite(piv)) {
      if (threadIdx.x == 0) bad = 1;
      piv = 1.0;
    }
    double l = sqrt(piv);
    __syncthreads();
    int i = j + 1 + (int)threadIdx.x;
    if (threadIdx.x == 0) A[j + j * ld] = l;
    if (i < w) A[i + j * ld] /= l;
    __syncthreads();
    // rank-1 update of the trailing lower triangle
    int rem = w - j - 1;
    for (int t = threadIdx.x; t < rem * rem; t += blockDim.x) {
      int ii = j + 1 + t % rem, kk = j + 1 + t / rem;
      if (ii >= kk) A[ii + kk * ld] -= A[ii + j * ld] * A[kk + j * ld];
    }
    __syncthreads();
  }
  return bad == 0;
}

// One 64-column panel step of every small front of a level. Tile 0 of an item factors the
// diagonal block in LDS, stores L_jj and publishes the factorization epoch in the item's flag;
// the row tiles below wait for that flag and read L_jj, which they then apply to their rows:
// X = A L_jj⁻ᵀ. L_jj moves with the device-coherent policy (sc1 buffer stores and loads), as the
// tiles of the persistent kernels do: after an acquire, ordinary loads of L_jj returned, now and
// then, lines a CU's cache held from before tile 0's store (results differed from run to run
// when other work shared the GPU). The diagonal workgroup precedes its row workgroups in the
// grid, so it runs before (or beside) them and the wait ends; it is bounded all the same, and
// raises the abort flag at its limit.
__global__ void __launch_bounds__(256) k_panel(const FrontDesc* fronts, const StepDesc* steps, int nstep,
                                               const InstDev* insts, int* flags, int epoch, int* abortf) {
  __shared__ double Ld[NB * (NB + 1)];
  __shared__ double X[NB * (NB + 1)];
  __shared__ int okw;
  const int ld = NB + 1;
  int it = find_item(steps, nstep, blockIdx.x);
  StepDesc S = steps[it];
  int tile = blockIdx.x - S.tile0;  // 0 = diag, 1.. = row tiles below kb+w
  const FrontDesc& Fd = fronts[S.front];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  if (!abortf) abortf = reinterpret_cast<int*>(I.sc + S_ABORT);
  double* F = Fd.F;
  int LD = Fd.ld, kb = S.kb, w = S.w;
  if (tile == 0) {
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
      int i = t % NB, j = t / NB;
      Ld[i + j * ld] = (i < w && j < w && i >= j) ? F[(size_t)(kb + i) + (size_t)(kb + j) * LD] : 0.0;
    }
    __syncthreads();
    bool ok = lds_potrf(Ld, w);
    if (!ok && threadIdx.x == 0) atomicAdd(&I.sc[S_FAIL], 1.0);
    store_tile<Coh::Dev>(F, LD, kb, kb, w, w, Ld, true);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&flags[it], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (threadIdx.x == 0) {
    long spins = 0;
    int good = 1;
    while (__hip_atomic_load(&flags[it], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
      if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > (1L << 24)) {
        __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    okw = good;
  }
  __syncthreads();
  if (!okw) return;
  load_tile<Coh::Dev>(Ld, F, LD, kb, kb, w, w);  // its upper part is not read
  __syncthreads();
  int r0 = kb + w + (tile - 1) * NB;
  int nr = min(NB, LD - r0);
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int i = t % NB, j = t / NB;
    X[i + j * ld] = (i < nr && j < w) ? F[(size_t)(r0 + i) + (size_t)(kb + j) * LD] : 0.0;
  }
  __syncthreads();
  // column sweep: X[:, j] /= L[j,j]; X[:, k] -= X[:, j] L[k, j] for k > j
  for (int j = 0; j < w; ++j) {
    if (threadIdx.x < NB) X[threadIdx.x + j * ld] /= Ld[j + j * ld];
    __syncthreads();
    int rem = w - j - 1;
    for (int t = threadIdx.x; t < rem * NB; t += blockDim.x) {
      int i = t % NB, k = j + 1 + t / NB;
      X[i + k * ld] -= X[i + j * ld] * Ld[k + j * ld];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int i = t % NB, j = t / NB;
    if (i < nr && j < w) F[(size_t)(r0 + i) + (size_t)(kb + j) * LD] = X[i + j * ld];
  }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// Trailing update C[ti, tj] -= X[ti] X[tj]ᵀ over lower tile pairs of [kb+w, ld)², FP64 MFMA
// (v_mfma_f64_16x16x4f64). Four waves, each a 32×32 quadrant = 2×2 MFMA tiles.
__global__ void __launch_bounds__(256) k_update(const FrontDesc* fronts, const StepDesc* steps, int nstep,
                                                const InstDev* insts) {
  __shared__ double As[NB * (NB + 1)];
  __shared__ double Bs[NB * (NB + 1)];
  const int ld = NB + 1;
  int it = find_item(steps, nstep, blockIdx.x);
  StepDesc S = steps[it];
  int tile = blockIdx.x - S.tile0;
  const FrontDesc& Fd = fronts[S.front];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  int ti = (int)((sqrt(8.0 * tile + 1.0) - 1.0) / 2.0);
  while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
  while (ti * (ti + 1) / 2 > tile) --ti;
  int tj = tile - ti * (ti + 1) / 2;
  double* F = Fd.F;
  int LD = Fd.ld, kb = S.kb, w = S.w;
  int base = kb + w;
  int ri = base + ti * NB, rj = base + tj * NB;
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    int i = t % NB, k = t / NB;
    bool kin = k < w;
    As[i + k * ld] = (kin && ri + i < LD) ? F[(size_t)(ri + i) + (size_t)(kb + k) * LD] : 0.0;
    Bs[i + k * ld] = (kin && rj + i < LD) ? F[(size_t)(rj + i) + (size_t)(kb + k) * LD] : 0.0;
  }
  __syncthreads();
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int qi = (wave >> 1) * 32, qj = (wave & 1) * 32;
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  int li = lane & 15, lk = lane >> 4;
  int kmax = (w + 3) & ~3;
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    double av[2], bv[2];
    for (int a = 0; a < 2; ++a) av[a] = As[(qi + a * 16 + li) + (k0 + lk) * ld];
    for (int b = 0; b < 2; ++b) bv[b] = Bs[(qj + b * 16 + li) + (k0 + lk) * ld];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
  // D layout of v_mfma_f64_16x16x4f64: lane holds rows (lane/16) + 4t, column lane%16
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int t = 0; t < 4; ++t) {
        int i = ri + qi + a * 16 + (lane >> 4) + 4 * t;
        int j = rj + qj + b * 16 + (lane & 15);
        if (i < LD && j < LD && i >= j) F[(size_t)i + (size_t)j * LD] -= acc[a][b][t];
      }
}

// ------------------------------------------------------------------------------------------
// triangular solves, one block per front per level
// ------------------------------------------------------------------------------------------

// forward: v[res] ← L11⁻¹ (v[res] − child contributions); uvec ← child pass-through + L21 x
__global__ void __launch_bounds__(256) k_fsolve(const FrontDesc* fronts, const int* items, const int* childlist,
                                                const InstDev* insts, int kv) {
  __shared__ double xs[NB];
  int f = items[blockIdx.x];
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* v = nvec(I, kv) + Fd.res0;
  double* u = Fd.uvec;
  int r = Fd.r, s = Fd.s, LD = Fd.ld;
  const double* F = Fd.F;
  for (int k = threadIdx.x; k < s; k += blockDim.x) u[k] = 0.0;
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
  for (int jb = 0; jb < r; jb += NB) {
    int w = min(NB, r - jb);
    // diagonal block solve by wave 0: lane i holds row i of the block in registers
    if (threadIdx.x < 64) {
      int lane = threadIdx.x;
      double Lr[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        Lr[j] = (lane < w && j <= lane) ? F[(size_t)(jb + lane) + (size_t)(jb + j) * LD] : 0.0;
      double xi = lane < w ? v[jb + lane] : 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < w) {
          double xj = __shfl(xi / Lr[j], j, 64);
          if (lane == j) xi = xj;
          if (lane > j) xi -= Lr[j] * xj;
        }
      }
      if (lane < w) {
        v[jb + lane] = xi;
        xs[lane] = xi;
      }
    }
    __syncthreads();
    // update the rest of res and the sep part
    for (int i = jb + w + (int)threadIdx.x; i < r + s; i += blockDim.x) {
      double acc = 0.0;
      for (int j = 0; j < w; ++j) acc += F[(size_t)i + (size_t)(jb + j) * LD] * xs[j];
      if (i < r) v[i] -= acc;
      else u[i - r] += acc;
    }
    __syncthreads();
  }
}

// backward: v[res] ← L11⁻ᵀ (v[res] − L21ᵀ v[sep])
__global__ void __launch_bounds__(256) k_bsolve(const FrontDesc* fronts, const int* items, const InstDev* insts,
                                                int kv) {
  __shared__ double xs[NB];
  __shared__ double red[256];
  int f = items[blockIdx.x];
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* vall = nvec(I, kv);
  double* v = vall + Fd.res0;
  int r = Fd.r, s = Fd.s, LD = Fd.ld;
  const double* F = Fd.F;
  // v[res] -= L21ᵀ v[sep]: one wave per column
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (s > 0) {
    for (int j = wave; j < r; j += blockDim.x >> 6) {
      double acc = 0.0;
      for (int k = lane; k < s; k += 64) acc += F[(size_t)(r + k) + (size_t)j * LD] * vall[Fd.sepcols[k]];
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) v[j] -= acc;
    }
  }
  __syncthreads();
  int nblk = (r + NB - 1) / NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    int jb = bi * NB;
    int w = min(NB, r - jb);
    if (threadIdx.x < 64) {
      // solve L_bbᵀ x = v_b, bottom-up: lane i holds column i of the block in registers
      double Lc[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        Lc[j] = (lane < w && j < w && j >= lane) ? F[(size_t)(jb + j) + (size_t)(jb + lane) * LD] : 0.0;
      double xi = lane < w ? v[jb + lane] : 0.0;
#pragma unroll
      for (int j = NB - 1; j >= 0; --j) {
        if (j < w) {
          // lane j divides by its own diagonal Lc[j] = L[j,j]
          double xj = __shfl(xi / Lc[j], j, 64);
          if (lane == j) xi = xj;
          if (lane < j) xi -= Lc[j] * xj;
        }
      }
      if (lane < w) {
        v[jb + lane] = xi;
        xs[lane] = xi;
      }
    }
    __syncthreads();
    // v[0:jb] -= L[jb:jb+w, 0:jb]ᵀ x_b: one wave per column
    for (int j = wave; j < jb; j += blockDim.x >> 6) {
      double acc = lane < w ? F[(size_t)(jb + lane) + (size_t)j * LD] * xs[lane] : 0.0;
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) v[j] -= acc;
    }
    __syncthreads();
  }
  (void)red;
}

// ------------------------------------------------------------------------------------------
// scalar kernels (one thread per instance)
// ------------------------------------------------------------------------------------------

// CG coefficient updates (cg.jl:108-160). op 0: α = ν²/ε² (ε² in S_CG_EP2), sets S_CG_ALPHA;
// op 1: β = ν²_new/ν²_old with new ν² in S_NORM0, sets S_CG_BETA, S_CG_NU2, S_CG_NP2
__global__ void k_cg_scalars(const InstDev* insts, int ninst, int op) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst) return;
  double* sc = insts[k].sc;
  if (sc[S_ACTIVE] == 0.0) return;
  if (op == 0) {
    sc[S_CG_ALPHA] = sc[S_CG_NU2] / sc[S_CG_EP2];
  } else {
    double pnu2 = sc[S_CG_NU2];
    double nu2 = sc[S_NORM0];
    double beta = nu2 / pnu2;
    sc[S_CG_BETA] = beta;
    sc[S_CG_NU2] = nu2;
    sc[S_CG_NP2] = nu2 + beta * beta * sc[S_CG_NP2];
  }
}

// layout probe: D = A·B for one v_mfma_f64_16x16x4f64 with A[i][k] = i + 100 k,
// B[k][j] = j + 1000 k; every lane writes its 4 outputs to out[lane*4 + t]
__global__ void k_probe_mfma(double* out) {
  int lane = threadIdx.x & 63;
  double a = (double)((lane & 15) + 100 * (lane >> 4));
  double b = (double)((lane & 15) + 1000 * (lane >> 4));
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  for (int t = 0; t < 4; ++t) out[lane * 4 + t] = acc[t];
}

// the same operands through v_mfma_f32_16x16x4f32 (every product and sum is exact in FP32)
__global__ void k_probe_mfma_f32(double* out) {
  int lane = threadIdx.x & 63;
  float a = (float)((lane & 15) + 100 * (lane >> 4));
  float b = (float)((lane & 15) + 1000 * (lane >> 4));
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  for (int t = 0; t < 4; ++t) out[lane * 4 + t] = (double)acc[t];
}

// ------------------------------------------------------------------------------------------
// host-side launch wrappers
// ------------------------------------------------------------------------------------------

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

void launch_axpby(hipStream_t st, const InstDev* I, const Tile* T, int nt, int space, int kx, int ky, int kz,
                  Coef a, Coef b) {
  if (nt) k_axpby<<<nt, TILE, 0, st>>>(I, T, space, kx, ky, kz, a, b);
}
void launch_axpbypcz(hipStream_t st, const InstDev* I, const Tile* T, int nt, int space, int kx, int ky, int kw,
                     int kz, Coef a, Coef b, Coef c) {
  if (nt) k_axpbypcz<<<nt, TILE, 0, st>>>(I, T, space, kx, ky, kw, kz, a, b, c);
}
void launch_dot(hipStream_t st, const InstDev* I, const Tile* T, int nt, const int* tile_ofs, int ninst,
                int space, int kx, int ky, int mode, int kw0, int kw1, double* partial, int* counter, int slot,
                int op) {
  if (nt && ninst == 1) {
    k_dot<<<nt, TILE, 0, st>>>(I, T, tile_ofs, space, kx, ky, mode, kw0, kw1, partial, counter, slot, op);
    return;
  }
  if (nt) k_dot_partial<<<nt, TILE, 0, st>>>(I, T, space, kx, ky, mode, kw0, kw1, partial);
  k_reduce_slot<<<ninst, TILE, 0, st>>>(I, tile_ofs, partial, slot, op);
}
void launch_dots(hipStream_t st, const InstDev* I, int inst, const DotBatch& B, int ninst) {
  if (B.n > 0 && ninst > 0) k_dots<<<dim3(B.n, ninst), DOTS_THREADS, 0, st>>>(I, inst, B);
}
void launch_lin(hipStream_t st, const InstDev* I, const Tile* ntiles, int nnt, const Tile* mtiles, int nmt,
                const LinSet& N, const LinSet& M) {
  const int nn = N.n ? nnt : 0, nm = M.n ? nmt : 0;
  if (nn + nm) k_lin<<<nn + nm, TILE, 0, st>>>(I, ntiles, nn, mtiles, N, M);
}
void launch_spmv(hipStream_t st, const InstDev* I, const Tile* T, int nt, int which, int kx, int ky, Coef a,
                 Coef b) {
  if (nt) k_spmv<<<nt, TILE, 0, st>>>(I, T, which, kx, ky, a, b);
}
void launch_spmv_nm(hipStream_t st, const InstDev* I, int ntm, const Tile* ntiles, int nn, const Tile* mtiles, int nm,
                    const SpmvN& N, const SpmvM& M) {
  if (nn + nm) k_spmv_nm<<<nn + nm, TILE, 0, st>>>(I, ntm, ntiles, nn, mtiles, N, M);
}
void launch_cone_matvec(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, int kx, int ky, Coef a) {
  if (nc) k_cone_matvec<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kx, ky, a);
}
void launch_cone_scale(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, const int* cone_ofs,
                       int ninst, double* out, const int* psd, int npsd, const BigPsd& big, const int* soc,
                       int nsoc) {
  if (nc) k_cone_scale<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, out);
  if (nsoc) k_soc_scale<<<nsoc, 64, 0, st>>>(I, C, soc, out);
  launch_psd_scale(st, I, C, psd, npsd, out);
  launch_bpsd_scale(st, I, C, big, out);
  k_cone_reduce<<<ninst, TILE, 0, st>>>(I, cone_ofs, out, 0);
}
void launch_cone_hnorm2(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, const int* cone_ofs,
                        int ninst, double* out, int slot, const int* dense, int ndense, int dense_nmax,
                        double* dpart, const BigPsd& big, double* bpart) {
  (void)dense_nmax;
  if (nc) k_cone_hnorm2<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, out);
  launch_cone_hnorm2_dense(st, I, C, dense, ndense, out, dpart);
  launch_bpsd_hnorm2(st, I, C, big, out, bpart);
  k_cone_sum<<<ninst, TILE, 0, st>>>(I, cone_ofs, out, slot);
}
void launch_cone_corr0(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, int kr) {
  if (nc) k_cone_corr0<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kr);
}
void launch_cone_corr(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, int kr, int kdp, int kdd) {
  if (nc) k_cone_corr<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kr, kdp, kdd);
}
void launch_cone_maxstep(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, const int* cone_ofs,
                         int ninst, int kdp, int kdd, double step_frac, double* out, const int* psd, int npsd,
                         const BigPsd& big) {
  if (nc) k_cone_maxstep<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kdp, kdd, step_frac, out);
  launch_psd_maxstep(st, I, C, psd, npsd, kdp, kdd, step_frac, out);
  launch_bpsd_maxstep(st, I, C, big, kdp, kdd, step_frac, out);
  k_cone_reduce<<<ninst, TILE, 0, st>>>(I, cone_ofs, out, 1);
}
void launch_cone_dshadow(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, int kx, int ky) {
  if (nc) k_cone_dshadow<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kx, ky);
}
void launch_zerofree(hipStream_t st, const InstDev* I, const ConeDesc* C, int nc, int kx) {
  if (nc) k_zerofree<<<cdiv(nc, 64), 64, 0, st>>>(I, C, nc, kx);
}
void launch_fill(hipStream_t st, double* x, size_t n, double v) {
  if (n) k_fill<<<std::max(cdiv((long)(n / 2), TILE), 1), TILE, 0, st>>>(x, n, v);
}

// The pieces of a staged setup (copies out of one uploaded image, and zero fills) in one launch:
// workgroup b finds its piece by bisection over blk0 and moves its STAGE_BLOCK bytes of it,
// 16 bytes per thread and step where both ends are 16-byte aligned, else 4 (and single bytes last).
__global__ void k_stage_scatter(const char* image, const StageSeg* segs, int nseg) {
  const long long b = blockIdx.x;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (segs[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const StageSeg S = segs[lo];
  const long long a = (b - S.blk0) * STAGE_BLOCK, e = min(S.bytes, a + STAGE_BLOCK);
  if (a >= e) return;
  char* d = S.dst + a;
  const char* src = S.src >= 0 ? image + S.src + a : nullptr;
  const long long n = e - a;
  const bool v16 = ((reinterpret_cast<uintptr_t>(d) | (src ? reinterpret_cast<uintptr_t>(src) : 0)) & 15) == 0;
  if (v16) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const long long n16 = n / 16;
    for (long long i = threadIdx.x; i < n16; i += blockDim.x)
      reinterpret_cast<u4*>(d)[i] = src ? reinterpret_cast<const u4*>(src)[i] : u4{0, 0, 0, 0};
    for (long long i = n16 * 4 + threadIdx.x; i < n / 4; i += blockDim.x)
      reinterpret_cast<unsigned*>(d)[i] = src ? reinterpret_cast<const unsigned*>(src)[i] : 0u;
  } else {
    for (long long i = threadIdx.x; i < n / 4; i += blockDim.x)
      reinterpret_cast<unsigned*>(d)[i] = src ? reinterpret_cast<const unsigned*>(src)[i] : 0u;
  }
  if (threadIdx.x == 0)
    for (long long i = n & ~3ll; i < n; ++i) d[i] = src ? src[i] : 0;  // a piece of odd size
}
void launch_stage_scatter(hipStream_t st, const char* image, const StageSeg* segs, int nseg, long long nblk) {
  if (nseg > 0 && nblk > 0) k_stage_scatter<<<(unsigned)nblk, 256, 0, st>>>(image, segs, nseg);
}
void launch_assemble_union(hipStream_t st, const InstDev* I, int inst, const int64_t* pos, const double* lval,
                           const double* qval, int64_t nent, int withq) {
  if (nent) k_assemble_union<<<cdiv(nent, TILE), TILE, 0, st>>>(I, inst, pos, lval, qval, nent, withq);
}
void launch_scatter_add(hipStream_t st, const InstDev* I, int inst, const int64_t* pos, const double* val,
                        int64_t nent) {
  if (nent) k_scatter_add<<<cdiv(nent, TILE), TILE, 0, st>>>(I, inst, pos, val, nent);
}
void launch_assemble_cones(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* cone_front,
                           const FrontDesc* F, int nc, int nmax) {
  const int groups = std::max(1, std::min(32, nmax / 64));
  if (nc) k_assemble_cones<<<dim3(nc, groups), 256, 0, st>>>(I, C, cone_front, F, nc);
}
void launch_add_diag(hipStream_t st, const FrontDesc* F, int nf, const InstDev* I) {
  if (nf) k_add_diag<<<nf, TILE, 0, st>>>(F, nf, I);
}
void launch_front_norm2(hipStream_t st, const FrontDesc* F, int nf, double* partial) {
  if (nf) k_front_norm2<<<nf, TILE, 0, st>>>(F, nf, partial);
}
void launch_extend_add(hipStream_t st, const FrontDesc* F, const int* items, const int* item_tile0, int nitem,
                       int ntile, const InstDev* I) {
  if (ntile) k_extend_add<<<ntile, TILE, 0, st>>>(F, items, item_tile0, nitem, I);
}
void launch_panel(hipStream_t st, const FrontDesc* F, const StepDesc* S, int nstep, int nblk, const InstDev* I,
                  int* flags, int epoch, int* abortf) {
  if (nblk) k_panel<<<nblk, 256, 0, st>>>(F, S, nstep, I, flags, epoch, abortf);
}
void launch_update(hipStream_t st, const FrontDesc* F, const StepDesc* S, int nstep, int nblk, const InstDev* I) {
  if (nblk) k_update<<<nblk, 256, 0, st>>>(F, S, nstep, I);
}
void launch_fsolve(hipStream_t st, const FrontDesc* F, const int* items, int nitem, const int* childlist,
                   const InstDev* I, int kv) {
  if (nitem) k_fsolve<<<nitem, 256, 0, st>>>(F, items, childlist, I, kv);
}
void launch_bsolve(hipStream_t st, const FrontDesc* F, const int* items, int nitem, const InstDev* I, int kv) {
  if (nitem) k_bsolve<<<nitem, 256, 0, st>>>(F, items, I, kv);
}
void launch_cg_scalars(hipStream_t st, const InstDev* I, int ninst, int op) {
  k_cg_scalars<<<cdiv(ninst, 64), 64, 0, st>>>(I, ninst, op);
}
void launch_probe_mfma(hipStream_t st, double* out, bool fp32) {
  if (fp32) k_probe_mfma_f32<<<1, 64, 0, st>>>(out);
  else k_probe_mfma<<<1, 64, 0, st>>>(out);
}

__global__ void k_set_scalar(double* sc, int slot, double v) { sc[slot] = v; }
__global__ void k_set_scalar2(double* sc, int s1, double v1, int s2, double v2) {
  sc[s1] = v1;
  sc[s2] = v2;
}
void launch_set_scalar2(hipStream_t st, double* sc, int s1, double v1, int s2, double v2) {
  k_set_scalar2<<<1, 1, 0, st>>>(sc, s1, v1, s2, v2);
}

// solver/ipm.jl:181-200: σμ from the affine centrality μa, on the device so that the corrector's
// right-hand side needs no read-back
__device__ __forceinline__ void smu_value(double* sc, double nu, double mu, double mut) {
  double smu = sc[S_TMP2];
  const double mua = smu / nu;
  if (mut > 0) {
    const double dmu = mu - mut;
    if (dmu > 0) {
      const double dmua = fmin(fmax(mua - mut, 0.0), dmu);
      smu = mut + fmin(fmax(dmua * (dmua / dmu) * (dmua / dmu), 0.0), dmu);
    } else {
      smu = mut;
    }
  } else {
    smu = fmin(fmax(mua * (mua / mu) * (mua / mu), 0.0), mu);
  }
  sc[S_SMU] = smu;
}
__global__ void k_smu(double* sc, double nu, double mu, double mut) { smu_value(sc, nu, mu, mut); }
void launch_smu(hipStream_t st, double* sc, double nu, double mu, double mut) {
  k_smu<<<1, 1, 0, st>>>(sc, nu, mu, mut);
}
void launch_set_scalar(hipStream_t st, double* sc, int slot, double v) { k_set_scalar<<<1, 1, 0, st>>>(sc, slot, v); }

// ------------------------------------------------------------------------------------------
// gang (lockstep batch) kernels: one thread per member slot of the combined instance table
// ------------------------------------------------------------------------------------------

// sc[s1] = a[k], sc[s2] = b[k] (s2 < 0: only s1) for the members taking part (S_ACTIVE)
__global__ void k_gang_set2(const InstDev* insts, int k, int s1, int s2, GangVals v) {
  const int i = threadIdx.x;
  if (i >= k) return;
  double* sc = insts[i].sc;
  if (sc[S_ACTIVE] == 0.0) return;
  sc[s1] = v.a[i];
  if (s2 >= 0) sc[s2] = v.b[i];
}
void launch_gang_set2(hipStream_t st, const InstDev* I, int k, int s1, int s2, const GangVals& v) {
  k_gang_set2<<<1, GANG_SLOTS, 0, st>>>(I, k, s1, s2, v);
}

// σμ of each member taking part (k_smu with ν = a, μ = b, μt = c)
__global__ void k_gang_smu(const InstDev* insts, int k, GangVals v) {
  const int i = threadIdx.x;
  if (i >= k) return;
  double* sc = insts[i].sc;
  if (sc[S_ACTIVE] == 0.0) return;
  smu_value(sc, v.a[i], v.b[i], v.c[i]);
}
void launch_gang_smu(hipStream_t st, const InstDev* I, int k, const GangVals& v) {
  k_gang_smu<<<1, GANG_SLOTS, 0, st>>>(I, k, v);
}

// the fronts of every member taking part set to zero, then BᵀB + δQ scattered into them (the
// members' own k_fill and k_assemble_union in one launch each); block b belongs to the member i
// with A[i].fb0 ≤ b < A[i+1].fb0 (ub0 likewise), A holding k+1 entries
__device__ __forceinline__ int gang_member(const GangAsm* A, int k, long long b, bool fill) {
  int i = 0;
  while (i + 1 < k && (fill ? A[i + 1].fb0 : A[i + 1].ub0) <= b) ++i;
  return i;
}
__global__ void k_gang_fill(const InstDev* insts, const GangAsm* A, int k) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const long long b = blockIdx.x;
  const int i = gang_member(A, k, b, true);
  if (insts[i].sc[S_ACTIVE] == 0.0) return;
  const GangAsm& M = A[i];
  const long long t = (b - M.fb0) * blockDim.x + threadIdx.x;
  if (t < M.nF / 2) reinterpret_cast<d2*>(M.F)[t] = d2{0.0, 0.0};
  if ((M.nF & 1) && t == 0) M.F[M.nF - 1] = 0.0;
}
__global__ void k_gang_union(const InstDev* insts, const GangAsm* A, int k) {
  const long long b = blockIdx.x;
  const int i = gang_member(A, k, b, false);
  const InstDev& I = insts[i];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const GangAsm& M = A[i];
  const long long t = (b - M.ub0) * blockDim.x + threadIdx.x;
  if (t >= M.nent) return;
  M.F[M.pos[t]] = M.l[t] + I.sc[S_DELTA] * M.q[t];
}
void launch_gang_assemble(hipStream_t st, const InstDev* I, const GangAsm* A, int k, long long fblocks,
                          long long ublocks) {
  if (fblocks > 0) k_gang_fill<<<(unsigned)fblocks, TILE, 0, st>>>(I, A, k);
  if (ublocks > 0) k_gang_union<<<(unsigned)ublocks, TILE, 0, st>>>(I, A, k);
}

// S_ACTIVE = bit i of mask for the occupied slots
__global__ void k_gang_active(const InstDev* insts, int k, unsigned long long mask, unsigned long long occupied) {
  const int i = threadIdx.x;
  if (i >= k || !((occupied >> i) & 1ull)) return;
  insts[i].sc[S_ACTIVE] = ((mask >> i) & 1ull) ? 1.0 : 0.0;
}
void launch_gang_active(hipStream_t st, const InstDev* I, int k, unsigned long long mask,
                        unsigned long long occupied) {
  k_gang_active<<<1, GANG_SLOTS, 0, st>>>(I, k, mask, occupied);
}

// one entry of the instance table (stream-ordered with the launches that read it)
__global__ void k_gang_setinst(InstDev* insts, int slot, InstDev v) {
  if (threadIdx.x == 0) insts[slot] = v;
}
void launch_gang_setinst(hipStream_t st, InstDev* I, int slot, const InstDev& v) {
  k_gang_setinst<<<1, 64, 0, st>>>(I, slot, v);
}

// the scalar blocks of the slots in mask, side by side (slot i at out + i·NSCALAR)
__global__ void k_gang_gather(const InstDev* insts, unsigned long long mask, double* out) {
  const int i = blockIdx.x;
  if (!((mask >> i) & 1ull)) return;
  const double* sc = insts[i].sc;
  for (int t = threadIdx.x; t < NSCALAR; t += blockDim.x) out[(size_t)i * NSCALAR + t] = sc[t];
}
void launch_gang_gather(hipStream_t st, const InstDev* I, int k, unsigned long long mask, double* out) {
  if (k > 0) k_gang_gather<<<k, 64, 0, st>>>(I, mask, out);
}

}  // namespace mb
