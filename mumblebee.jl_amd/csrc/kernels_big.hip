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

#include <cmath>
#include <cstdint>

#include "device.h"
#include "kernels.h"
#include "kernels_big.h"

namespace mb {

#define LDT (NB + 1)           // column stride of a column-major tile in LDS
constexpr int LROW = NB + 2;   // row stride of a row-major tile in LDS (16-byte aligned rows)
constexpr int NT = 256;        // threads per workgroup of every kernel here
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

// task timeline (optional), TRACE_SLOTS stamps per task: 0 claimed, 1 updates done, 2 inputs of
// the last step ready, 3 done; a diagonal task also stamps 4 factor panels done, 5 diagonal-block
// inverses done, 6 inverse complete, 7 inverse published
constexpr int TRACE_SLOTS = 8;
__device__ __forceinline__ void stamp(unsigned long long* trace, int t, int slot) {
  if (trace != nullptr && threadIdx.x == 0)
    trace[TRACE_SLOTS * (size_t)t + slot] = (unsigned long long)wall_clock64();
}

// ------------------------------------------------------------------------------------------
// tile movement (256 threads, 16 elements each)
// ------------------------------------------------------------------------------------------

// Tiles move through raw buffer instructions over a resource spanning exactly the tile's
// extent in F: a lane outside the h×w tile gets an offset past that extent, which the hardware
// range check turns into a zero load or a dropped store. With the device-coherent cache policy
// (sc1, the same policy as the agent-scope atomics above) the loads are ordinary loads to the
// compiler, so a thread's 16 loads are in flight together instead of one after another.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int CPOL_SC1 = 16;           // buffer cache policy bit: coherent at device scope
constexpr unsigned OFF_OUT = 0x80000000u;  // an offset past every tile extent

template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const T* F, int LD, int r0, int c0, int h, int w) {
  const T* base = F + (size_t)r0 + (size_t)c0 * LD;
  const unsigned bytes = (unsigned)(((size_t)(w - 1) * LD + h) * sizeof(T));
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), 0, (int)bytes, 0x00020000);
}

template <bool COH>
__device__ __forceinline__ void buf_load(double& v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, COH ? CPOL_SC1 : 0));
}
template <bool COH>
__device__ __forceinline__ void buf_load(float& v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, COH ? CPOL_SC1 : 0));
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, (int)off, 0, CPOL_SC1);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)off, 0, CPOL_SC1);
}

// A thread's share of a tile in flight: element t = threadIdx.x + u·NT is (t % NB, t / NB).
template <class T>
struct TileRegs {
  T v[NB * NB / NT];
};

// issue the loads of the h×w tile at (r0, c0) of F (leading dimension LD), zero padded.
// COH: coherent loads (the tile was written by another workgroup of the running kernel).
template <bool COH, class T>
__device__ __forceinline__ void fetch_tile(TileRegs<T>& R, const T* F, int LD, int r0, int c0, int h, int w) {
  const __amdgpu_buffer_rsrc_t rs = tile_rsrc(F, LD, r0, c0, h, w);
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    int t = threadIdx.x + u * NT;
    int i = t % NB, j = t / NB;
    unsigned off = i < h && j < w ? (unsigned)((i + (size_t)j * LD) * sizeof(T)) : OFF_OUT;
    buf_load<COH>(R.v[u], rs, off);
  }
}

// fetched tile → LDS column-major (stride LDT), converted to the LDS element type
template <class TD, class TS>
__device__ __forceinline__ void put_tile(TD* S, const TileRegs<TS>& R) {
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    int t = threadIdx.x + u * NT;
    S[(t % NB) + (t / NB) * LDT] = (TD)R.v[u];
  }
}

template <bool COH, class TD, class TS>
__device__ __forceinline__ void load_tile(TD* S, const TS* F, int LD, int r0, int c0, int h, int w) {
  TileRegs<TS> R;
  fetch_tile<COH>(R, F, LD, r0, c0, h, w);
  put_tile(S, R);
}

// LDS column-major tile → h×w tile at (r0, c0) of F with coherent stores (lower part only if
// lower), converted to F's element type
template <class TD, class TS>
__device__ __forceinline__ void store_tile(TD* F, int LD, int r0, int c0, int h, int w, const TS* S, bool lower) {
  const __amdgpu_buffer_rsrc_t rs = tile_rsrc(F, LD, r0, c0, h, w);
  TD v[NB * NB / NT];
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    int t = threadIdx.x + u * NT;
    v[u] = (TD)S[(t % NB) + (t / NB) * LDT];
  }
#pragma unroll
  for (int u = 0; u < NB * NB / NT; ++u) {
    int t = threadIdx.x + u * NT;
    int i = t % NB, j = t / NB;
    unsigned off = i < h && j < w && (!lower || i >= j) ? (unsigned)((i + (size_t)j * LD) * sizeof(TD)) : OFF_OUT;
    buf_store(v[u], rs, off);
  }
}

// 16×16×4 MFMA of the element type. Both take A[i][k] from lane i + 16k and B[k][j] from lane
// j + 16k; they differ in where the 16×16 result lands: element t of lane l is row
// (l>>4) + 4t for FP64 and row 4(l>>4) + t for FP32, column l & 15 for both (the lane-layout
// probes mb_probe_mfma_f64 / _f32 check this on the device).
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int t) { return (lane >> 4) + 4 * t; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int t) { return 4 * (lane >> 4) + t; }
};

// The 64×64 accumulator tile of a workgroup: wave w holds the 32×32 quadrant (qi, qj) =
// (32·(w>>1), 32·(w&1)) as 2×2 blocks of 16×16 MFMA outputs; element t of block (a, b) in
// lane l is row qi + 16a + Mfma<T>::row(l, t), column qj + 16b + (l&15).
template <class T>
struct Acc {
  typename Mfma<T>::v4 v[2][2];
};

template <class T>
__device__ __forceinline__ void acc_from_lds(Acc<T>& A, const T* S) {
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int qi = (wave >> 1) * 32, qj = (wave & 1) * 32;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        A.v[a][b][t] = S[(qi + a * 16 + Mfma<T>::row(lane, t)) + (qj + b * 16 + (lane & 15)) * LDT];
}

template <class T>
__device__ __forceinline__ void acc_to_lds(const Acc<T>& A, T* S) {
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int qi = (wave >> 1) * 32, qj = (wave & 1) * 32;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        S[(qi + a * 16 + Mfma<T>::row(lane, t)) + (qj + b * 16 + (lane & 15)) * LDT] = A.v[a][b][t];
}

template <class T>
__device__ __forceinline__ void acc_zero(Acc<T>& A) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) A.v[a][b] = typename Mfma<T>::v4{0, 0, 0, 0};
}

// A ← A ± X Yᵀ for 64×64 LDS tiles X, Y (column-major, zero padded past column w)
template <bool NEG, class T>
__device__ __forceinline__ void mfma_acc(Acc<T>& A, const T* X, const T* Y, int w) {
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int qi = (wave >> 1) * 32, qj = (wave & 1) * 32;
  int li = lane & 15, lk = lane >> 4;
  int kmax = (w + 3) & ~3;
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    T xv[2], yv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      T x = X[(qi + a * 16 + li) + (k0 + lk) * LDT];
      xv[a] = NEG ? -x : x;
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) yv[b] = Y[(qj + b * 16 + li) + (k0 + lk) * LDT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) A.v[a][b] = Mfma<T>::op(xv[a], yv[b], A.v[a][b]);
  }
}

__device__ __forceinline__ double readlane(double v, int lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// 1/√d to full precision from the hardware estimate y (relative error about 2^-23). FP64: with
// e = 1 − d y², 1/√d = y (1 − e)^(−1/2) = y (1 + e/2 + 3e²/8 + O(e³)), one correction whose
// truncation error (about 2^-66) is below the rounding, four dependent operations instead of the
// six of two Newton steps (the factorization's pivot chain runs through it). FP32: one Newton step.
__device__ __forceinline__ double rsqrt_nr(double d) {
  const double y = __builtin_amdgcn_rsq(d);
  const double e = __builtin_fma(-d * y, y, 1.0);  // 1 − d y²
  const double p = __builtin_fma(0.375, e, 0.5);   // 1/2 + 3e/8
  return __builtin_fma(y * e, p, y);
}
__device__ __forceinline__ float rsqrt_nr(float d) {
  float y = __builtin_amdgcn_rsqf(d);
  float e = __builtin_fmaf(-d * y, y, 1.0f);
  return __builtin_fmaf(0.5f * y, e, y);
}

// ------------------------------------------------------------------------------------------
// Cholesky factor and inverse of a 64×64 diagonal tile. The factor is right-looking over four
// 16-column panels: wave 0 factors a panel with one row per lane (all 64 rows at once), the
// other waves wait, then all four waves subtract the panel's outer product from the trailing
// tiles on MFMA. The inverse follows from the factor: wave b inverts diagonal block b (one column
// per lane), then forms X_{b+s,b} = −X_{b+s,b+s} Σ_k L_{b+s,k} X_{k,b} for s = 1, 2, 3 in its
// own block column, the products on MFMA 16×16 tiles.
// ------------------------------------------------------------------------------------------

// 16×16 MFMA tile: acc ± Σ_{k<K} A[ra+i][ca+k] · op(B)[k][j] from LDS (stride LDT), with
// op(B)[k][j] = B[rb+j][cb+k] if BT, else B[rb+k][cb+j]
template <int K, bool BT, bool NEG, class T>
__device__ __forceinline__ typename Mfma<T>::v4 mfma16(typename Mfma<T>::v4 acc, const T* A, int ra, int ca,
                                                       const T* Bm, int rb, int cb) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 4) {
    const T a = A[(ra + li) + (ca + k0 + lk) * LDT];
    const T b = BT ? Bm[(rb + li) + (cb + k0 + lk) * LDT] : Bm[(rb + k0 + lk) + (cb + li) * LDT];
    acc = Mfma<T>::op(NEG ? -a : a, b, acc);
  }
  return acc;
}

template <class T>
__device__ __forceinline__ typename Mfma<T>::v4 tile16_load(const T* C, int rc, int cc) {
  const int lane = threadIdx.x & 63;
  typename Mfma<T>::v4 v;
#pragma unroll
  for (int t = 0; t < 4; ++t) v[t] = C[(rc + Mfma<T>::row(lane, t)) + (cc + (lane & 15)) * LDT];
  return v;
}

template <class T>
__device__ __forceinline__ void tile16_store(T* C, int rc, int cc, typename Mfma<T>::v4 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 4; ++t) C[(rc + Mfma<T>::row(lane, t)) + (cc + (lane & 15)) * LDT] = v[t];
}

// LDS writes of one wave become visible to the wave's later LDS reads in other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Panel c0..c0+15 of S, by wave 0: lane r holds row r (columns c0..c0+15) in registers. Step j
// takes the pivot from lane c0+j, scales column j by 1/√d, and updates the trailing columns:
// the next two columns' entries come by v_readlane (they feed the next pivots), the others
// through a broadcast read of the scaled column from LDS. Rows ≥ c0 are written back, and
// rv[c0+j] ← 1/L_jj.
// Returns 1 on a non-positive or non-finite pivot: the factor then holds NaNs and is discarded
// by the caller, which retries with a diagonal shift (the check stays off the step's chain).
template <class T>
__device__ __forceinline__ int panel16(T* S, T* rv, T* colj, int c0) {
  const int lane = threadIdx.x & 63;
  T row[16], rr[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) row[k] = S[lane + (c0 + k) * LDT];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const T d = readlane(row[j], c0 + j);
    bad |= !(d > T(0)) || !isfinite(d);
    const T r = rsqrt_nr(d);
    rr[j] = r;
    const T lij = row[j] * r;
    row[j] = lij;
    if (j + 1 < 16) {
      T* cb = colj + (j & 1) * 64;  // two buffers: step j+1 writes while step j's reads drain
      cb[lane] = lij;
      row[j + 1] -= lij * readlane(lij, c0 + j + 1);
      if (j + 2 < 16) row[j + 2] -= lij * readlane(lij, c0 + j + 2);  // slack for the LDS reads
      wave_lds_sync();
#pragma unroll
      for (int k = j + 3; k < 16; ++k) row[k] -= lij * cb[c0 + k];
    }
  }
  T mine = rr[0];  // lane j < 16 stores 1/L_jj
#pragma unroll
  for (int j = 1; j < 16; ++j) mine = lane == j ? rr[j] : mine;
  if (lane < 16) rv[c0 + lane] = mine;
  if (lane >= c0) {
#pragma unroll
    for (int k = 0; k < 16; ++k) S[lane + (c0 + k) * LDT] = row[k];
  }
  return bad ? 1 : 0;
}

// X_{b+SL,b} = −X_{b+SL,b+SL} Σ_{k} L_{b+SL,k} X_{k,b} (block indices, k = b..b+SL−1), by one
// wave; the partial sum goes through X_{b+SL,b}'s place in LDS
template <int SL, class T>
__device__ __forceinline__ void inv_level(const T* S, T* X, int b) {
  typedef typename Mfma<T>::v4 v4;
  const v4 t = mfma16<16 * SL, false, false>(v4{0, 0, 0, 0}, S, 16 * (b + SL), 16 * b, X, 16 * b, 16 * b);
  tile16_store(X, 16 * (b + SL), 16 * b, t);
  wave_lds_sync();
  const v4 a = mfma16<16, false, true>(v4{0, 0, 0, 0}, X, 16 * (b + SL), 16 * (b + SL), X, 16 * (b + SL), 16 * b);
  wave_lds_sync();
  tile16_store(X, 16 * (b + SL), 16 * b, a);
}

// Cholesky factor L (in S: column-major, zero upper part, identity padded past h) and its
// inverse X of the diagonal tile of order h held in S (zero padded past h). W: 192 elements of
// LDS scratch, best below 64 KiB of LDS so that the panel's broadcast reads take immediate
// offsets. Returns false on a non-positive pivot.
template <class T>
__device__ __forceinline__ bool chol_inv64(T* S, T* X, T* W, int h, unsigned long long* trace, int task) {
  __shared__ int okflag;
  T* rv = W;
  T* colj = W + NB;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x >= h && threadIdx.x < NB) S[threadIdx.x * (LDT + 1)] = T(1);
  __syncthreads();
  int bad = 0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int c0 = 16 * p;
    if (w == 0) bad |= panel16(S, rv, colj, c0);
    __syncthreads();
    if (p < 3) {
      // the lower 16×16 tiles (ti ≥ tk) of the trailing blocks p+1..3, spread over the waves
      const int nb = 3 - p, ntile = nb * (nb + 1) / 2;
      for (int t = w; t < ntile; t += NT / 64) {
        int ti = 0, tk = t;
        while (tk > ti) {
          tk -= ti + 1;
          ++ti;
        }
        const int ri = c0 + 16 + 16 * ti, rk = c0 + 16 + 16 * tk;
        typename Mfma<T>::v4 acc = tile16_load(S, ri, rk);
        acc = mfma16<16, true, true>(acc, S, ri, c0, S, rk, c0);
        tile16_store(S, ri, rk, acc);
      }
      __syncthreads();
    }
  }
  stamp(trace, task, 4);
  {
    // wave w inverts diagonal block w: lane c < 16 forms column c by forward substitution
    const int o = 16 * w, c = lane & 15;
    T x[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) x[m] = (m == c) ? T(1) : T(0);
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
      x[cc] *= rv[o + cc];
#pragma unroll
      for (int m = cc + 1; m < 16; ++m) x[m] -= S[(o + m) + (o + cc) * LDT] * x[cc];
    }
    if (lane < 16) {
#pragma unroll
      for (int m = 0; m < 16; ++m) X[(o + m) + (o + c) * LDT] = x[m];
    }
  }
  __syncthreads();
  stamp(trace, task, 5);
  if (w <= 2) inv_level<1>(S, X, w);
  if (w <= 1) inv_level<2>(S, X, w);
  if (w == 0) inv_level<3>(S, X, w);
  if (threadIdx.x == 0) okflag = !bad;
  __syncthreads();
  for (int t = threadIdx.x; t < NB * NB; t += NT) {
    const int r = t % NB, c = t / NB;
    if (r < c) {
      S[r + c * LDT] = T(0);
      X[r + c * LDT] = T(0);
    }
  }
  __syncthreads();
  return okflag != 0;
}

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
__device__ __forceinline__ double* factor_of<double>(const BigFront& B) { return B.F; }
template <>
__device__ __forceinline__ float* factor_of<float>(const BigFront& B) { return B.F32; }
template <class T>
__device__ __forceinline__ T* dinv_of(const BigFront& B);
template <>
__device__ __forceinline__ double* dinv_of<double>(const BigFront& B) { return B.Dinv; }
template <>
__device__ __forceinline__ float* dinv_of<float>(const BigFront& B) { return B.Dinv32; }

template <class T>
__global__ void __launch_bounds__(NT) k_dag_factor(BigFront B, const InstDev* insts, int epoch) {
  __shared__ __align__(16) T SA[NB * LROW];
  __shared__ __align__(16) T SB[NB * LROW];
  __shared__ __align__(16) T SC[NB * LROW];
  __shared__ __align__(16) T SD[NB * LROW];
  __shared__ int task_s;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  double* F = B.F;
  T* Lf = factor_of<T>(B);
  T* Di = dinv_of<T>(B);
  const int LD = B.ld;
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
    stamp(B.trace, t, 0);
    const int4 task = B.tasks[t];
    const int type = task.x, i = task.y, j = task.z, kmax = task.w;
    const int r0i = B.tb[i], hi = B.tb[i + 1] - r0i;
    const int r0j = B.tb[j], hj = B.tb[j + 1] - r0j;
    // the assembled tile (written by earlier launches; only this task writes it here)
    load_tile<false>(SA, F, LD, r0i, r0j, hi, hj);
    __syncthreads();
    Acc<T> C;
    acc_from_lds(C, SA);
    __syncthreads();
    // C −= Σ_k L_ik L_jkᵀ, software pipelined: the tiles of column k+1 are in flight while the
    // MFMAs of column k run from the other LDS buffer pair
    bool alive = true;
    TileRegs<T> RA, RB;
    if (kmax > 0) {
      if (!wait_geq(&B.state[tri_index(i, 0)], epoch, B.abortf) ||
          (i != j && !wait_geq(&B.state[tri_index(j, 0)], epoch, B.abortf)))
        break;
      const int c00 = B.tb[0], w0 = B.tb[1] - c00;
      fetch_tile<true>(RA, Lf, LD, r0i, c00, hi, w0);
      if (i != j) fetch_tile<true>(RB, Lf, LD, r0j, c00, hj, w0);
    }
    for (int k = 0; k < kmax; ++k) {
      T* XA = (k & 1) ? SC : SA;
      T* XB = (k & 1) ? SD : SB;
      put_tile(XA, RA);
      if (i != j) put_tile(XB, RB);
      __syncthreads();  // also: the buffer pair written next was last read before this point
      if (k + 1 < kmax) {
        if (!wait_geq(&B.state[tri_index(i, k + 1)], epoch, B.abortf) ||
            (i != j && !wait_geq(&B.state[tri_index(j, k + 1)], epoch, B.abortf))) {
          alive = false;
          break;
        }
        const int c0n = B.tb[k + 1], wn = B.tb[k + 2] - c0n;
        fetch_tile<true>(RA, Lf, LD, r0i, c0n, hi, wn);
        if (i != j) fetch_tile<true>(RB, Lf, LD, r0j, c0n, hj, wn);
      }
      mfma_acc<true>(C, XA, i != j ? XB : XA, B.tb[k + 1] - B.tb[k]);
    }
    __syncthreads();
    if (!alive) break;
    stamp(B.trace, t, 1);
    if (type == 0) {
      if (j > 0) {
        // L_{j,j−1} = C_{j,j−1} L_{j−1,j−1}⁻ᵀ from the staged tile (fetched before the inverse is
        // ready), then C_jj −= L_{j,j−1} L_{j,j−1}ᵀ
        const int c0p = B.tb[j - 1], wp = B.tb[j] - c0p;
        if (!wait_geq(&B.stage[j - 1], epoch, B.abortf)) break;
        load_tile<true>(SA, Lf, LD, r0j, c0p, hj, wp);
        if (!wait_geq(&B.state[tri_index(j - 1, j - 1)], epoch, B.abortf)) break;
        load_tile<true>(SB, Di + (size_t)(j - 1) * NB * NB, NB, 0, 0, NB, NB);
        __syncthreads();
        Acc<T> X;
        acc_zero(X);
        mfma_acc<false>(X, SA, SB, NB);
        __syncthreads();
        acc_to_lds(X, SA);
        __syncthreads();
        store_tile(Lf, LD, r0j, c0p, hj, wp, SA, false);
        mfma_acc<true>(C, SA, SA, wp);
        publish(&B.state[tri_index(j, j - 1)], epoch);  // the stores drained during the update
      }
      stamp(B.trace, t, 2);
      // factor and invert; L_jj⁻¹ → Dinv and publish, then L_jj → F (read only by later launches)
      acc_to_lds(C, SA);
      __syncthreads();
      bool ok = chol_inv64(SA, SC, SB, hj, B.trace, t);  // SB is free here and lies below 64 KiB
      stamp(B.trace, t, 6);
      if (!ok && threadIdx.x == 0) atomicAdd(&I.sc[S_FAIL], 1.0);
      store_tile(Di + (size_t)j * NB * NB, NB, 0, 0, NB, NB, SC, false);
      publish(&B.state[tri_index(j, j)], epoch);
      stamp(B.trace, t, 7);
      store_tile(Lf, LD, r0j, r0j, hj, hj, SA, true);
    } else if (type == 3) {
      // sub-diagonal tile: hand the accumulated tile to the next diagonal task
      acc_to_lds(C, SA);
      __syncthreads();
      stamp(B.trace, t, 2);
      store_tile(Lf, LD, r0i, r0j, hi, hj, SA, false);
      publish(&B.stage[j], epoch);
    } else if (type == 1) {
      // L_ij = C_ij L_jj⁻ᵀ
      acc_to_lds(C, SA);
      if (!wait_geq(&B.state[tri_index(j, j)], epoch, B.abortf)) break;
      stamp(B.trace, t, 2);
      load_tile<true>(SB, Di + (size_t)j * NB * NB, NB, 0, 0, NB, NB);
      __syncthreads();
      Acc<T> X;
      acc_zero(X);
      mfma_acc<false>(X, SA, SB, NB);
      __syncthreads();
      acc_to_lds(X, SA);
      __syncthreads();
      store_tile(Lf, LD, r0i, r0j, hi, hj, SA, false);
      publish(&B.state[tri_index(i, j)], epoch);
    } else {
      // update-matrix tile: final once every factor column has been subtracted (FP64 for the
      // parent's extend-add)
      acc_to_lds(C, SA);
      __syncthreads();
      stamp(B.trace, t, 2);
      store_tile(F, LD, r0i, r0j, hi, hj, SA, i == j);
    }
    stamp(B.trace, t, 3);
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
  load_tile<false>(S, F, LD, r0, c0, h, w);
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
__global__ void __launch_bounds__(NT) k_big_mtiles(BigFront B, const InstDev* insts) {
  __shared__ __align__(16) double SX[NB * LDT];
  __shared__ __align__(16) double SY[NB * LDT];
  // the factorization that just ended took its last ticket: the next one starts from 0
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.ticket = 0;
  const InstDev& I = insts[B.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  // workgroups [0, Tr) form M_b, workgroups [Tr, 2 Tr) form N_b
  const bool mpart = (int)blockIdx.x < B.Tr;
  const int b = mpart ? (int)blockIdx.x : (int)blockIdx.x - B.Tr;
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

// forward: tile b (ticket order) accumulates L[b, j] x_j for the complete j < b−1 while the
// earlier tiles solve; the factor tiles stream through registers one block ahead of the solution,
// and every block's contribution is summed in block order whatever the batches (reproducible
// results). A factor tile then forms w_b = L_bb⁻¹ (v_b − Σ) before block b−1 arrives, and wave 0
// finishes x_b = w_b − M_b x_{b−1} after it. A separator tile adds its Σ over all factor blocks
// into the front's update vector. TF: element type of the stored factor; the arithmetic is FP64.
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
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  const TF* F = factor_of<TF>(B);
  const int LD = B.ld;
  const int row = threadIdx.x & 63, q = threadIdx.x >> 6;
  const bool diag = b < B.Tr;
  const int jacc = diag ? max(b - 1, 0) : B.Tr;  // blocks summed before the last step
  const int r0 = B.tb[b], h = B.tb[b + 1] - r0;
  if (diag) load_tile<false>(Sinv, dinv_of<TF>(B) + (size_t)b * NB * NB, NB, 0, 0, NB, NB);
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

// L[sep, b]ᵀ x_sep for one factor column block b and one 64-row chunk of the separator rows
// (grid: Tr × ⌈s/64⌉): the tile goes through LDS so that the global reads run down its columns.
// Separator solution values are final before the backward solve of a front starts.
template <class TF>
__global__ void __launch_bounds__(NT) k_big_sepgemv(BigFront B, const FrontDesc* fronts, int f, const InstDev* insts,
                                                    int kv) {
  __shared__ double S[NB * LDT];
  __shared__ double xsep[NB];
  __shared__ double part[4][NB];
  const FrontDesc& Fd = fronts[f];
  const InstDev& I = insts[Fd.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int b = blockIdx.x, rs0 = Fd.r + (int)blockIdx.y * NB;
  const int h = min(NB, B.ld - rs0), c0 = B.tb[b], w = B.tb[b + 1] - c0;
  const double* vall = I.nv + (size_t)kv * I.n;
  load_tile<false>(S, factor_of<TF>(B), B.ld, rs0, c0, h, w);
  if (threadIdx.x < NB) xsep[threadIdx.x] = threadIdx.x < h ? vall[Fd.sepcols[rs0 - Fd.r + threadIdx.x]] : 0.0;
  __syncthreads();
  const int col = threadIdx.x & 63, q = threadIdx.x >> 6;
  double a = 0.0;
#pragma unroll
  for (int rr = q; rr < NB; rr += 4) a += S[rr + col * LDT] * xsep[rr];
  part[q][col] = a;
  __syncthreads();
  if (threadIdx.x < NB)
    B.spart[((size_t)blockIdx.y * B.Tr + b) * NB + col] = (part[0][col] + part[1][col]) + (part[2][col] + part[3][col]);
}

// backward: column block b = Tr−1−ticket accumulates L[j, b]ᵀ x_j for the complete j > b+1:
// thread (r, q) sums L[j, b][r][q + 4i] x_j[r] over the blocks (in descending order whatever the
// batches), one transpose through LDS then sums over r. With L[sep, b]ᵀ x_sep (the chunk partials
// of k_big_sepgemv) it forms w_b = L_bbᵀ⁻¹ (v_b − Σ) before block b+1 arrives, and wave 0
// finishes x_b = w_b − N_bᵀ x_{b+1} after it.
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
  const int b = B.Tr - 1 - tkt;
  double* v = (I.nv + (size_t)kv * I.n) + Fd.res0;
  const TF* F = factor_of<TF>(B);
  const int LD = B.ld;
  const int c0 = B.tb[b], w = B.tb[b + 1] - c0;
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  load_tile<false>(Sinv, dinv_of<TF>(B) + (size_t)b * NB * NB, NB, 0, 0, NB, NB);
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

// ------------------------------------------------------------------------------------------

// The ticket starts at 0 (zeroed with the front's buffers, then put back by k_big_mtiles after
// each factorization, which launch_big_mtiles must follow); the abort flag stays raised once set.
void launch_dag_factor(hipStream_t st, const BigFront& B, const InstDev* I, int grid, int epoch, bool fp32) {
  if (fp32) k_dag_factor<float><<<grid, NT, 0, st>>>(B, I, epoch);
  else k_dag_factor<double><<<grid, NT, 0, st>>>(B, I, epoch);
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
