// 64×64 tile helpers shared by the persistent front kernels and the large PSD cone kernels:
// tile loads and stores through buffer resources spanning exactly the tile (ragged edges read
// zero and drop their stores), LDS staging, and the 64×64 FP64/FP32 MFMA accumulator tile of a
// 256-thread workgroup.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace mb {

#define LDT (NB + 1)           // column stride of a column-major tile in LDS
constexpr int NT = 256;        // threads per workgroup of the tile kernels

// Tiles move through raw buffer instructions over a resource spanning exactly the tile's
// extent in F: a lane outside the h×w tile gets an offset past that extent, which the hardware
// range check turns into a zero load or a dropped store. The cache policy says who wrote the
// data (Coh): None — earlier launches, ordinary cached loads; Dev — another workgroup of the
// running kernel: loads and stores with sc1 (device scope, the policy of the agent-scope
// atomics). A CU's L1 keeps a line it read earlier even after another CU of the same XCD stored
// to it, and no cheaper load or cache invalidate reaches that store, so device scope is needed
// within one XCD too. Either way the loads are ordinary loads to the compiler, so a thread's 16
// loads are in flight together instead of one after another.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int CPOL_SC1 = 16;               // buffer cache policy bit: device scope
constexpr unsigned OFF_OUT = 0x80000000u;  // an offset past every tile extent

enum class Coh { None, Dev };
template <Coh C>
constexpr int load_pol() { return C == Coh::Dev ? CPOL_SC1 : 0; }
template <Coh C>
constexpr int store_pol() { return C == Coh::Dev ? CPOL_SC1 : 0; }

template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const T* F, int LD, int r0, int c0, int h, int w) {
  const T* base = F + (size_t)r0 + (size_t)c0 * LD;
  const unsigned bytes = (unsigned)(((size_t)(w - 1) * LD + h) * sizeof(T));
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), 0, (int)bytes, 0x00020000);
}

template <Coh C>
__device__ __forceinline__ void buf_load(double& v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, load_pol<C>()));
}
template <Coh C>
__device__ __forceinline__ void buf_load(float& v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, load_pol<C>()));
}
template <Coh C>
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, (int)off, 0, store_pol<C>());
}
template <Coh C>
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t rs, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)off, 0, store_pol<C>());
}

// A thread's share of a tile in flight: element t = threadIdx.x + u·NT is (t % NB, t / NB).
template <class T>
struct TileRegs {
  T v[NB * NB / NT];
};

// issue the loads of the h×w tile at (r0, c0) of F (leading dimension LD), zero padded, with the
// cache policy of the tile's writer
template <Coh COH, class T>
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

template <Coh COH, class TD, class TS>
__device__ __forceinline__ void load_tile(TD* S, const TS* F, int LD, int r0, int c0, int h, int w) {
  TileRegs<TS> R;
  fetch_tile<COH>(R, F, LD, r0, c0, h, w);
  put_tile(S, R);
}

// LDS column-major tile → h×w tile at (r0, c0) of F (lower part only if lower), converted to F's
// element type, with the cache policy its readers in the running kernel need
template <Coh COH, class TD, class TS>
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
    buf_store<COH>(v[u], rs, off);
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


}  // namespace mb
