// PSD cones of any order above the LDS kernels' limit (d > PSD_MAX): the dynamic path of
// src/IPM/src/cone/sdp.jl (sdpscaledynamic!, sdpcorrdynamic!, sdpcorr0dynamic!,
// sdpmaxstepdynamic) on d×d matrices in HBM, as a pipeline of batched launches over the large PSD
// cones of an instance:
//   * 64×64 tile products on FP64 MFMA, one workgroup per output tile and cone (k_bp_gemm), for
//     every product of the path (Pᵀ D, P U, R Σ Rᵀ, the corrector's seven products, the max
//     step's congruence, W X W);
//   * blocked Cholesky with inverted 64×64 diagonal blocks and the triangular solve R = P⁻ᵀ U on
//     the same tiles, one workgroup per cone;
//   * one-sided Jacobi SVD (round-robin pairs, one wave per pair, columns in HBM) and the
//     Householder tridiagonalisation + Sturm λmin, one workgroup per cone.
// The scaling block H_v = W ⊗ₛ W (order n = d(d+1)/2, utils.jl:475-513) is never stored: the
// cache keeps W = R Σ Rᵀ, H_v x is svec(W smat(x) W), and the KKT assembly adds δ (W ⊗ₛ W)[a, b]
// into the front as it forms each entry.
//
// Per cone: cache R (d²), S (d²), s (d), W (d²); workspace X0, X1, X2 (d² each), the inverted
// diagonal blocks of the last Cholesky (⌈d/64⌉·64²) and BP_AUX flags / step values.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.h"
#include "kernels.h"
#include "tile64.h"

namespace mb {

namespace {

constexpr double IR2 = 0.70710678118654752440;
constexpr double R2 = 1.41421356237309504880;

// aux slots
enum : int { AX_OK = 0, AX_CHOL0, AX_CHOL1, AX_TAU0, AX_TAU1, AX_LAM, AX_COUNT = 16 };

__device__ __forceinline__ int bp_order(int n) { return (int)((sqrt(1.0 + 8.0 * n) - 1.0) * 0.5 + 0.5); }
__device__ __forceinline__ int svpos(int i, int j, int d) { return j * d - (j * (j - 1)) / 2 + (i - j); }
__device__ __forceinline__ double* cvec(const InstDev& I, int kind) { return I.nv + (size_t)kind * I.n; }

struct BV {
  int d, n, T;
  double *X0, *X1, *X2, *R, *S, *s, *W, *Dinv, *aux;
};

__device__ __forceinline__ BV bview(const ConeDesc& C) {
  BV v;
  v.n = C.n;
  v.d = bp_order(C.n);
  v.T = (v.d + NB - 1) / NB;
  const size_t dd = (size_t)v.d * v.d;
  v.R = C.cache;
  v.S = C.cache + dd;
  v.s = C.cache + 2 * dd;
  v.W = C.cache + 2 * dd + v.d;
  v.X0 = C.ws;
  v.X1 = C.ws + dd;
  v.X2 = C.ws + 2 * dd;
  v.Dinv = C.ws + 3 * dd;
  v.aux = v.Dinv + (size_t)v.T * NB * NB;
  return v;
}

__device__ __forceinline__ double* slot_ptr(const BV& v, int s) {
  switch (s) {
    case BP_X0: return v.X0;
    case BP_X1: return v.X1;
    case BP_X2: return v.X2;
    case BP_R: return v.R;
    case BP_S: return v.S;
    default: return v.W;
  }
}

// (i, j), i ≥ j, of svec position a of a d×d matrix (column j starts at j d − j(j−1)/2)
__device__ __forceinline__ void svec_ij(int a, int d, int& i, int& j) {
  const double b = 2.0 * d + 1.0;
  j = (int)floor(0.5 * (b - sqrt(fmax(b * b - 8.0 * a, 0.0))));
  j = j < 0 ? 0 : (j >= d ? d - 1 : j);
  while (j + 1 < d && svpos(j + 1, j + 1, d) <= a) ++j;
  while (j > 0 && svpos(j, j, d) > a) --j;
  i = j + (a - svpos(j, j, d));
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ double block_sum256(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  const double s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

// S[i + k·LDT] = op(M)[r0 + i][c0 + k] for the 64×64 tile of the nr×nc column-major matrix M
// (leading dimension ld), zero outside it; op = transpose if TR. The loads run along M's columns.
template <bool TR>
__device__ __forceinline__ void gload(double* S, const double* M, int ld, int nr, int nc, int r0, int c0) {
  for (int t = threadIdx.x; t < NB * NB; t += NT) {
    const int fast = t % NB, slow = t / NB;
    if (!TR) {  // element (r0+fast, c0+slow) of M
      const int r = r0 + fast, c = c0 + slow;
      S[fast + slow * LDT] = (r < nr && c < nc) ? M[r + (size_t)c * ld] : 0.0;
    } else {    // op(M)[r0+slow][c0+fast] = M[c0+fast][r0+slow]
      const int r = c0 + fast, c = r0 + slow;
      S[slow + fast * LDT] = (r < nr && c < nc) ? M[r + (size_t)c * ld] : 0.0;
    }
  }
}

// the h×w corner of the LDS tile S (stride LDT) → M at (r0, c0), scaled by a
__device__ __forceinline__ void gstore(double* M, int ld, int nr, int nc, int r0, int c0, const double* S, double a) {
  for (int t = threadIdx.x; t < NB * NB; t += NT) {
    const int i = t % NB, k = t / NB, r = r0 + i, c = c0 + k;
    if (r < nr && c < nc) M[r + (size_t)c * ld] = a * S[i + k * LDT];
  }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// elementwise kernels: grid (cone, chunk of 256·8 entries)
// ------------------------------------------------------------------------------------------

// M ← smat(x) (both triangles), x = vector kind kv at the cone's columns
__global__ void __launch_bounds__(NT) k_bp_smat(const InstDev* insts, const ConeDesc* cones, const int* list, int kv,
                                                int slot) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  const double* x = cvec(I, kv) + C.col0;
  double* M = slot_ptr(v, slot);
  const int dd = v.d * v.d;
  for (int t = blockIdx.y * NT * 8 + threadIdx.x; t < dd && t < (int)(blockIdx.y + 1) * NT * 8; t += NT) {
    const int i = t % v.d, j = t / v.d;
    const int a = i >= j ? i : j, b = i >= j ? j : i;
    const double e = x[svpos(a, b, v.d)];
    M[t] = a == b ? e : e * IR2;
  }
}

// op 0: X1 = R·diag(√s) (the factor Y of W = Y Yᵀ); op 1: X0 = R·diag(σμ − s²) (corr0); op 4: X1 = I;
// op 2: X0 = Ŵ from X2 = A B (corr: Ŵ_ij = −(s_i X_ij + s_j X_ji)/(s_i + s_j), Ŵ_jj = σμ − s_j² −
// X_jj); op 3: the max step's M (X2 symmetrised; side 1: ÷ s_i s_j) into X0 and I + M into X1
__global__ void __launch_bounds__(NT) k_bp_ew(const InstDev* insts, const ConeDesc* cones, const int* list, int op,
                                              int side) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (op == 0 && v.aux[AX_OK] == 0.0) return;
  const double smu = I.sc[S_SMU];
  const int d = v.d, dd = d * d;
  for (int t = blockIdx.y * NT * 8 + threadIdx.x; t < dd && t < (int)(blockIdx.y + 1) * NT * 8; t += NT) {
    const int i = t % d, j = t / d;
    if (op == 0) {
      v.X1[t] = v.R[t] * sqrt(v.s[j]);
    } else if (op == 1) {
      v.X0[t] = v.R[t] * (smu - v.s[j] * v.s[j]);
    } else if (op == 2) {
      double w;
      if (i == j) w = smu - v.s[j] * v.s[j] - v.X2[t];
      else w = -(v.s[i] * v.X2[t] + v.s[j] * v.X2[j + (size_t)i * d]) / (v.s[i] + v.s[j]);
      v.X0[t] = w;
    } else if (op == 3) {
      double m = 0.5 * (v.X2[t] + v.X2[j + (size_t)i * d]);
      if (side == 1) m /= v.s[i] * v.s[j];
      v.X0[t] = m;
      v.X1[t] = m + (i == j ? 1.0 : 0.0);
    } else if (op == 4) {
      v.X1[t] = i == j ? 1.0 : 0.0;
    }
  }
}

// r[col0 + a] = svec(½(M + Mᵀ))_a (mode 0) or y[col0 + a] += c·svec(½(M + Mᵀ))_a (mode 1)
__global__ void __launch_bounds__(NT) k_bp_svec(const InstDev* insts, const ConeDesc* cones, const int* list, int slot,
                                                int kr, int mode, Coef c) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  const double* M = slot_ptr(v, slot);
  double* r = cvec(I, kr) + C.col0;
  const double cv = c.slot < 0 ? c.c : (c.inv ? c.c / I.sc[c.slot] : c.c * I.sc[c.slot]);
  for (int a = blockIdx.y * NT * 8 + threadIdx.x; a < v.n && a < (int)(blockIdx.y + 1) * NT * 8; a += NT) {
    int i, j;
    svec_ij(a, v.d, i, j);
    const double x = 0.5 * (M[i + (size_t)j * v.d] + M[j + (size_t)i * v.d]);
    const double e = i == j ? x : R2 * x;
    if (mode == 0) r[a] = e;
    else r[a] += cv * e;
  }
}

// ------------------------------------------------------------------------------------------
// tile products: C = α op(A) op(B), grid (cone, output tile)
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(NT) k_bp_gemm(const InstDev* insts, const ConeDesc* cones, const int* list, int sa,
                                                int ta, int sb, int tb, int sc, double alpha, int tmax) {
  __shared__ __align__(16) double X[NB * LDT];
  __shared__ __align__(16) double Y[NB * LDT];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  const int ti = blockIdx.y % tmax, tj = blockIdx.y / tmax;
  if (ti >= v.T || tj >= v.T) return;
  const double* A = slot_ptr(v, sa);
  const double* B = slot_ptr(v, sb);
  double* Cm = slot_ptr(v, sc);
  const int d = v.d;
  Acc<double> acc;
  acc_zero(acc);
  for (int kb = 0; kb < v.T; ++kb) {
    if (ta) gload<true>(X, A, d, d, d, ti * NB, kb * NB);
    else gload<false>(X, A, d, d, d, ti * NB, kb * NB);
    if (tb) gload<false>(Y, B, d, d, d, tj * NB, kb * NB);
    else gload<true>(Y, B, d, d, d, tj * NB, kb * NB);
    __syncthreads();
    mfma_acc<false>(acc, X, Y, NB);
    __syncthreads();
  }
  acc_to_lds(acc, X);
  __syncthreads();
  gstore(Cm, d, d, d, ti * NB, tj * NB, X, alpha);
}

// ------------------------------------------------------------------------------------------
// blocked Cholesky of the matrix in `slot` (lower; the strict upper triangle is zeroed), the
// inverted diagonal blocks into Dinv; aux[flag] = 1 on success, 0 on a non-positive pivot. One
// workgroup per cone. skip_if_failed: do nothing when the scaling already failed.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_bp_chol(const InstDev* insts, const ConeDesc* cones, const int* list, int slot,
                                                int flag, int skip_if_failed) {
  __shared__ __align__(16) double SA[NB * LDT];
  __shared__ __align__(16) double SB[NB * LDT];
  __shared__ __align__(16) double SX[NB * LDT];
  __shared__ double SW[3 * NB];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (skip_if_failed && v.aux[AX_OK] == 0.0) return;
  double* M = slot_ptr(v, slot);
  const int d = v.d, T = v.T;
  bool ok = true;
  for (int kb = 0; kb < T && ok; ++kb) {
    const int h = d - kb * NB < NB ? d - kb * NB : NB;
    gload<false>(SA, M, d, d, d, kb * NB, kb * NB);
    __syncthreads();
    ok = chol_inv64(SA, SX, SW, h, nullptr, 0);
    if (!ok) break;
    // L_kk (lower part) and L_kk⁻¹
    for (int t = threadIdx.x; t < NB * NB; t += NT) {
      const int i = t % NB, k = t / NB;
      if (i < h && k < h && i >= k) M[(kb * NB + i) + (size_t)(kb * NB + k) * d] = SA[i + k * LDT];
      v.Dinv[(size_t)kb * NB * NB + t] = SX[i + k * LDT];
    }
    // L_ik = A_ik L_kk⁻ᵀ below the diagonal block
    for (int ib = kb + 1; ib < T; ++ib) {
      gload<false>(SB, M, d, d, d, ib * NB, kb * NB);
      __syncthreads();
      Acc<double> acc;
      acc_zero(acc);
      mfma_acc<false>(acc, SB, SX, NB);
      __syncthreads();
      acc_to_lds(acc, SB);
      __syncthreads();
      gstore(M, d, d, d, ib * NB, kb * NB, SB, 1.0);
      __syncthreads();
    }
    // trailing update A_ij −= L_ik L_jkᵀ (whole tiles, i ≥ j)
    for (int jb = kb + 1; jb < T; ++jb)
      for (int ib = jb; ib < T; ++ib) {
        gload<false>(SX, M, d, d, d, ib * NB, jb * NB);
        gload<false>(SA, M, d, d, d, ib * NB, kb * NB);
        gload<false>(SB, M, d, d, d, jb * NB, kb * NB);
        __syncthreads();
        Acc<double> acc;
        acc_from_lds(acc, SX);
        mfma_acc<true>(acc, SA, SB, NB);
        __syncthreads();
        acc_to_lds(acc, SX);
        __syncthreads();
        gstore(M, d, d, d, ib * NB, jb * NB, SX, 1.0);
        __syncthreads();
      }
  }
  if (ok)
    for (int t = threadIdx.x; t < d * d; t += NT) {
      const int i = t % d, j = t / d;
      if (i < j) M[t] = 0.0;
    }
  if (threadIdx.x == 0) {
    v.aux[flag] = ok ? 1.0 : 0.0;
    if (flag == AX_CHOL0 && !ok) v.aux[AX_OK] = 0.0;
  }
}

// the scaling's flag starts at 1 each scale!, and Cholesky of D also feeds it
__global__ void k_bp_flag_init(const InstDev* insts, const ConeDesc* cones, const int* list) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || threadIdx.x != 0) return;
  const BV v = bview(C);
  v.aux[AX_OK] = 1.0;
}
__global__ void k_bp_flag_and(const InstDev* insts, const ConeDesc* cones, const int* list, int flag) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || threadIdx.x != 0) return;
  const BV v = bview(C);
  if (v.aux[flag] == 0.0) v.aux[AX_OK] = 0.0;
}

// ------------------------------------------------------------------------------------------
// R = P⁻ᵀ U: out = L⁻ᵀ B with L lower (slot sl, its inverted diagonal blocks in Dinv), backward
// over tile rows; one workgroup per cone
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_bp_trsm_lt(const InstDev* insts, const ConeDesc* cones, const int* list, int sl,
                                                   int sbm, int so, int flag) {
  __shared__ __align__(16) double SA[NB * LDT];
  __shared__ __align__(16) double SB[NB * LDT];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (v.aux[flag] == 0.0) return;  // the Cholesky factor L is not there
  const double* L = slot_ptr(v, sl);
  const double* Bm = slot_ptr(v, sbm);
  double* O = slot_ptr(v, so);
  const int d = v.d, T = v.T;
  for (int ib = T - 1; ib >= 0; --ib)
    for (int jb = 0; jb < T; ++jb) {
      gload<false>(SA, Bm, d, d, d, ib * NB, jb * NB);
      __syncthreads();
      Acc<double> acc;
      acc_from_lds(acc, SA);
      __syncthreads();
      for (int k = ib + 1; k < T; ++k) {  // acc −= L(k, ib)ᵀ O(k, jb)
        gload<true>(SA, L, d, d, d, ib * NB, k * NB);
        gload<true>(SB, O, d, d, d, jb * NB, k * NB);
        __syncthreads();
        mfma_acc<true>(acc, SA, SB, NB);
        __syncthreads();
      }
      // O(ib, jb) = L_ib⁻ᵀ acc: X = Dinv_ibᵀ, Y = accᵀ
      acc_to_lds(acc, SB);
      __syncthreads();
      gload<true>(SA, v.Dinv + (size_t)ib * NB * NB, NB, NB, NB, 0, 0);
      for (int t = threadIdx.x; t < NB * NB; t += NT) {  // SB ← SBᵀ in place (pairs swapped once)
        const int i = t % NB, k = t / NB;
        if (i < k) {
          const double a = SB[i + k * LDT], b = SB[k + i * LDT];
          SB[i + k * LDT] = b;
          SB[k + i * LDT] = a;
        }
      }
      __syncthreads();
      Acc<double> o;
      acc_zero(o);
      mfma_acc<false>(o, SA, SB, NB);
      __syncthreads();
      acc_to_lds(o, SA);
      __syncthreads();
      gstore(O, d, d, d, ib * NB, jb * NB, SA, 1.0);
      __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// one-sided Jacobi SVD of the columns of X2 (svdjacobi!, eig.jl:1-174): round-robin pairs, one
// wave per pair (lanes along the rows), ≤ 30 sweeps, rotation threshold 1e-15·√(αβ); then s =
// column norms, X2 = normalised columns (U). Failure (no convergence, a zero singular value)
// clears aux[AX_OK]. One workgroup per cone; slot: LDS round-robin table of d+1 entries.
// ------------------------------------------------------------------------------------------
constexpr int BP_DMAX_JACOBI = 4096;
constexpr int BP_JACOBI_LDS = 128;  // orders up to this run the LDS kernel

__global__ void __launch_bounds__(NT) k_bp_jacobi(const InstDev* insts, const ConeDesc* cones, const int* list) {
  __shared__ int slot[BP_DMAX_JACOBI + 2];
  __shared__ int rotated, zero;
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (v.d <= BP_JACOBI_LDS || v.aux[AX_OK] == 0.0) return;  // k_bp_jacobi_lds
  const int d = v.d, dd = d + (d & 1), npair = dd / 2;
  double* A = v.X2;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < dd; t += NT) slot[t] = t;  // index d (odd d) is a blank
  __syncthreads();
  bool conv = false;
  for (int sweep = 0; sweep < 30 && !conv; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < dd - 1; ++round) {
      for (int q = w; q < npair; q += NT / 64) {
        const int a = slot[q], b = slot[dd - 1 - q];
        if (a >= d || b >= d) continue;
        double* xa = A + (size_t)a * d;
        double* xb = A + (size_t)b * d;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < d; i += 64) {
          const double x = xa[i], y = xb[i];
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double tn = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + tn * tn), sn = c * tn;
          for (int i = lane; i < d; i += 64) {
            const double x = xa[i], y = xb[i];
            xa[i] = c * x - sn * y;
            xb[i] = sn * x + c * y;
          }
          if (lane == 0) rotated = 1;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {  // rotate every slot but the first
        const int last = slot[dd - 1];
        for (int k = dd - 1; k > 1; --k) slot[k] = slot[k - 1];
        slot[1] = last;
      }
      __syncthreads();
    }
    conv = rotated == 0;
    __syncthreads();
  }
  if (threadIdx.x == 0) zero = 0;
  __syncthreads();
  for (int j = w; j < d; j += NT / 64) {  // s_j = ‖column j‖, one wave per column
    double nr = 0.0;
    for (int i = lane; i < d; i += 64) nr = fma(A[i + (size_t)j * d], A[i + (size_t)j * d], nr);
    nr = sqrt(wave_sum(nr));
    if (lane == 0) {
      v.s[j] = nr;
      if (!(nr > 0.0)) zero = 1;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < d * d; t += NT) {
    const double sj = v.s[t / d];
    if (sj > 0.0) A[t] /= sj;
  }
  if (threadIdx.x == 0 && !(conv && zero == 0)) v.aux[AX_OK] = 0.0;
}

// The same one-sided Jacobi for cones of order ≤ BP_JACOBI_LDS with the matrix in LDS (128 KiB at
// order 128) and sixteen waves: the column pairs, their rotations and the norms are the HBM
// kernel's, in the same order, so the results are the same bit for bit.
constexpr int NTJ = 1024;
__global__ void __launch_bounds__(NTJ) k_bp_jacobi_lds(const InstDev* insts, const ConeDesc* cones, const int* list) {
  __shared__ double S[BP_JACOBI_LDS * BP_JACOBI_LDS];
  __shared__ int slot[BP_JACOBI_LDS + 2];
  __shared__ int rotated, zero;
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (v.d > BP_JACOBI_LDS || v.aux[AX_OK] == 0.0) return;
  const int d = v.d, dd = d + (d & 1), npair = dd / 2;
  double* A = v.X2;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < d * d; t += NTJ) S[t] = A[t];
  for (int t = threadIdx.x; t < dd; t += NTJ) slot[t] = t;  // index d (odd d) is a blank
  __syncthreads();
  bool conv = false;
  for (int sweep = 0; sweep < 30 && !conv; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < dd - 1; ++round) {
      for (int q = w; q < npair; q += NTJ / 64) {
        const int a = slot[q], b = slot[dd - 1 - q];
        if (a >= d || b >= d) continue;
        double* xa = S + a * d;
        double* xb = S + b * d;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < d; i += 64) {
          const double x = xa[i], y = xb[i];
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double tn = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + tn * tn), sn = c * tn;
          for (int i = lane; i < d; i += 64) {
            const double x = xa[i], y = xb[i];
            xa[i] = c * x - sn * y;
            xb[i] = sn * x + c * y;
          }
          if (lane == 0) rotated = 1;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {  // rotate every slot but the first
        const int last = slot[dd - 1];
        for (int k = dd - 1; k > 1; --k) slot[k] = slot[k - 1];
        slot[1] = last;
      }
      __syncthreads();
    }
    conv = rotated == 0;
    __syncthreads();
  }
  if (threadIdx.x == 0) zero = 0;
  __syncthreads();
  for (int j = w; j < d; j += NTJ / 64) {  // s_j = ‖column j‖, one wave per column
    double nr = 0.0;
    for (int i = lane; i < d; i += 64) nr = fma(S[i + j * d], S[i + j * d], nr);
    nr = sqrt(wave_sum(nr));
    if (lane == 0) {
      v.s[j] = nr;
      if (!(nr > 0.0)) zero = 1;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < d * d; t += NTJ) {
    const double sj = v.s[t / d];
    A[t] = sj > 0.0 ? S[t] / sj : S[t];
  }
  if (threadIdx.x == 0 && !(conv && zero == 0)) v.aux[AX_OK] = 0.0;
}

// the scaling's results: out[2c] = flag, out[2c+1] = Σ 1/s_i² (sdp.jl:253-262)
__global__ void k_bp_scale_out(const InstDev* insts, const ConeDesc* cones, const int* list, double* out) {
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || threadIdx.x != 0) return;
  const BV v = bview(C);
  const bool ok = v.aux[AX_OK] != 0.0;
  double spsd = 0.0;
  if (ok)
    for (int k = 0; k < v.d; ++k) spsd += 1.0 / (v.s[k] * v.s[k]);
  out[2 * c] = ok ? 1.0 : 0.0;
  out[2 * c + 1] = spsd;
}

// ------------------------------------------------------------------------------------------
// λmin of the symmetric matrix X0 (destroyed), when the Cholesky test of I + M failed: Householder
// tridiagonalisation, then Sturm counts by multisection over the Gershgorin interval
// (eigminsturm!, eig.jl:180-589); aux[AX_TAU0 + side] = 1 if the test passed, else
// 1/max(1, −λmin). One workgroup per cone; v_, p_: 2d doubles of X1 after the test.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_bp_eigmin(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                  int side) {
  __shared__ double red[4];
  __shared__ double sh[4];
  __shared__ double lohi[2];
  __shared__ int first;
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (v.aux[AX_CHOL1] != 0.0) {
    if (threadIdx.x == 0) v.aux[AX_TAU0 + side] = 1.0;
    return;
  }
  const int d = v.d;
  double* M = v.X0;
  double* vv = v.X1;      // Householder vector
  double* pp = v.X1 + d;  // p, then w
  for (int k = 0; k + 2 < d; ++k) {
    const int m = d - k - 1;
    double* colk = M + (k + 1) + (size_t)k * d;  // M[k+1 .. d−1, k]
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += NT) s = fma(colk[i], colk[i], s);
    s = block_sum256(s, red);
    const double nx = sqrt(s);
    if (nx == 0.0) continue;  // uniform: every thread has the same sum
    const double x0 = colk[0];
    const double alpha = x0 >= 0.0 ? -nx : nx;
    double vsq = 0.0;
    for (int i = threadIdx.x; i < m; i += NT) {
      const double e = colk[i] - (i == 0 ? alpha : 0.0);
      vv[i] = e;
      vsq = fma(e, e, vsq);
    }
    vsq = block_sum256(vsq, red);
    if (vsq == 0.0) continue;
    const double beta = 2.0 / vsq;
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += NT) {  // p = β M22 v (M22 symmetric: read column i)
      const double* col = M + (k + 1) + (size_t)(k + 1 + i) * d;
      double a = 0.0;
      for (int j = 0; j < m; ++j) a = fma(col[j], vv[j], a);
      pp[i] = beta * a;
    }
    __syncthreads();
    double vp = 0.0;
    for (int i = threadIdx.x; i < m; i += NT) vp = fma(vv[i], pp[i], vp);
    vp = block_sum256(vp, red);
    const double K = 0.5 * beta * vp;
    for (int i = threadIdx.x; i < m; i += NT) pp[i] -= K * vv[i];  // w
    __syncthreads();
    for (int t = threadIdx.x; t < m * m; t += NT) {
      const int i = t % m, j = t / m;
      M[(k + 1 + i) + (size_t)(k + 1 + j) * d] -= vv[i] * pp[j] + pp[i] * vv[j];
    }
    for (int i = threadIdx.x; i < m; i += NT) {
      const double x = i == 0 ? alpha : 0.0;
      colk[i] = x;
      M[k + (size_t)(k + 1 + i) * d] = x;
    }
    __syncthreads();
  }
  // Gershgorin interval of the tridiagonal (diagonal M_ii, off-diagonal M_{i,i−1})
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < d; i += NT) {
    const double r = (i > 0 ? fabs(M[i + (size_t)(i - 1) * d]) : 0.0) + (i + 1 < d ? fabs(M[i + 1 + (size_t)i * d]) : 0.0);
    lo = fmin(lo, M[i + (size_t)i * d] - r);
    hi = fmax(hi, M[i + (size_t)i * d] + r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = lo;
    red[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    lohi[0] = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
    lohi[1] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  }
  __syncthreads();
  for (int round = 0; round < 8; ++round) {
    const double a = lohi[0], b = lohi[1];
    if (threadIdx.x == 0) first = NT;
    __syncthreads();
    const double x = a + (b - a) * (double)(threadIdx.x + 1) / (double)(NT + 1);
    int cnt = 0;
    double q = M[0] - x;
    if (q < 0.0) ++cnt;
    for (int i = 1; i < d; ++i) {
      if (q == 0.0) q = 1e-300;
      const double e = M[i + (size_t)(i - 1) * d];
      q = M[i + (size_t)i * d] - x - e * e / q;
      if (q < 0.0) ++cnt;
    }
    if (cnt >= 1) atomicMin(&first, (int)threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int f = first;
      const double step = (b - a) / (double)(NT + 1);
      lohi[0] = a + step * (double)f;
      lohi[1] = f < NT ? a + step * (double)(f + 1) : b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double lam = 0.5 * (lohi[0] + lohi[1]);
    v.aux[AX_TAU0 + side] = 1.0 / fmax(1.0, -lam);
  }
}

__global__ void k_bp_step_out(const InstDev* insts, const ConeDesc* cones, const int* list, double step_frac,
                              double* out) {
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0 || threadIdx.x != 0) return;
  const BV v = bview(C);
  out[c] = step_frac * fmin(v.aux[AX_TAU0], v.aux[AX_TAU1]);
}

// ------------------------------------------------------------------------------------------
// ‖Q_vv + W ⊗ₛ W‖²: with no Q block, ½(‖W‖⁴ + ‖W²‖²) (the eigenvalues of W ⊗ₛ W are λ_iλ_j,
// i ≤ j), X0 = W² formed before; with a Q block, entry by entry. grid (cone, DENSE_SPLIT chunks).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_bp_hnorm2(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                  double* part) {
  __shared__ double red[4];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  double s0 = 0.0, s1 = 0.0;
  const BV v = bview(C);
  if (I.sc[S_ACTIVE] != 0.0) {
    if (!C.Qd) {
      const size_t dd = (size_t)v.d * v.d;
      const size_t chunk = (dd + DENSE_SPLIT - 1) / DENSE_SPLIT;
      const size_t lo = chunk * blockIdx.y, hi = lo + chunk < dd ? lo + chunk : dd;
      for (size_t t = lo + threadIdx.x; t < hi; t += NT) {
        s0 = fma(v.W[t], v.W[t], s0);
        s1 = fma(v.X0[t], v.X0[t], s1);
      }
    } else {
      const size_t nn = (size_t)v.n * v.n;
      const size_t chunk = (nn + DENSE_SPLIT - 1) / DENSE_SPLIT;
      const size_t lo = chunk * blockIdx.y, hi = lo + chunk < nn ? lo + chunk : nn;
      for (size_t t = lo + threadIdx.x; t < hi; t += NT) {
        int i, j, k, l;
        svec_ij((int)(t % v.n), v.d, i, j);
        svec_ij((int)(t / v.n), v.d, k, l);
        const double* W = v.W;
        const size_t d = v.d;
        double h = 0.5 * (W[i + k * d] * W[j + l * d] + W[i + l * d] * W[j + k * d]);
        if (i != j) h *= R2;
        if (k != l) h *= R2;
        h += C.Qd[t];
        s0 = fma(h, h, s0);
      }
    }
  }
  s0 = block_sum256(s0, red);
  s1 = block_sum256(s1, red);
  if (threadIdx.x == 0) {
    part[2 * ((size_t)blockIdx.x * DENSE_SPLIT + blockIdx.y)] = s0;
    part[2 * ((size_t)blockIdx.x * DENSE_SPLIT + blockIdx.y) + 1] = s1;
  }
}

__global__ void k_bp_hnorm2_sum(const InstDev* insts, const ConeDesc* cones, const int* list, const double* part,
                                double* out) {
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (threadIdx.x != 0 || I.sc[S_ACTIVE] == 0.0) return;
  double s0 = 0.0, s1 = 0.0;
  for (int k = 0; k < DENSE_SPLIT; ++k) {
    s0 += part[2 * ((size_t)blockIdx.x * DENSE_SPLIT + k)];
    s1 += part[2 * ((size_t)blockIdx.x * DENSE_SPLIT + k) + 1];
  }
  out[c] = C.Qd ? s0 : 0.5 * (s0 * s0 + s1);
}

// ------------------------------------------------------------------------------------------
// assembly: F[front] += δ (W ⊗ₛ W)[a, b] for a ≥ b over the cone's columns; grid (cone, column
// group), threads along a column. blocks: the whole n×n block (column-major) into D instead.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_bp_assemble(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                    const int* cone_front, const FrontDesc* fronts, double* D) {
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  const size_t d = v.d;
  const double* W = v.W;
  const FrontDesc* Fd = D ? nullptr : &fronts[cone_front[c]];
  const double delta = D ? 1.0 : I.sc[S_DELTA];
  const int off = D ? 0 : C.col0 - Fd->res0;
  for (int b = blockIdx.y; b < v.n; b += gridDim.y) {
    int k, l;
    svec_ij(b, v.d, k, l);
    const double cb = k != l ? R2 : 1.0;
    double* col = D ? D + (size_t)b * v.n : Fd->F + (size_t)(off + b) * Fd->ld + off;
    for (int a = (D ? 0 : b) + threadIdx.x; a < v.n; a += NT) {
      int i, j;
      svec_ij(a, v.d, i, j);
      double h = 0.5 * (W[i + k * d] * W[j + l * d] + W[i + l * d] * W[j + k * d]) * cb;
      if (i != j) h *= R2;
      if (D) col[a] = h;
      else col[a] += delta * h;
    }
  }
}

// ------------------------------------------------------------------------------------------
// launch sequences
// ------------------------------------------------------------------------------------------

namespace {
int ew_chunks(int dmax) { return std::max(1, (dmax * dmax + NT * 8 - 1) / (NT * 8)); }
int sv_chunks(int dmax) { return std::max(1, (dmax * (dmax + 1) / 2 + NT * 8 - 1) / (NT * 8)); }
int tiles(int dmax) { return (dmax + NB - 1) / NB; }

void gemm(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int sa, int ta, int sb, int tb, int sc,
          double alpha = 1.0) {
  const int t = tiles(B.dmax);
  k_bp_gemm<<<dim3(B.n, t * t), NT, 0, st>>>(I, C, B.list, sa, ta, sb, tb, sc, alpha, t);
}
void smat(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kv, int slot) {
  k_bp_smat<<<dim3(B.n, ew_chunks(B.dmax)), NT, 0, st>>>(I, C, B.list, kv, slot);
}
void ew(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int op, int side) {
  k_bp_ew<<<dim3(B.n, ew_chunks(B.dmax)), NT, 0, st>>>(I, C, B.list, op, side);
}
void svec(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int slot, int kr, int mode, Coef c) {
  k_bp_svec<<<dim3(B.n, sv_chunks(B.dmax)), NT, 0, st>>>(I, C, B.list, slot, kr, mode, c);
}
}  // namespace

void launch_bpsd_scale(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, double* out) {
  if (!B.n) return;
  k_bp_flag_init<<<B.n, 64, 0, st>>>(I, C, B.list);
  smat(st, I, C, B, NV_P, BP_X0);
  smat(st, I, C, B, NV_D, BP_X1);
  k_bp_chol<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X1, AX_CHOL1, 1);  // LD (its inverse blocks are not kept)
  k_bp_flag_and<<<B.n, 64, 0, st>>>(I, C, B.list, AX_CHOL1);
  k_bp_chol<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X0, AX_CHOL0, 1);  // LP, inverse blocks for the solve
  gemm(st, I, C, B, BP_X0, 1, BP_X1, 0, BP_X2);                       // LPᵀ LD
  k_bp_jacobi_lds<<<B.n, NTJ, 0, st>>>(I, C, B.list);                // X2 = U, s = Σ (d ≤ 128)
  if (B.dmax > BP_JACOBI_LDS) k_bp_jacobi<<<B.n, NT, 0, st>>>(I, C, B.list);  // the larger cones
  gemm(st, I, C, B, BP_X0, 0, BP_X2, 0, BP_S);                        // S = LP U
  k_bp_trsm_lt<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X0, BP_X2, BP_R, AX_OK);  // R = LP⁻ᵀ U
  ew(st, I, C, B, 0, 0);                                              // X1 = R Σ^½
  gemm(st, I, C, B, BP_X1, 0, BP_X1, 1, BP_W);                        // W = X1 X1ᵀ
  k_bp_scale_out<<<B.n, 64, 0, st>>>(I, C, B.list, out);
}

void launch_bpsd_corr0(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kr) {
  if (!B.n) return;
  ew(st, I, C, B, 1, 0);                              // X0 = R (σμ I − Σ²)
  gemm(st, I, C, B, BP_X0, 0, BP_R, 1, BP_X1);        // X1 = X0 Rᵀ
  svec(st, I, C, B, BP_X1, kr, 0, K(0.0));
}

void launch_bpsd_corr(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kr, int kdp,
                      int kdd) {
  if (!B.n) return;
  smat(st, I, C, B, kdp, BP_X0);
  smat(st, I, C, B, kdd, BP_X1);
  gemm(st, I, C, B, BP_X0, 0, BP_R, 0, BP_X2);  // ΔP R
  gemm(st, I, C, B, BP_R, 1, BP_X2, 0, BP_X0);  // A = Rᵀ ΔP R
  gemm(st, I, C, B, BP_X1, 0, BP_S, 0, BP_X2);  // ΔD S
  gemm(st, I, C, B, BP_S, 1, BP_X2, 0, BP_X1);  // B = Sᵀ ΔD S
  gemm(st, I, C, B, BP_X0, 0, BP_X1, 0, BP_X2); // A B
  ew(st, I, C, B, 2, 0);                        // X0 = Ŵ
  gemm(st, I, C, B, BP_X0, 0, BP_R, 1, BP_X1);  // Ŵ Rᵀ
  gemm(st, I, C, B, BP_R, 0, BP_X1, 0, BP_X2);  // R Ŵ Rᵀ
  svec(st, I, C, B, BP_X2, kr, 0, K(0.0));
}

void launch_bpsd_maxstep(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kdp, int kdd,
                         double step_frac, double* out) {
  if (!B.n) return;
  for (int side = 0; side < 2; ++side) {
    const int F = side == 0 ? BP_R : BP_S;
    smat(st, I, C, B, side == 0 ? kdp : kdd, BP_X0);
    gemm(st, I, C, B, BP_X0, 0, F, 0, BP_X1);  // ΔX F
    gemm(st, I, C, B, F, 1, BP_X1, 0, BP_X2);  // Fᵀ ΔX F
    ew(st, I, C, B, 3, side);                  // X0 = M, X1 = I + M
    k_bp_chol<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X1, AX_CHOL1, 0);
    k_bp_eigmin<<<B.n, NT, 0, st>>>(I, C, B.list, side);
  }
  k_bp_step_out<<<B.n, 64, 0, st>>>(I, C, B.list, step_frac, out);
}

void launch_bpsd_matvec(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kx, int ky, Coef a) {
  if (!B.n) return;
  smat(st, I, C, B, kx, BP_X0);
  gemm(st, I, C, B, BP_W, 0, BP_X0, 0, BP_X1);  // W X
  gemm(st, I, C, B, BP_X1, 0, BP_W, 0, BP_X2);  // W X W
  svec(st, I, C, B, BP_X2, ky, 1, a);
}

void launch_bpsd_hnorm2(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, double* out,
                        double* part) {
  if (!B.n) return;
  gemm(st, I, C, B, BP_W, 0, BP_W, 0, BP_X0);  // W²
  k_bp_hnorm2<<<dim3(B.n, DENSE_SPLIT), NT, 0, st>>>(I, C, B.list, part);
  k_bp_hnorm2_sum<<<B.n, 64, 0, st>>>(I, C, B.list, part, out);
}

void launch_bpsd_assemble(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, const int* cone_front,
                          const FrontDesc* F) {
  if (!B.n) return;
  const int groups = std::max(1, std::min(1024, B.dmax * (B.dmax + 1) / 2 / 4));
  k_bp_assemble<<<dim3(B.n, groups), NT, 0, st>>>(I, C, B.list, cone_front, F, nullptr);
}

// sdpshadow! on the large cones: X = smat(x) = L Lᵀ, Y = L⁻ᵀ (the solve on I), X⁻¹ = Y Yᵀ → y;
// y stays untouched where X is not positive definite
__global__ void k_bp_svec_if(const InstDev* insts, const ConeDesc* cones, const int* list, int slot, int kr, int flag) {
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const BV v = bview(C);
  if (v.aux[flag] == 0.0) return;
  const double* M = slot_ptr(v, slot);
  double* r = cvec(I, kr) + C.col0;
  for (int a = blockIdx.y * NT * 8 + threadIdx.x; a < v.n && a < (int)(blockIdx.y + 1) * NT * 8; a += NT) {
    int i, j;
    svec_ij(a, v.d, i, j);
    const double x = 0.5 * (M[i + (size_t)j * v.d] + M[j + (size_t)i * v.d]);
    r[a] = i == j ? x : R2 * x;
  }
}

void launch_bpsd_dshadow(hipStream_t st, const InstDev* I, const ConeDesc* C, const BigPsd& B, int kx, int ky) {
  if (!B.n) return;
  smat(st, I, C, B, kx, BP_X0);
  k_bp_chol<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X0, AX_CHOL1, 0);
  ew(st, I, C, B, 4, 0);  // X1 = I
  k_bp_trsm_lt<<<B.n, NT, 0, st>>>(I, C, B.list, BP_X0, BP_X1, BP_X2, AX_CHOL1);
  gemm(st, I, C, B, BP_X2, 0, BP_X2, 1, BP_X1);  // L⁻ᵀ L⁻¹
  k_bp_svec_if<<<dim3(B.n, sv_chunks(B.dmax)), NT, 0, st>>>(I, C, B.list, BP_X1, ky, AX_CHOL1);
}

void launch_bpsd_block(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* one, int n, double* D) {
  const int groups = std::max(1, std::min(1024, n));
  k_bp_assemble<<<dim3(1, groups), NT, 0, st>>>(I, C, one, nullptr, nullptr, D);
}

}  // namespace mb
