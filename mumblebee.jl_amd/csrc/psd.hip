// Positive-semidefinite cone kernels (src/IPM/src/cone/sdp.jl) and the dense-block kernels
// shared by every cone whose scaling block is a full matrix (second-order and PSD).
//
// One workgroup per PSD cone; the cone's d×d matrices live in LDS (d ≤ PSD_MAX). Vectors are in
// svec form: the lower triangle column by column with off-diagonal entries times √2
// (cone/utils.jl:570-664). Per-cone cache: R (d², column-major), S (d²), s (d), where with
// P = LP LPᵀ, D = LD LDᵀ and LPᵀ LD = U Σ Vᵀ: R = LP⁻ᵀ U, S = LP U, s = Σ (sdp.jl:10-30).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.h"
#include "kernels.h"

namespace mb {

constexpr int PT = 256;           // threads per workgroup
constexpr int PLD = PSD_MAX + 1;  // LDS column stride
constexpr int PM = PSD_MAX * PLD; // doubles per LDS matrix
constexpr double IR2 = 0.70710678118654752440;
constexpr double R2 = 1.41421356237309504880;

__device__ __forceinline__ double* cvec(const InstDev& I, int kind) { return I.nv + (size_t)kind * I.n; }

// position of (i, j), i ≥ j, in the svec vector of a d×d matrix
__device__ __forceinline__ int svpos(int i, int j, int d) { return j * d - (j * (j - 1)) / 2 + (i - j); }

__device__ __forceinline__ int psd_order(int n) { return (int)((sqrt(1.0 + 8.0 * n) - 1.0) * 0.5 + 0.5); }

// M ← smat(v), both triangles
__device__ void smat_lds(double* M, const double* v, int d) {
  for (int t = threadIdx.x; t < d * d; t += PT) {
    int i = t % d, j = t / d;
    int a = i >= j ? i : j, b = i >= j ? j : i;
    double x = v[svpos(a, b, d)];
    M[i + j * PLD] = a == b ? x : x * IR2;
  }
}

// C ← op(A) op(B) for d×d LDS matrices (TA / TB: transpose)
template <bool TA, bool TB>
__device__ void gemm_lds(double* C, const double* A, const double* B, int d) {
  for (int t = threadIdx.x; t < d * d; t += PT) {
    int i = t % d, j = t / d;
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
      double a = TA ? A[k + i * PLD] : A[i + k * PLD];
      double b = TB ? B[j + k * PLD] : B[k + j * PLD];
      s = fma(a, b, s);
    }
    C[i + j * PLD] = s;
  }
}

// In-place lower Cholesky of A (and of B when given), right-looking; false on a non-positive
// pivot of either.
__device__ bool chol_lds(double* A, double* B, int d) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int nm = B ? 2 : 1;
  for (int k = 0; k < d; ++k) {
    if (threadIdx.x < nm) {
      double* M = threadIdx.x == 0 ? A : B;
      double p = M[k + k * PLD];
      if (!(p > 0.0) || !isfinite(p)) {
        bad = 1;
        p = 1.0;
      }
      M[k + k * PLD] = sqrt(p);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nm * d; t += PT) {
      int i = t % d;
      double* M = t < d ? A : B;
      if (i > k) M[i + k * PLD] /= M[k + k * PLD];
    }
    __syncthreads();
    const int m = d - k - 1;
    for (int t = threadIdx.x; t < nm * m * m; t += PT) {
      int r = t % (m * m);
      double* M = t < m * m ? A : B;
      int i = k + 1 + r % m, j = k + 1 + r / m;
      if (i >= j) M[i + j * PLD] -= M[i + k * PLD] * M[j + k * PLD];
    }
    __syncthreads();
  }
  return bad == 0;
}

__device__ __forceinline__ double sum8(double v) {
  for (int o = 4; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One-sided Jacobi SVD of the columns of A (sdp.jl's svdjacobi!, eig.jl:1-174): rotate column
// pairs until every pair is orthogonal to working precision; then s = column norms and A = the
// normalised columns (left singular vectors). Round-robin pair schedule, d/2 disjoint pairs per
// round, 8 threads per pair. False if a singular value is zero or the sweeps do not converge.
__device__ bool svd_jacobi(double* A, double* s, int d, int* slot) {
  __shared__ int rotated;
  const int dd = d + (d & 1), npair = dd / 2;
  for (int t = threadIdx.x; t < dd; t += PT) slot[t] = t;  // index d (odd d) is a blank
  __syncthreads();
  bool conv = false;
  for (int sweep = 0; sweep < 30 && !conv; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < dd - 1; ++round) {
      const int q = threadIdx.x >> 3, l = threadIdx.x & 7;
      const int a = q < npair ? slot[q] : d, b = q < npair ? slot[dd - 1 - q] : d;
      const bool real = a < d && b < d;
      double al = 0.0, be = 0.0, ga = 0.0;
      if (real)
        for (int i = l; i < d; i += 8) {
          double x = A[i + a * PLD], y = A[i + b * PLD];
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
      al = sum8(al);
      be = sum8(be);
      ga = sum8(ga);
      if (real && ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
        const double zeta = (be - al) / (2.0 * ga);
        const double tn = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tn * tn), sn = c * tn;
        for (int i = l; i < d; i += 8) {
          double x = A[i + a * PLD], y = A[i + b * PLD];
          A[i + a * PLD] = c * x - sn * y;
          A[i + b * PLD] = sn * x + c * y;
        }
        if (l == 0) rotated = 1;
      }
      __syncthreads();
      if (threadIdx.x == 0) {  // rotate every slot but the first
        int last = slot[dd - 1];
        for (int k = dd - 1; k > 1; --k) slot[k] = slot[k - 1];
        slot[1] = last;
      }
      __syncthreads();
    }
    conv = rotated == 0;
    __syncthreads();
  }
  __shared__ int zero;
  if (threadIdx.x == 0) zero = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < d; j += PT) {
    double nr = 0.0;
    for (int i = 0; i < d; ++i) nr = fma(A[i + j * PLD], A[i + j * PLD], nr);
    s[j] = sqrt(nr);
    if (!(s[j] > 0.0)) zero = 1;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < d * d; t += PT) {
    int i = t % d, j = t / d;
    if (s[j] > 0.0) A[i + j * PLD] /= s[j];
  }
  __syncthreads();
  return conv && zero == 0;
}

// λmin of the symmetric matrix M (destroyed): Householder tridiagonalisation, then Sturm counts
// by multisection over the Gershgorin interval (the reference tridiagonalises and bisects with
// Sturm sequences too, eig.jl:180-589).
__device__ double eigmin_lds(double* M, double* v, double* p, int d) {
  __shared__ double red[4];
  __shared__ double lohi[2];
  __shared__ int first;
  for (int k = 0; k + 2 < d; ++k) {
    const int m = d - k - 1;  // trailing order
    if (threadIdx.x == 0) {
      double nx = 0.0;
      for (int i = 0; i < m; ++i) nx = fma(M[k + 1 + i + k * PLD], M[k + 1 + i + k * PLD], nx);
      nx = sqrt(nx);
      double x0 = M[k + 1 + k * PLD];
      double alpha = x0 >= 0.0 ? -nx : nx;
      red[0] = alpha;
      red[1] = nx;
    }
    __syncthreads();
    const double alpha = red[0];
    if (red[1] == 0.0) {
      __syncthreads();
      continue;
    }
    for (int i = threadIdx.x; i < m; i += PT) v[i] = M[k + 1 + i + k * PLD] - (i == 0 ? alpha : 0.0);
    __syncthreads();
    if (threadIdx.x == 0) {
      double vv = 0.0;
      for (int i = 0; i < m; ++i) vv = fma(v[i], v[i], vv);
      red[2] = vv;
    }
    __syncthreads();
    const double vv = red[2];
    if (vv == 0.0) {
      __syncthreads();
      continue;
    }
    const double beta = 2.0 / vv;
    for (int i = threadIdx.x; i < m; i += PT) {
      double s = 0.0;
      for (int j = 0; j < m; ++j) s = fma(M[k + 1 + i + (k + 1 + j) * PLD], v[j], s);
      p[i] = beta * s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double vp = 0.0;
      for (int i = 0; i < m; ++i) vp = fma(v[i], p[i], vp);
      red[3] = 0.5 * beta * vp;
    }
    __syncthreads();
    const double K = red[3];
    for (int i = threadIdx.x; i < m; i += PT) p[i] -= K * v[i];  // w
    __syncthreads();
    for (int t = threadIdx.x; t < m * m; t += PT) {
      int i = t % m, j = t / m;
      M[k + 1 + i + (k + 1 + j) * PLD] -= v[i] * p[j] + p[i] * v[j];
    }
    for (int i = threadIdx.x; i < m; i += PT) {
      double x = i == 0 ? alpha : 0.0;
      M[k + 1 + i + k * PLD] = x;
      M[k + (k + 1 + i) * PLD] = x;
    }
    __syncthreads();
  }
  // Gershgorin interval
  if (threadIdx.x == 0) {
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < d; ++i) {
      double r = (i > 0 ? fabs(M[i + (i - 1) * PLD]) : 0.0) + (i + 1 < d ? fabs(M[i + 1 + i * PLD]) : 0.0);
      lo = fmin(lo, M[i + i * PLD] - r);
      hi = fmax(hi, M[i + i * PLD] + r);
    }
    lohi[0] = lo;
    lohi[1] = hi;
  }
  __syncthreads();
  for (int round = 0; round < 8; ++round) {
    const double lo = lohi[0], hi = lohi[1];
    if (threadIdx.x == 0) first = PT;
    __syncthreads();
    const double x = lo + (hi - lo) * (double)(threadIdx.x + 1) / (double)(PT + 1);
    int cnt = 0;
    double q = M[0] - x;
    if (q < 0.0) ++cnt;
    for (int i = 1; i < d; ++i) {
      if (q == 0.0) q = 1e-300;
      double e = M[i + (i - 1) * PLD];
      q = M[i + i * PLD] - x - e * e / q;
      if (q < 0.0) ++cnt;
    }
    if (cnt >= 1) atomicMin(&first, (int)threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int f = first;
      const double step = (hi - lo) / (double)(PT + 1);
      lohi[0] = lo + step * (double)f;
      lohi[1] = f < PT ? lo + step * (double)(f + 1) : hi;
    }
    __syncthreads();
  }
  return 0.5 * (lohi[0] + lohi[1]);
}

// ------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------

// sdpscaledynamic! (sdp.jl:212-262): LP, LD = chol(smat p), chol(smat d); U Σ = svd(LPᵀ LD);
// S = LP U, R = LP⁻ᵀ U, W = R Σ Rᵀ, D ← W ⊗ₛ W (utils.jl:475-513); ⟨p*, d*⟩ = Σ 1/sᵢ².
// out[2c] = flag, out[2c+1] = spsd.
__global__ void __launch_bounds__(PT) k_psd_scale(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                  double* out) {
  __shared__ double L0[PM], L1[PM], L2[PM], L3[PM];
  __shared__ double sv[PSD_MAX];
  __shared__ int slot[PSD_MAX + 1];
  __shared__ short ij[PSD_MAX * (PSD_MAX + 1) / 2][2];
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int n = C.n, d = psd_order(n);
  const double* p = cvec(I, NV_P) + C.col0;
  const double* dv = cvec(I, NV_D) + C.col0;
  smat_lds(L0, p, d);
  smat_lds(L1, dv, d);
  __syncthreads();
  bool ok = chol_lds(L0, L1, d);  // L0 = LP, L1 = LD
  if (ok) {
    for (int t = threadIdx.x; t < d * d; t += PT) {  // L2 = LPᵀ LD (both lower triangular)
      int i = t % d, j = t / d;
      double s = 0.0;
      for (int k = i > j ? i : j; k < d; ++k) s = fma(L0[k + i * PLD], L1[k + j * PLD], s);
      L2[i + j * PLD] = s;
    }
    __syncthreads();
    ok = svd_jacobi(L2, sv, d, slot);  // L2 = U
  }
  double spsd = 0.0;
  if (ok) {
    double* R = C.cache;
    double* S = C.cache + d * d;
    double* s = C.cache + 2 * d * d;
    for (int t = threadIdx.x; t < d * d; t += PT) {  // S = LP U
      int i = t % d, j = t / d;
      double a = 0.0;
      for (int k = 0; k <= i; ++k) a = fma(L0[i + k * PLD], L2[k + j * PLD], a);
      S[i + j * d] = a;
    }
    for (int j = threadIdx.x; j < d; j += PT) {  // L3 = R = LP⁻ᵀ U, column by column
      for (int i = d - 1; i >= 0; --i) {
        double a = L2[i + j * PLD];
        for (int k = i + 1; k < d; ++k) a -= L0[k + i * PLD] * L3[k + j * PLD];
        L3[i + j * PLD] = a / L0[i + i * PLD];
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < d * d; t += PT) {
      int i = t % d, j = t / d;
      R[i + j * d] = L3[i + j * PLD];
      double a = 0.0;  // W = R Σ Rᵀ into L1
      for (int k = 0; k < d; ++k) a = fma(L3[i + k * PLD] * sv[k], L3[j + k * PLD], a);
      L1[i + j * PLD] = a;
    }
    for (int j = threadIdx.x; j < d; j += PT) {
      s[j] = sv[j];
      for (int i = j; i < d; ++i) {
        int a = svpos(i, j, d);
        ij[a][0] = (short)i;
        ij[a][1] = (short)j;
      }
    }
    __syncthreads();
    // D[a, b] = c_a c_b (W_ik W_jl + W_il W_jk) / 2 for a = (i, j), b = (k, l)
    for (size_t t = threadIdx.x; t < (size_t)n * n; t += PT) {
      const int a = (int)(t % n), b = (int)(t / n);
      const int i = ij[a][0], j = ij[a][1], k = ij[b][0], l = ij[b][1];
      double h = 0.5 * (L1[i + k * PLD] * L1[j + l * PLD] + L1[i + l * PLD] * L1[j + k * PLD]);
      if (i != j) h *= R2;
      if (k != l) h *= R2;
      C.D[t] = h;
    }
    if (threadIdx.x == 0)
      for (int k = 0; k < d; ++k) spsd += 1.0 / (sv[k] * sv[k]);
  }
  if (threadIdx.x == 0) {
    out[2 * c] = ok ? 1.0 : 0.0;
    out[2 * c + 1] = spsd;
  }
}

// sdpcorr0dynamic! (sdp.jl:490-509): r = svec(R (σμ I − Σ²) Rᵀ)
__global__ void __launch_bounds__(PT) k_psd_corr0(const InstDev* insts, const ConeDesc* cones, const int* list, int kr) {
  __shared__ double L0[PM];
  __shared__ double g[PSD_MAX];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int n = C.n, d = psd_order(n);
  const double smu = I.sc[S_SMU];
  const double* R = C.cache;
  const double* s = C.cache + 2 * d * d;
  for (int t = threadIdx.x; t < d * d; t += PT) L0[(t % d) + (t / d) * PLD] = R[t];
  for (int k = threadIdx.x; k < d; k += PT) g[k] = smu - s[k] * s[k];
  __syncthreads();
  double* r = cvec(I, kr) + C.col0;
  for (int j = 0; j < d; ++j)
    for (int i = j + (int)threadIdx.x; i < d; i += PT) {
      double a = 0.0;
      for (int k = 0; k < d; ++k) a = fma(L0[i + k * PLD] * g[k], L0[j + k * PLD], a);
      r[svpos(i, j, d)] = i == j ? a : R2 * a;
    }
}

// sdpcorrdynamic! (sdp.jl:374-433): A = Rᵀ ΔP R, B = Sᵀ ΔD S, X = A B,
// Ŵ_ij = −(s_i X_ij + s_j X_ji)/(s_i + s_j) (i ≠ j), Ŵ_jj = σμ − s_j² − X_jj, r = svec(R Ŵ Rᵀ)
__global__ void __launch_bounds__(PT) k_psd_corr(const InstDev* insts, const ConeDesc* cones, const int* list, int kr,
                                                 int kdp, int kdd) {
  __shared__ double L0[PM], L1[PM], L2[PM], L3[PM], L4[PM];
  __shared__ double sv[PSD_MAX];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int n = C.n, d = psd_order(n);
  const double smu = I.sc[S_SMU];
  smat_lds(L0, cvec(I, kdp) + C.col0, d);
  smat_lds(L1, cvec(I, kdd) + C.col0, d);
  for (int t = threadIdx.x; t < d * d; t += PT) {
    L2[(t % d) + (t / d) * PLD] = C.cache[t];          // R
    L3[(t % d) + (t / d) * PLD] = C.cache[d * d + t];  // S
  }
  for (int k = threadIdx.x; k < d; k += PT) sv[k] = C.cache[2 * d * d + k];
  __syncthreads();
  gemm_lds<false, false>(L4, L0, L2, d);  // ΔP R
  __syncthreads();
  gemm_lds<true, false>(L0, L2, L4, d);   // A = Rᵀ ΔP R
  __syncthreads();
  gemm_lds<false, false>(L4, L1, L3, d);  // ΔD S
  __syncthreads();
  gemm_lds<true, false>(L1, L3, L4, d);   // B = Sᵀ ΔD S
  __syncthreads();
  gemm_lds<false, false>(L4, L0, L1, d);  // X = A B
  __syncthreads();
  for (int t = threadIdx.x; t < d * d; t += PT) {  // Ŵ into L3
    int i = t % d, j = t / d;
    double w;
    if (i == j) w = smu - sv[j] * sv[j] - L4[j + j * PLD];
    else w = -(sv[i] * L4[i + j * PLD] + sv[j] * L4[j + i * PLD]) / (sv[i] + sv[j]);
    L3[i + j * PLD] = w;
  }
  __syncthreads();
  gemm_lds<false, true>(L4, L3, L2, d);   // Ŵ Rᵀ
  __syncthreads();
  gemm_lds<false, false>(L0, L2, L4, d);  // R Ŵ Rᵀ
  __syncthreads();
  double* r = cvec(I, kr) + C.col0;
  for (int t = threadIdx.x; t < n; t += PT) {
    // svec position t → (i, j): walk the columns
    int j = 0, a = t;
    while (a >= d - j) {
      a -= d - j;
      ++j;
    }
    int i = j + a;
    double x = 0.5 * (L0[i + j * PLD] + L0[j + i * PLD]);
    r[t] = i == j ? x : R2 * x;
  }
}

// sdpmaxstepdynamic (sdp.jl:583-630) for the primal (F = R) and dual (F = S, M / s_i s_j)
// directions: M = Fᵀ smat(Δx) F; τ = 1 if I + M is positive definite, else 1/max(1, −λmin(M)).
// out[c] = step_frac · min(τp, τd).
__global__ void __launch_bounds__(PT) k_psd_maxstep(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                    int kdp, int kdd, double step_frac, double* out) {
  __shared__ double L0[PM], L1[PM], L2[PM], L4[PM];
  __shared__ double sv[PSD_MAX], v[PSD_MAX], pw[PSD_MAX];
  const int c = list[blockIdx.x];
  const ConeDesc C = cones[c];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int n = C.n, d = psd_order(n);
  double tau[2];
  for (int side = 0; side < 2; ++side) {
    smat_lds(L0, cvec(I, side == 0 ? kdp : kdd) + C.col0, d);
    for (int t = threadIdx.x; t < d * d; t += PT) L2[(t % d) + (t / d) * PLD] = C.cache[side * d * d + t];
    for (int k = threadIdx.x; k < d; k += PT) sv[k] = C.cache[2 * d * d + k];
    __syncthreads();
    gemm_lds<false, false>(L4, L0, L2, d);
    __syncthreads();
    gemm_lds<true, false>(L0, L2, L4, d);  // Fᵀ ΔX F
    __syncthreads();
    // M (symmetrised, scaled) into L4 and I + M into L1: L0 is only read here, so no thread
    // sees another's result in place of an input
    for (int t = threadIdx.x; t < d * d; t += PT) {
      int i = t % d, j = t / d;
      double m = 0.5 * (L0[i + j * PLD] + L0[j + i * PLD]);
      if (side == 1) m /= sv[i] * sv[j];
      L4[i + j * PLD] = m;
      L1[i + j * PLD] = m + (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (chol_lds(L1, nullptr, d)) {
      tau[side] = 1.0;
    } else {
      const double lam = eigmin_lds(L4, v, pw, d);
      tau[side] = 1.0 / fmax(1.0, -lam);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = step_frac * fmin(tau[0], tau[1]);
}

// sdpshadow! (sdp.jl:666-678): y = svec(X⁻¹) with X = smat(x) = L Lᵀ; X⁻¹ = L⁻ᵀ L⁻¹. y is left
// untouched when X is not positive definite.
__global__ void __launch_bounds__(PT) k_psd_dshadow(const InstDev* insts, const ConeDesc* cones, const int* list,
                                                    int kx, int ky) {
  __shared__ double L0[PM], L1[PM], L2[PM];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int n = C.n, d = psd_order(n);
  smat_lds(L0, cvec(I, kx) + C.col0, d);
  __syncthreads();
  if (!chol_lds(L0, nullptr, d)) return;
  for (int j = threadIdx.x; j < d; j += PT)  // L1 = L⁻¹ column by column (forward substitution)
    for (int i = 0; i < d; ++i) {
      double a = i == j ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) a -= L0[i + k * PLD] * L1[k + j * PLD];
      L1[i + j * PLD] = i < j ? 0.0 : a / L0[i + i * PLD];
    }
  __syncthreads();
  gemm_lds<true, false>(L2, L1, L1, d);  // L⁻ᵀ L⁻¹
  __syncthreads();
  double* y = cvec(I, ky) + C.col0;
  for (int j = 0; j < d; ++j)
    for (int i = j + (int)threadIdx.x; i < d; i += PT) {
      const double x = 0.5 * (L2[i + j * PLD] + L2[j + i * PLD]);
      y[svpos(i, j, d)] = i == j ? x : R2 * x;
    }
}

void launch_psd_dshadow(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, int kx, int ky) {
  if (n) k_psd_dshadow<<<n, PT, 0, st>>>(I, C, list, kx, ky);
}

// y += a D x on the dense scaling blocks (second-order and PSD cones), one workgroup per cone
// y += a·D x over the dense-block cones: grid (cone, 64-row slice); each wave sums a quarter of
// the columns (j ≡ wave mod 4) for the slice's rows (coalesced along the rows), and the four
// partial sums are added in a fixed order
__global__ void __launch_bounds__(PT) k_cone_matvec_dense(const InstDev* insts, const ConeDesc* cones,
                                                          const int* list, int kx, int ky, Coef a) {
  __shared__ double red[PT / 64][64];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  if (I.sc[S_ACTIVE] == 0.0) return;
  const int i0 = blockIdx.y * 64;
  if (i0 >= C.n) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = i0 + lane;
  const double* x = cvec(I, kx) + C.col0;
  double acc = 0.0;
  if (i < C.n)
    for (int j = w; j < C.n; j += PT / 64) acc = fma(C.D[i + (size_t)j * C.n], x[j], acc);
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && i < C.n) {
    const double av = a.slot < 0 ? a.c : (a.inv ? a.c / I.sc[a.slot] : a.c * I.sc[a.slot]);
    cvec(I, ky)[C.col0 + i] += av * ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
  }
}

// ‖Q_vv + D_v‖² of the dense-block cones: grid (cone, chunk) → part[DENSE_SPLIT·listed + chunk]
__global__ void __launch_bounds__(PT) k_cone_hnorm2_dense(const InstDev* insts, const ConeDesc* cones,
                                                          const int* list, double* part) {
  __shared__ double red[PT / 64];
  const ConeDesc C = cones[list[blockIdx.x]];
  const InstDev& I = insts[C.inst];
  double s = 0.0;
  if (I.sc[S_ACTIVE] != 0.0) {
    const size_t nn = (size_t)C.n * C.n;
    const size_t chunk = (nn + DENSE_SPLIT - 1) / DENSE_SPLIT;
    const size_t lo = chunk * blockIdx.y, hi = lo + chunk < nn ? lo + chunk : nn;
    for (size_t t = lo + threadIdx.x; t < hi; t += PT) {
      const double h = (C.Qd ? C.Qd[t] : 0.0) + C.D[t];
      s = fma(h, h, s);
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)blockIdx.x * DENSE_SPLIT + blockIdx.y] = red[0] + red[1] + red[2] + red[3];
}

// out[cone] = Σ of its DENSE_SPLIT partials, in order
__global__ void k_cone_dense_sum(const InstDev* insts, const ConeDesc* cones, const int* list, const double* part,
                                 double* out) {
  const int c = list[blockIdx.x];
  const InstDev& I = insts[cones[c].inst];
  if (threadIdx.x != 0 || I.sc[S_ACTIVE] == 0.0) return;
  double s = 0.0;
  for (int k = 0; k < DENSE_SPLIT; ++k) s += part[(size_t)blockIdx.x * DENSE_SPLIT + k];
  out[c] = s;
}

void launch_psd_scale(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, double* out) {
  if (n) k_psd_scale<<<n, PT, 0, st>>>(I, C, list, out);
}
void launch_psd_corr0(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, int kr) {
  if (n) k_psd_corr0<<<n, PT, 0, st>>>(I, C, list, kr);
}
void launch_psd_corr(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, int kr, int kdp,
                     int kdd) {
  if (n) k_psd_corr<<<n, PT, 0, st>>>(I, C, list, kr, kdp, kdd);
}
void launch_psd_maxstep(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, int kdp, int kdd,
                        double step_frac, double* out) {
  if (n) k_psd_maxstep<<<n, PT, 0, st>>>(I, C, list, kdp, kdd, step_frac, out);
}
void launch_cone_matvec_dense(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n, int nmax,
                              int kx, int ky, Coef a) {
  if (n) k_cone_matvec_dense<<<dim3(n, (nmax + 63) / 64), PT, 0, st>>>(I, C, list, kx, ky, a);
}
void launch_cone_hnorm2_dense(hipStream_t st, const InstDev* I, const ConeDesc* C, const int* list, int n,
                              double* out, double* part) {
  if (!n) return;
  k_cone_hnorm2_dense<<<dim3(n, DENSE_SPLIT), PT, 0, st>>>(I, C, list, part);
  k_cone_dense_sum<<<n, 64, 0, st>>>(I, C, list, part, out);
}

}  // namespace mb
