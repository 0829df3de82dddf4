// Launch wrappers of the HIP kernels and the per-instance device table they read.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device.h"
#include "kinds.h"

namespace mb {

// n-vector kinds (one arena of NV_COUNT × n doubles per instance)
enum NVKind : int {
  NV_P = 0, NV_D, NV_F, NV_DF, NV_FR, NV_QP, NV_DPA, NV_DDA, NV_DP, NV_DD, NV_PSCL,
  NV_KF, NV_KX, NV_CW, NV_T0, NV_T1, NV_T2, NV_COUNT
};
// m-vector kinds
enum MVKind : int {
  MV_Y = 0, MV_G, MV_DG, MV_DYA, MV_DY, MV_YSCL, MV_KG, MV_KY, MV_CP, MV_T0, MV_T1, MV_COUNT
};

struct InstDev {
  double* nv;       // NV_COUNT × n
  double* mv;       // MV_COUNT × m
  double* sc;       // NSCALAR
  double* scratch;  // 3n doubles for cone correctors
  double* F;        // frontal storage base
  CSRDev B, Bt, Q;  // B (m rows), Bᵀ (n rows), Q (n rows)
  int n, m;
};

struct Tile {
  int inst;
  int start;
};
constexpr int TILE_THREADS = 256;  // threads of a tile-list workgroup
// A sparse-product tile list holds tiles of 256/lanes consecutive rows (start ≥ 0) and, after them,
// one tile per row longer than SPMV_LONG × lanes (start = −1 − row), which a whole workgroup sums;
// the row tiles leave such rows to it.
constexpr int SPMV_LONG = 64;

// coefficient value: c (slot < 0), c·sc[slot], or c/sc[slot] (inv)
struct Coef {
  double c;
  int slot;
  int inv;
};
inline Coef K(double c) { return Coef{c, -1, 0}; }
inline Coef KS(double c, int slot) { return Coef{c, slot, 0}; }
inline Coef KI(double c, int slot) { return Coef{c, slot, 1}; }

void launch_axpby(hipStream_t, const InstDev*, const Tile*, int nt, int space, int kx, int ky, int kz, Coef a,
                  Coef b);
void launch_axpbypcz(hipStream_t, const InstDev*, const Tile*, int nt, int space, int kx, int ky, int kw, int kz,
                     Coef a, Coef b, Coef c);
// counter: ninst ints, zero between launches (the kernel leaves them at zero)
void launch_dot(hipStream_t, const InstDev*, const Tile*, int nt, const int* tile_ofs, int ninst, int space,
                int kx, int ky, int mode, int kw0, int kw1, double* partial, int* counter, int slot, int op);
// A batch of dot products of one instance in one launch, one workgroup per dot over the whole
// vector (for vectors of at most DOTS_LEN_MAX entries): sc[slot] = op(Σ term), the term as in
// launch_dot's mode, op 1 = square root.
struct DotSpec {
  int space, kx, ky, mode, kw0, kw1, slot, op;
};
constexpr int DOTS_MAX = 16;
constexpr int DOTS_LEN_MAX = 1 << 14;
struct DotBatch {
  DotSpec s[DOTS_MAX];
  int n = 0;
  DotBatch& add(int space, int kx, int ky, int mode, int slot, int op, int kw0 = 0, int kw1 = 0) {
    s[n++] = DotSpec{space, kx, ky, mode, kw0, kw1, slot, op};
    return *this;
  }
};
// ninst > 1: the same batch for instances inst .. inst+ninst−1 of the table in one launch
void launch_dots(hipStream_t, const InstDev*, int inst, const DotBatch&, int ninst = 1);
// z = a x + b y + c w; up to LIN_MAX of them on the n-space and on the m-space in one launch,
// applied in order at each index (a later one sees the earlier ones' results)
struct Lin {
  int kx, ky, kw, kz;
  Coef a, b, c;
};
constexpr int LIN_MAX = 3;
struct LinSet {
  Lin op[LIN_MAX];
  int n = 0;
  LinSet& add(const Lin& l) {
    op[n++] = l;
    return *this;
  }
};
void launch_lin(hipStream_t, const InstDev*, const Tile* ntiles, int nnt, const Tile* mtiles, int nmt,
                const LinSet& N, const LinSet& M);
void launch_spmv(hipStream_t, const InstDev*, const Tile*, int nt, int which, int kx, int ky, Coef a, Coef b);
// Sparse products over the n rows and over the m rows in one launch. n rows (the row tiles of
// matrix ntm: 1 = Bᵀ, 2 = Q): y = aq Q xq + at Bᵀ xt + c z + d w + e v, and Q xq alone into kyq;
// m rows (the row tiles of B): y = a B x + c z, then y2 += e y (y2 = e y if y2set). A term or
// output whose kind is < 0 is absent.
struct SpmvN {
  int kxq, kxt, kz, kw, kv, ky;
  Coef aq, at, c, d, e;
  int kyq = -1;
};
struct SpmvM {
  int kx, kz, ky, ky2;
  Coef a, c, e;
  int y2set;
};
void launch_spmv_nm(hipStream_t, const InstDev*, int ntm, const Tile* ntiles, int nn, const Tile* mtiles, int nm,
                    const SpmvN& N, const SpmvM& M);
void launch_cone_matvec(hipStream_t, const InstDev*, const ConeDesc*, int nc, int kx, int ky, Coef a);
// psd / dense: lists of the PSD cones and of the cones with dense scaling blocks (SOC, PSD)
// PSD cones of order above PSD_MAX (psd_big.hip): the list, its length and the largest order
struct BigPsd {
  const int* list = nullptr;
  int n = 0;
  int dmax = 0;
};
// soc: the second-order cones (one wavefront each)
void launch_cone_scale(hipStream_t, const InstDev*, const ConeDesc*, int nc, const int* cone_ofs, int ninst,
                       double* out, const int* psd, int npsd, const BigPsd& big, const int* soc, int nsoc);
// dense_nmax: the largest order among the dense-block cones; dpart: DENSE_SPLIT doubles per dense cone
void launch_cone_hnorm2(hipStream_t, const InstDev*, const ConeDesc*, int nc, const int* cone_ofs, int ninst,
                        double* out, int slot, const int* dense, int ndense, int dense_nmax, double* dpart,
                        const BigPsd& big, double* bpart);
void launch_cone_corr0(hipStream_t, const InstDev*, const ConeDesc*, int nc, int kr);
void launch_cone_corr(hipStream_t, const InstDev*, const ConeDesc*, int nc, int kr, int kdp, int kdd);
void launch_cone_maxstep(hipStream_t, const InstDev*, const ConeDesc*, int nc, const int* cone_ofs, int ninst,
                         int kdp, int kdd, double step_frac, double* out, const int* psd, int npsd,
                         const BigPsd& big);

// PSD cones (order d ≤ PSD_MAX) and dense-block cones: one workgroup per listed cone (psd.hip)
constexpr int PSD_MAX = 48;
void launch_psd_scale(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, double* out);
void launch_psd_corr0(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, int kr);
void launch_psd_corr(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, int kr, int kdp, int kdd);
void launch_psd_maxstep(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, int kdp, int kdd,
                        double step_frac, double* out);
// PSD cones above PSD_MAX (psd_big.hip), order ≤ PSD_BIG_MAX. Per cone: cache R, S (d² each),
// s (d), W (d²) — psd_big_cache(d) doubles; workspace psd_big_ws(d) doubles. Their scaling block
// W ⊗ₛ W is not stored (ConeDesc::D is null): the products, the KKT assembly and ‖H‖ form it
// from W. Matrix slots of the launches: workspace X0..X2, cache R, S, W.
constexpr int PSD_BIG_MAX = 4096;
enum BigPsdSlot : int { BP_X0 = 0, BP_X1, BP_X2, BP_R, BP_S, BP_W };
inline long long psd_big_cache(long long d) { return 3 * d * d + d; }
inline long long psd_big_ws(long long d) { return 3 * d * d + ((d + NB - 1) / NB) * NB * NB + 16; }
void launch_bpsd_scale(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, double* out);
void launch_bpsd_corr0(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, int kr);
void launch_bpsd_corr(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, int kr, int kdp, int kdd);
void launch_bpsd_maxstep(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, int kdp, int kdd,
                         double step_frac, double* out);
// y += a·svec(W smat(x) W) over the listed cones
void launch_bpsd_matvec(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, int kx, int ky, Coef a);
// out[c] = ‖Q_vv + W ⊗ₛ W‖²; part: 2·DENSE_SPLIT doubles per listed cone
void launch_bpsd_hnorm2(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, double* out, double* part);
void launch_bpsd_assemble(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, const int* cone_front,
                          const FrontDesc*);
// the n×n block W ⊗ₛ W of the cone at one[0] into D (column-major)
void launch_bpsd_block(hipStream_t, const InstDev*, const ConeDesc*, const int* one, int n, double* D);
// the dense-block cones split over many workgroups: 64-row slices for D·x, DENSE_SPLIT chunks per
// cone for ‖Q_vv + D_v‖² (partials in part, summed in a fixed order); nmax: the largest order
constexpr int DENSE_SPLIT = 32;
void launch_cone_matvec_dense(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, int nmax, int kx,
                              int ky, Coef a);
void launch_cone_hnorm2_dense(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, double* out,
                              double* part);
void launch_zerofree(hipStream_t, const InstDev*, const ConeDesc*, int nc, int kx);
// dualshadow! y = −f′(x) (cone.jl:556-566) of every cone but the PSD ones (launch_psd_dshadow,
// launch_bpsd_dshadow)
void launch_cone_dshadow(hipStream_t, const InstDev*, const ConeDesc*, int nc, int kx, int ky);
void launch_psd_dshadow(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n, int kx, int ky);
void launch_bpsd_dshadow(hipStream_t, const InstDev*, const ConeDesc*, const BigPsd&, int kx, int ky);
// the shadow-solve seed of every listed exponential / power cone (initcache!)
void launch_td_initcache(hipStream_t, const InstDev*, const ConeDesc*, const int* list, int n);
void launch_fill(hipStream_t, double* x, size_t n, double v);
// One piece of a staged setup: bytes at dst (4-byte aligned) from src bytes into the staged image,
// or zeros (src < 0); the piece is copied by workgroups blk0 .. blk0 + ⌈bytes / STAGE_BLOCK⌉ − 1.
struct StageSeg {
  char* dst;
  long long src;
  long long bytes;
  long long blk0;
};
constexpr long long STAGE_BLOCK = 64 << 10;
// every piece of segs[0, nseg) in one launch of nblk workgroups
void launch_stage_scatter(hipStream_t, const char* image, const StageSeg* segs, int nseg, long long nblk);
// F[pos] = lval + δ qval (withq) or lval
void launch_assemble_union(hipStream_t, const InstDev*, int inst, const int64_t* pos, const double* lval,
                           const double* qval, int64_t nent, int withq);
// F[pos] += δ val
void launch_scatter_add(hipStream_t, const InstDev*, int inst, const int64_t* pos, const double* val, int64_t nent);
// nmax: the largest cone order (sets how many workgroups share one cone's columns)
void launch_assemble_cones(hipStream_t, const InstDev*, const ConeDesc*, const int* cone_front, const FrontDesc*,
                           int nc, int nmax);
void launch_add_diag(hipStream_t, const FrontDesc*, int nf, const InstDev*);
void launch_front_norm2(hipStream_t, const FrontDesc*, int nf, double* partial);
void launch_extend_add(hipStream_t, const FrontDesc*, const int* items, const int* item_tile0, int nitem,
                       int ntile, const InstDev*);
// flags: one int per step item (epoch-valued, zeroed once); abortf: the instance's abort word
// (nullptr: each front's own instance's)
void launch_panel(hipStream_t, const FrontDesc*, const StepDesc*, int nstep, int nblk, const InstDev*, int* flags,
                  int epoch, int* abortf);
void launch_update(hipStream_t, const FrontDesc*, const StepDesc*, int nstep, int nblk, const InstDev*);
void launch_fsolve(hipStream_t, const FrontDesc*, const int* items, int nitem, const int* childlist,
                   const InstDev*, int kv);
void launch_bsolve(hipStream_t, const FrontDesc*, const int* items, int nitem, const InstDev*, int kv);
void launch_cg_scalars(hipStream_t, const InstDev*, int ninst, int op);
void launch_probe_mfma(hipStream_t, double* out, bool fp32);
void launch_set_scalar(hipStream_t, double* sc, int slot, double v);
void launch_set_scalar2(hipStream_t, double* sc, int s1, double v1, int s2, double v2);
// the corrector's σμ (solver/ipm.jl:181-200) from μa = sc[S_TMP2]/ν, into sc[S_SMU]
void launch_smu(hipStream_t, double* sc, double nu, double mu, double mut);

// gang kernels over a combined instance table of k ≤ GANG_SLOTS slots (csrc/gang.cpp)
constexpr int GANG_SLOTS = 32;
struct GangVals {
  double a[GANG_SLOTS], b[GANG_SLOTS], c[GANG_SLOTS];
};
void launch_gang_set2(hipStream_t, const InstDev*, int k, int s1, int s2, const GangVals&);
void launch_gang_smu(hipStream_t, const InstDev*, int k, const GangVals&);
void launch_gang_active(hipStream_t, const InstDev*, int k, unsigned long long mask, unsigned long long occupied);
void launch_gang_setinst(hipStream_t, InstDev*, int slot, const InstDev&);
void launch_gang_gather(hipStream_t, const InstDev*, int k, unsigned long long mask, double* out);
// one member's assembly arrays: its front storage (nF doubles) and the BᵀB + δQ scatter map; fb0 /
// ub0: its first block of the fill and of the scatter launch (TILE threads a block, two doubles a
// thread for the fill)
struct GangAsm {
  double* F;
  long long nF;
  const int64_t* pos;
  const double* l;
  const double* q;
  long long nent;
  long long fb0, ub0;
};
// the fill and the scatter of all slots of the table A (k entries) in one launch each
void launch_gang_assemble(hipStream_t, const InstDev*, const GangAsm* A, int k, long long fblocks,
                          long long ublocks);

}  // namespace mb
