// Persistent tile-DAG factorization and pipelined triangular solves for large fronts.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace mb {

constexpr int BIG_LD = 512;  // fronts with ld ≥ BIG_LD take the persistent kernels

struct BigFront {
  double* F;            // frontal storage (ld × ld, column-major)
  int ld, r, s;
  int T, Tr;            // tiles (res tiles first, then separator tiles from column r)
  const int* tb;        // T+1 tile boundaries
  int* state;           // T(T+1)/2 tile flags: the factorization epoch once the tile is final
  int* ticket;          // task ticket
  int* abortf;          // abort flag (an int in the instance's S_ABORT scalar)
  const int4* tasks;    // (type, i, j, kmax) in column order: 0 diagonal, 1 factor tile,
                        // 2 update-matrix tile, 3 sub-diagonal tile (staged for the diagonal)
  int ntask;
  int nstate;
  double* Dinv;         // Tr inverted diagonal blocks, 64×64 column-major each
  float* F32;           // mixed precision: the FP32 factor (ld × r, same indexing as F), or null
  float* Dinv32;        // mixed precision: FP32 inverted diagonal blocks, or null
  int* stage;           // Tr flags: sub-diagonal tile (j+1, j) accumulated (epoch-valued)
  unsigned long long* trace;  // optional: 8 timestamps per task (100 MHz wall clock), or null
  unsigned long long* sll;  // solve hand-off: T blocks × 64 entries × 2 (value half, epoch) words
  double* spart;        // backward solve: L[sep, b]ᵀ x_sep by 64-row separator chunk, ⌈s/64⌉·Tr·64 values
  double* Mf;           // Tr FP64 64×64 blocks M_b = L_bb⁻¹ L_{b,b−1} (forward-solve chain step)
  double* Nt;           // Tr FP64 64×64 blocks N_bᵀ, N_b = L_{b+1,b} L_bb⁻¹ (backward-solve chain step)
  long long* sticket;   // 64-bit monotone ticket of the solve kernels
  int inst;
  // Write-once factor storage (both set, or both null): L holds the FP64 factor columns (ld × r,
  // same indexing as F) and stg the staged sub-diagonal tiles (Tr 64×64 tiles of the factor's
  // element type), so the assembled front stays intact. With null, the FP64 factor overwrites F
  // in place. Either way, tiles finalized in the running launch are read device-coherent.
  double* L;
  double* stg;
};

// factorization workgroups resident per CU: launch grids are given in CUs and multiplied by this
constexpr int DAG_PER_CU = 2;

// fp32: the factor in FP32 on FP32 MFMA (B.F32 / B.Dinv32 must be allocated); the solves read
// the factor of the same precision and compute in FP64
void launch_dag_factor(hipStream_t st, const BigFront& B, const InstDev* I, int grid, int epoch, bool fp32);
// after each launch_dag_factor: the solves' chain-step tiles, and the factorization ticket back to 0
void launch_big_mtiles(hipStream_t st, const BigFront& B, const InstDev* I, bool fp32);
// the solve launchers return their grid size: every workgroup consumes one ticket of B.sticket,
// so the caller advances its ticket base by that many
int launch_big_fsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const int* childlist,
                      const InstDev* I, int kv, int epoch, long long base, bool fp32);
int launch_big_bsolve(hipStream_t st, const BigFront& B, const FrontDesc* F, int f, const InstDev* I, int kv,
                      int epoch, long long base, bool fp32);

// The large fronts of a gang (several instances) in one launch each (write-once factor storage
// required). Bs: the nf fronts (device copies of their BigFront), bfront: their indices in the
// front table F; order: (front, task or block) pairs, each front's own items in their order
// (max2tr: 2·max Tr, the chain-step tile grid); the solves take one ticket of the 64-bit counter
// per workgroup (the caller advances base by ntot).
//
// The task queues of one launch: queue q is order[off[q], off[q+1]), served by the workgroups
// ≡ q (mod n); each front's tasks lie in one queue, in their topological order
constexpr int DAG_QUEUES = 8;
struct DagQueues {
  int n;
  int off[DAG_QUEUES + 1];
};
// ticket: DAG_QUEUES counters, 0 on entry and put back to 0; before / after (or null): events
// recorded around the factorization kernel alone
void launch_dag_factor_multi(hipStream_t st, const BigFront* Bs, int nf, const int2* order, const DagQueues& Q,
                             int* ticket, const InstDev* I, int grid, int epoch, bool fp32, int max2tr,
                             hipEvent_t before = nullptr, hipEvent_t after = nullptr);
void launch_big_fsolve_multi(hipStream_t st, const BigFront* Bs, const int* bfront, int nf, const int2* order,
                             int ntot, unsigned long long* ticket, const FrontDesc* F, const int* childlist,
                             const InstDev* I, int kv, int epoch, long long base, bool fp32);
// maxtr, maxch: the separator-product grid (largest Tr and number of 64-row separator chunks)
void launch_big_bsolve_multi(hipStream_t st, const BigFront* Bs, const int* bfront, int nf, const int2* order,
                             int ntot, unsigned long long* ticket, const FrontDesc* F, const InstDev* I, int kv,
                             int epoch, long long base, bool fp32, int maxtr, int maxch);

}  // namespace mb
