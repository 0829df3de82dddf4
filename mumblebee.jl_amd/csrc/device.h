// Descriptors shared by the host driver and the HIP kernels. Every batched kernel walks a
// list of work items (fronts, cones, vector segments); each item carries the instance index
// used to read that instance's device scalars, so one launch serves any number of
// independent instances (a batch is a block-diagonal problem with per-instance scalars).
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace mb {

constexpr int NB = 64;  // frontal tile / panel width

// Per-instance device scalars (double slots). Host-set slots are written by the driver before
// the kernels that read them; device-set slots are produced by reductions.
enum Scalar : int {
  S_DELTA = 0,   // augmentation δ
  S_ALPHA,       // 1/δ
  S_SMU,         // σμ for the correctors
  S_RHO,         // diagonal shift of the retried factorization
  S_NORM0,       // reduction outputs
  S_NORM1,
  S_NORM2,
  S_NORM3,
  S_NORM4,
  S_NORM5,
  S_CG_NU2,      // CG: ‖u‖²
  S_CG_EP2,      // CG: ‖w‖²
  S_CG_ALPHA,
  S_CG_BETA,
  S_CG_NP2,
  S_SPSD,        // Σ ⟨p*, d*⟩
  S_TAU,         // min max-step (the corrector's is the iterate update's step length)
  S_FAIL,        // factorization failure flag (> 0 failed)
  S_FLAGOK,      // scaling flag (1 ok, 0 failed)
  S_ACTIVE,      // 1 if this instance takes part in the current launch sequence
  S_CG_ACTIVE,   // 1 while this instance's CG loop continues
  S_IR_ACTIVE,
  S_TMP0,
  S_TMP1,
  S_TMP2,
  S_TMP3,
  S_R0,          // S_R0 .. S_R15: reduction outputs read back in one transfer
  S_R1, S_R2, S_R3, S_R4, S_R5, S_R6, S_R7,
  S_R8, S_R9, S_R10, S_R11, S_R12, S_R13, S_R14, S_R15,
  S_ABORT,       // the persistent kernels' abort flag: an int in this slot's low word (nonzero: raised)
  NSCALAR
};

// A frontal matrix: dense (r+s)×(r+s) column-major, leading dimension ld = r+s; the first r
// columns become the factor columns [L11; L21], the trailing s×s block the update matrix.
struct FrontDesc {
  double* F;       // frontal storage (FP64 path)
  float* F32;      // frontal storage (FP32 factor path), or nullptr
  int r, s, ld;
  int inst;
  int res0;        // first factor column (instance-local)
  int parent;      // index of the parent front in the instance's front table (-1 root)
  int nchild;
  int child0;      // first entry in the child list of the instance (global child table)
  const int* sepcols;  // s instance-local column indices
  const int* relidx;   // s positions in the parent's frontal index space
  double* uvec;    // s doubles: forward-solve update vector
};

// One (front, panel step) of the blocked factorization.
struct StepDesc {
  int front;       // index into the batch front table
  int kb, w;       // panel columns [kb, kb+w)
  int ntile;       // row tiles below kb: ceil((ld - kb)/NB)
  int tile0;       // prefix offset of this item's tiles within the launch
};

struct ConeDesc {
  int kind;        // ConeKind
  int inst;
  int col0;        // instance-local first column
  int n;           // columns
  double param;    // power-cone alpha
  double* D;       // n×n scaling block (column-major); n entries (its diagonal) for POS
  double* cache;   // per-cone cache slots
  const double* Qd;  // the n×n diagonal block of Q (for ‖H‖)
  double* ws;      // PSD cones above the LDS kernels' order: the d×d workspace (else null)
};

struct SegDesc {    // a contiguous range of one instance's vector
  int inst;
  int n;
  int tile0;        // prefix offset of tiles
};

// Sparse matrix (CSR) on device: y_i = Σ val[k] x[col[k]]. `lanes` (a power of two ≤ 64)
// lanes work on one row, so a 256-thread workgroup covers 256/lanes consecutive rows.
struct CSRDev {
  const int* ptr;
  const int* col;
  const double* val;
  int nrow;
  int lanes;
};

}  // namespace mb
