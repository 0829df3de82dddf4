// Dense frontal Cholesky on caller-owned device memory: the frontal POTRF of the supernodal
// factorization (the dense kernel inside CliqueTrees' cholesky!, called at
// src/IPM/src/kkt/uzawa.jl:93 and :111), exposed on its own so that a caller holding one dense
// front in HBM can factor it with the persistent tile kernel (k_dag_factor).
//
// Workspace of one factorization of order n (T = ⌈n/64⌉ tiles): the instance table and its
// scalars, the task ticket, T+1 tile bounds, T(T+1)/2 tasks and tile flags, T stage flags and
// T inverted 64×64 diagonal blocks. k_dense_prep writes all of it on the device (task list,
// bounds, cleared flags, scalars), so a call needs no host-to-device copy; the factorization then
// runs with epoch 1.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/mumblebee_hip.h"
#include "device.h"
#include "kernels.h"
#include "kernels_big.h"
#include "solver.h"

namespace mb {

namespace {

#define HIPCHK_F(x)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +         \
                                                   hipGetErrorString(e_) + " at " #x);   \
  } while (0)

constexpr size_t ALIGN = 256;
constexpr size_t up(size_t b) { return (b + ALIGN - 1) / ALIGN * ALIGN; }

// byte offsets of the workspace parts for T tiles
struct WsLayout {
  size_t inst = 0, sc = 0, ticket = 0, tb = 0, tasks = 0, state = 0, stage = 0, dinv = 0, total = 0;
  constexpr explicit WsLayout(int T) {
    const size_t nt = (size_t)T * (T + 1) / 2;
    size_t o = 0;
    inst = o; o += up(sizeof(InstDev));
    sc = o; o += up(sizeof(double) * NSCALAR);
    ticket = o; o += up(sizeof(int) * 2);
    tb = o; o += up(sizeof(int) * (T + 1));
    tasks = o; o += up(sizeof(int4) * (nt > 0 ? nt : 1));
    state = o; o += up(sizeof(int) * (nt > 0 ? nt : 1));
    stage = o; o += up(sizeof(int) * (T > 0 ? T : 1));
    dinv = o; o += up(sizeof(double) * NB * NB * (T > 0 ? T : 1));
    total = o;
  }
};

// One workgroup: the instance table, scalars, ticket, tile bounds, the task list (the same
// topological order as the solver's large fronts with no separator: diagonal j, then the factor
// tiles of column j−1 below row j, then the staged sub-diagonal tile (j+1, j)), and cleared flags.
__global__ void k_dense_prep(char* ws, WsLayout L, int n, int T) {
  InstDev* I = reinterpret_cast<InstDev*>(ws + L.inst);
  double* sc = reinterpret_cast<double*>(ws + L.sc);
  int* ticket = reinterpret_cast<int*>(ws + L.ticket);
  int* tb = reinterpret_cast<int*>(ws + L.tb);
  int4* tasks = reinterpret_cast<int4*>(ws + L.tasks);
  int* state = reinterpret_cast<int*>(ws + L.state);
  int* stage = reinterpret_cast<int*>(ws + L.stage);
  const int nt = T * (T + 1) / 2;
  if (threadIdx.x == 0) {
    InstDev z{};
    z.sc = sc;
    *I = z;
    ticket[0] = 0;
    ticket[1] = 0;
  }
  for (int k = threadIdx.x; k < NSCALAR; k += blockDim.x) sc[k] = k == S_ACTIVE ? 1.0 : 0.0;
  for (int k = threadIdx.x; k <= T; k += blockDim.x) tb[k] = k < T ? k * NB : n;
  for (int k = threadIdx.x; k < nt; k += blockDim.x) state[k] = 0;
  for (int k = threadIdx.x; k < T; k += blockDim.x) stage[k] = 0;
  // column j's tasks start after those of the columns before it: 1 + (T−1−j' if j' > 0) + (1 if
  // j'+1 < T) each
  for (int j = threadIdx.x; j < T; j += blockDim.x) {
    int o = 0;
    for (int q = 0; q < j; ++q) o += 1 + (q > 0 ? T - 1 - q : 0) + (q + 1 < T ? 1 : 0);
    tasks[o++] = make_int4(0, j, j, j > 0 ? j - 1 : 0);
    if (j > 0)
      for (int i = j + 1; i < T; ++i) tasks[o++] = make_int4(1, i, j - 1, j - 1);
    if (j + 1 < T) tasks[o++] = make_int4(3, j + 1, j, j);
  }
}

// factor A (device, n×n, column-major, leading dimension lda) in place with the workspace at ws;
// returns 0, or 1 if a pivot was not positive (the matrix is not positive definite)
int dense_factor(double* A, int64_t n, int64_t lda, int grid, char* ws, hipStream_t st) {
  const int T = (int)((n + NB - 1) / NB);
  const WsLayout L(T);
  hipLaunchKernelGGL(k_dense_prep, dim3(1), dim3(256), 0, st, ws, L, (int)n, T);
  HIPCHK_F(hipGetLastError());
  if (T == 0) return 0;
  BigFront B{};
  B.F = A;
  B.ld = (int)lda;
  B.r = (int)n;
  B.s = 0;
  B.T = T;
  B.Tr = T;
  B.tb = reinterpret_cast<const int*>(ws + L.tb);
  B.state = reinterpret_cast<int*>(ws + L.state);
  B.nstate = T * (T + 1) / 2;
  B.ticket = reinterpret_cast<int*>(ws + L.ticket);
  B.abortf = reinterpret_cast<int*>(ws + L.sc + sizeof(double) * S_ABORT);
  B.tasks = reinterpret_cast<const int4*>(ws + L.tasks);
  B.ntask = T * (T + 1) / 2;
  B.Dinv = reinterpret_cast<double*>(ws + L.dinv);
  B.stage = reinterpret_cast<int*>(ws + L.stage);
  B.inst = 0;
  if (grid <= 0) {
    int dev = 0, ncu = 256;
    HIPCHK_F(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
    grid = ncu;
  }
  grid = std::min(grid, B.ntask);
  launch_dag_factor(st, B, reinterpret_cast<const InstDev*>(ws + L.inst), grid, 1, false);
  HIPCHK_F(hipGetLastError());
  double flags[2];
  HIPCHK_F(hipMemcpyAsync(&flags[0], ws + L.sc + sizeof(double) * S_FAIL, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK_F(hipMemcpyAsync(&flags[1], ws + L.sc + sizeof(double) * S_ABORT, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK_F(hipStreamSynchronize(st));
  int32_t raised = 0;
  std::memcpy(&raised, &flags[1], sizeof(raised));
  if (raised != 0) throw std::runtime_error("persistent front kernel exceeded its wait limit");
  return flags[0] != 0.0 ? 1 : 0;
}

void check_args(const double* A, int64_t n, int64_t lda) {
  if (n < 0 || lda < (n > 0 ? n : 1)) throw std::invalid_argument("dense cholesky: need n >= 0 and lda >= max(n, 1)");
  if (n > 0 && A == nullptr) throw std::invalid_argument("dense cholesky: null matrix");
  if (lda * (n > 0 ? n - 1 : 0) + n > (int64_t)1 << 31)
    throw std::invalid_argument("dense cholesky: matrix extent beyond 2^31 elements");
}

template <class F>
int guard_front(F&& fn) {
  try {
    return fn();
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -1;
  }
}

// The no-allocation form: one workspace per GPU in a file-scope device array, sized for fronts
// of order up to NOALLOC_NMAX (4.4 MB; held by each GPU the library runs on until the process
// ends). In host code `scratch` only names the symbol: its device address on the current GPU comes
// from hipGetSymbolAddress, taken once per device at the entry point's first call there.
constexpr int64_t NOALLOC_NMAX = 8192;
constexpr size_t NOALLOC_BYTES = WsLayout((int)((NOALLOC_NMAX + NB - 1) / NB)).total;
__device__ __align__(256) char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[MAX_DEVICES] = {};

char* scratch_on_current_device() {
  int dev = 0;
  HIPCHK_F(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("dense cholesky: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK_F(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("dense cholesky: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace mb

extern "C" {

int mb_dense_cholesky_noalloc(double* A, int64_t n, int64_t lda, int32_t grid, void* stream) {
  return mb::guard_front([&]() {
    if (n > mb::NOALLOC_NMAX)
      throw std::invalid_argument("dense cholesky (no allocation): order " + std::to_string(n) +
                                  " exceeds the workspace limit " + std::to_string(mb::NOALLOC_NMAX));
    mb::check_args(A, n, lda);
    return mb::dense_factor(A, n, lda, grid, mb::scratch_on_current_device(), (hipStream_t)stream);
  });
}

int mb_dense_cholesky(double* A, int64_t n, int64_t lda, int32_t grid, void* stream) {
  return mb::guard_front([&]() {
    mb::check_args(A, n, lda);
    const mb::WsLayout L((int)((n + mb::NB - 1) / mb::NB));
    mb::DevBuf ws;
    ws.alloc(L.total);
    return mb::dense_factor(A, n, lda, grid, ws.as<char>(), (hipStream_t)stream);
  });
}

}  // extern "C"
