// The IPM driver over device-resident state: restates src/IPM/src/ipm/solver/ipm.jl (step!,
// predictor, corrector, termination) and src/IPM/src/kkt/uzawa.jl / cg.jl (Uzawa refinement,
// dual CG) with every vector operation, cone kernel, factorization and triangular solve on the
// GPU; only scalars cross to the host.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "device.h"
#include "host.h"
#include "kernels.h"
#include "kernels_big.h"

namespace mb {

struct Settings {
  int verbose = 0;
  double step_frac = 0.99;
  double feas_tol = 1e-8;
  double gap_tol = 1e-8;
  int max_iter = 100;
  double near_factor = 1000.0;
  double stall_tol = 1e-3;
  double refine_stall_tol = 0.5;
  int scale_max_iter = 10;
  int refine_max_iter = 10;
  int newton_max_iter = 100;
  double aug_tol = 1e-7;
  double infeas_abs = 1e-8;
  double infeas_rel = 1e-8;
  int amalgamate = 1;
  int big_front_ld = BIG_LD;  // fronts with ld ≥ this take the persistent kernels
  int dag_grid = 0;           // workgroups of the persistent factorization (0: one per CU)
  int precision = 0;          // 1: large fronts factored in FP32, refined in FP64
  double mixed_mu = 1e-5;     // precision 1: μ below which the factor returns to FP64
};

struct HistRow {
  double pres, dres, pobj, dobj, step;
  int piter, ppass, pstat, citer, cpass, cstat;
  double dmin, dmax, rho, delta, mu, chi;
};

enum IPMStatus {
  CONTINUE = 0, OPTIMAL, NEAR_OPTIMAL, STALLED, NUMERICAL_FAILURE, ITERATION_LIMIT, PRIMAL_INFEASIBLE,
  DUAL_INFEASIBLE, ILL_POSED, NEAR_PRIMAL_INFEASIBLE, NEAR_DUAL_INFEASIBLE, NEAR_ILL_POSED
};
enum KKTStatus { KKT_CONTINUE = 0, KKT_SOLVED, KKT_STAGNATED, KKT_ITMAX, KKT_NUMERICAL_FAILURE, KKT_IRMAX };

enum ProfKind {
  PROF_FACTOR = 0, PROF_UPDATE, PROF_PANEL, PROF_FSOLVE, PROF_BSOLVE, PROF_EXTEND, PROF_DAG, PROF_NKIND
};
constexpr int PROF_TRACE_BIT = 1 << 16;  // profile mask bit: record the large-front task timeline

// Freed device buffers stay cached for later allocations (hipFree waits for the whole device);
// release_cached_memory hands them back to the runtime and returns the bytes released.
size_t release_cached_memory();
size_t cached_memory();

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void alloc(size_t b);
  void clear(hipStream_t st);
  template <class T>
  T* as() const { return (T*)p; }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t st);
};

struct Level {
  std::vector<int> fronts;   // fronts at this level
  // extend-add items (children of this level's fronts) and their tile prefix, per round (round k:
  // the k-th child of each front)
  std::vector<int> ea_off, ea_n, ea_tiles;
  // panel / update steps: (offset into step table, count, blocks) per panel step k
  std::vector<int> pan_off, pan_n, pan_blk, upd_off, upd_n, upd_blk;
  std::vector<double> pan_fl, upd_fl;  // algorithmic flops of each launch
  int items_off = 0;         // small fronts of the level in the solve item table
  int nsmall = 0;
  std::vector<int> big;      // indices into Solver::big_ of this level's large fronts
  double solve_bytes = 0.0;  // factor bytes one triangular sweep reads at this level
  double big_flops = 0.0;    // algorithmic flops of this level's large fronts
};

class Gang;
struct GOp;

// Host copies of the solver's device lists, kept for the gang's combined tables (csrc/gang.cpp)
struct HostLists {
  std::vector<Tile> tn, tm, sp[3];  // n-space / m-space tiles; row tiles of B, Bᵀ, Q
  std::vector<int> psd, dense, soc, td, childlist, ea_items, ea_tile0, solve_items;
  std::vector<StepDesc> steps;
};

class Solver {
  friend class Gang;

 public:
  struct Deferred {};
  Solver(const Problem& prob, const Settings& st, const double* p0, const double* d0, const double* y0);
  // host analysis only (symbolic factor, permutation, equilibration); device_setup() completes it
  Solver(const Problem& prob, const Settings& st, Deferred);
  void device_setup(const double* p0, const double* d0, const double* y0);
  double estimate_device_bytes() const;     // upper bound of device_setup's allocations
  size_t device_bytes() const { return dev_bytes_; }  // what device_setup allocated
  ~Solver();
  int step();
  int solve();
  void result(double* p, double* d, double* y, double* info);  // info: mu pres dres pobj dobj niter nsolve status
  const std::vector<HistRow>& history();  // settles a pending step length first
  // diagnostics
  double exact_flops() const { return sym_.exact_flops; }
  double front_flops() const;
  int nfront() const { return (int)fr_.size(); }
  int nlevel() const { return (int)levels_.size(); }
  double last_factor_ms() const { return fac_ms_; }
  double total_factor_ms() const { return fac_ms_tot_; }
  int factor_calls() const { return fac_calls_; }
  int factor_calls_fp32() const { return fac_fp32_calls_; }
  double host_setup_ms() const { return host_ms_; }      // init on the host: analysis, scaling, BᵀB
  double device_setup_ms() const { return device_ms_; }  // init on the device: buffers, uploads, start
  double device_setup_part(int k) const { return dev_parts_[k]; }
  double build_device_mark(int k) const { return bd_parts_[k]; }
  void dense_factor(double* out);  // n×n lower factor (debug; small n)
  int kkt_factor(double delta, double* rho);  // initkkt! on the current H
  // initkkt!(kkt, H; δ) with the caller's H (internal vertex partition, both triangles); the KKT
  // solves then use it until the next scale / step / kkt_factor
  int kkt_factor_h(const BSM& H, double delta, double* rho);
  // frule! / rrule! (differential.jl:5-75, 252-334) at the last iterate, in the caller's space
  void frule(const double* df, const double* dg, double dmu, double* dp, double* dy, double* dd);
  double rrule(const double* dp, const double* dy, const double* dd, double* df, double* dg);
  // cone fan-outs on the iterate, in the solver's internal (permuted, scaled) coordinates
  // (cone.jl:146-553); scale replaces the iterate by (p, d)
  bool cone_scale(const double* p, const double* d, double* spsd);
  void cone_blocks(double* out);  // D_v per cone: n_v entries (nonneg), none (free), n_v² (others)
  size_t cone_blocks_size() const;
  const Problem& internal() const { return P_; }  // the permuted problem (before equilibration)
  const std::vector<double>& col_scale() const { return pscl_; }
  const std::vector<double>& row_scale() const { return yscl_; }
  void cone_corr0(double smu, double* r);
  void cone_corr(const double* dp, const double* dd, double smu, double* r);
  double cone_maxsteps(const double* dp, const double* dd, double step_frac);
  hipStream_t stream() const { return st_; }
  // live per-kernel timing with HIP events on the solver's stream (kinds: PROF_*)
  void set_profile(int mask) {
    prof_mask_ = mask;
    set_trace((mask & PROF_TRACE_BIT) != 0);
  }
  // per-task timeline of the persistent factorization of large front b (4 stamps per task)
  void set_trace(bool on);
  int get_trace(int b, long long* out, int len);
  void prof_flush();
  void prof_get(double* out) const;  // per kind: total ms, launches, algorithmic flops, bytes
  void reinit_public(const double* p0, const double* d0, const double* y0) { reinit(p0, d0, y0); }
  int n() const { return (int)n_; }
  int m() const { return (int)m_; }
  int last_status() const { return status_; }
  // lockstep batching: solve() inside a gang (join, IPM loop with every launch shared with the
  // other members, leave); the result is then read with result() as after solve()
  int solve_in(Gang& g);

 private:
  // setup
  void host_setup(const Problem& prob);
  void build_device();
  void build_schedule();
  // device ops. Inside a gang each of these is handed to the gang (one shared launch per phase
  // for all members); the raw_ forms launch on the solver's own tables
  bool ganged() const;
  // the gang's leader runs a member's own launches (on the leader's thread) between these
  static void set_issuing(const Solver* s);
  void gang_submit(GOp* op);  // the gang owns op from here (a blocking op: until it returns)
  void sync_scalars();
  void raw_sync_scalars();
  void set_scalar(int slot, double v);
  void set_scalar2(int s1, double v1, int s2, double v2);
  void smu(double nu, double mu, double mut);
  void cg_scalars(int op);
  void cone_scale_norm();
  void corr0_op(int kr);
  void corr_op(int kr, int kdp, int kdd);
  void zerofree(int kx);
  void factor_run(bool fp32);
  void raw_factor_run(bool fp32);
  void raw_assemble();
  void axpby(int space, int kx, int ky, int kz, Coef a, Coef b);
  void axpbypcz(int space, int kx, int ky, int kw, int kz, Coef a, Coef b, Coef c);
  void dot(int space, int kx, int ky, int mode, int slot, int op, int kw0 = 0, int kw1 = 0);
  void dots(const DotBatch& b);
  void lin2(const Lin& on, const Lin& om);
  void lin(const LinSet& N, const LinSet& M);
  void spmv_nm(int ntm, const SpmvN* N, const SpmvM* M);
  bool cg_runs(double atol, int itmax, bool stops) const;
  void spmv(int which, int kx, int ky, Coef a, Coef b);
  void lsolve(int kv);
  void usolve(int kv);
  void raw_lsolve(int kv, bool fp32);
  void raw_usolve(int kv, bool fp32);
  bool factor(double delta, double& rho);
  void run_factor(bool fp32);
  double maxsteps(int kdp, int kdd, double step_frac);
  void launch_maxsteps(int kdp, int kdd, double step_frac);
  void reinit(const double* p0, const double* d0, const double* y0);
  // KKT
 public:
  struct KKTOut {
    int ncg, nir, status;
    double dmin, dmax;
  };

 private:
  KKTOut solvekkt(int kx, int ky, int kf, int kg, bool warm, double ftol, double gtol, int irmin, bool stops);

 public:
  // solvekkt! (uzawa.jl:265-344) on the current H and factor, host vectors in internal coordinates
  KKTOut kkt_solve(double* x, double* y, const double* f, const double* g, bool warm, double ftol, double gtol,
                   int irmin, bool stops);

 private:
  int cg(int kx, int ky, int ku, double atol, int itmax, bool stops, int& status);
  double getaug();
  bool isstalled() const;
  int nearstatus(int status, double mu, double mus, double pobj, double dobj, double pres, double dres);
  bool isopt(double mu, double mus, double pobj, double dobj, double pres, double dres, double factor) const;
  bool isprimalinfeasible(double rtol, double atol);
  bool isdualinfeasible(double rtol, double atol);
  void compute_stats();

  Settings set_;
  Problem P_;          // permuted + scaled
  Symbolic sym_;
  // B of the permuted, scaled problem as stored nonzeros by rows and by columns, and the lower
  // triangle of BᵀB as scalar CSC
  std::vector<int> brp_, brc_, bcp_, bcr_, Li_;
  std::vector<double> brv_, bcv_, Lx_;
  std::vector<i64> Lp_;
  int64_t n_ = 0, m_ = 0;
  int nu_ = 0;
  double mu_t_ = 0.0;
  std::vector<double> pscl_, yscl_;
  double nB_ = 0, nf_ = 0, ng_ = 0, sf_ = 0, sg_ = 0, nH_ = 0, qoff2_ = 0;
  std::vector<HistRow> hist_;
  bool step_pending_ = false;  // the last update's step length is on the device only (S_TAU)
  bool step_fp32_ = false;     // that update's direction came from an FP32 factor
  void settle_step();
  std::vector<double> cghist_;
  int status_ = CONTINUE;

  // host copies of device descriptors
  std::vector<FrontDesc> fr_;
  std::vector<ConeDesc> cones_;
  std::vector<Level> levels_;
  std::vector<int> cone_front_;
  int ncone_ = 0;
  int64_t fsize_ = 0;

  hipStream_t st_ = nullptr;
  double* hsc_ = nullptr;  // pinned host scalar mirror
  InstDev idev_{};
  // device buffers
  DevBuf d_inst_, d_nv_, d_mv_, d_sc_, d_scr_, d_F_, d_fr_, d_cones_, d_cone_front_, d_cone_ofs_;
  DevBuf d_Bp_, d_Bc_, d_Bv_, d_Btp_, d_Btc_, d_Btv_, d_Qp_, d_Qc_, d_Qv_;
  DevBuf d_upos_, d_ul_, d_uq_, d_D_, d_cache_, d_uvec_, d_sepcols_, d_relidx_, d_childlist_;
  DevBuf d_steps_, d_ea_items_, d_ea_tile0_, d_solve_items_, d_ntiles_, d_mtiles_, d_ntofs_, d_mtofs_;
  DevBuf d_partial_, d_coneout_, d_qdiag_, d_dotcnt_, d_dpart_;
  int ntile_n_ = 0, ntile_m_ = 0;
  DevBuf d_sptiles_[3];  // row tiles of B, Bᵀ, Q for the group-per-row spmv
  DevBuf d_psdlist_, d_denselist_;  // PSD cones; cones with dense scaling blocks (SOC, PSD)
  DevBuf d_soclist_;                 // second-order cones
  int nsoc_ = 0;
  DevBuf d_bpsdlist_, d_bws_, d_bpart_;  // PSD cones above PSD_MAX, their workspaces, ‖H‖ partials
  int nbpsd_ = 0, bpsd_dmax_ = 0;
  BigPsd bpsd() const { return BigPsd{d_bpsdlist_.as<int>(), nbpsd_, bpsd_dmax_}; }
  DevBuf d_tdlist_;                 // exponential and power cones
  int npsd_ = 0, ndense_ = 0, ntd_ = 0;
  int cone_nmax_ = 0, dense_nmax_ = 0;  // largest cone order (non-free; dense-block cones)
  int nsptile_[3] = {0, 0, 0};
  int64_t nunion_ = 0;
  double fac_ms_ = 0.0, fac_ms_tot_ = 0.0;
  int fac_calls_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // profiling
  struct ProfRec {
    int ev;      // index of the event pair in the pool
    int kind;
    double flops, bytes;
  };
  int prof_mask_ = 0;
  std::vector<hipEvent_t> evpool_;
  std::vector<ProfRec> pending_;
  double prof_ms_[PROF_NKIND] = {0}, prof_n_[PROF_NKIND] = {0}, prof_fl_[PROF_NKIND] = {0},
         prof_by_[PROF_NKIND] = {0};
  int prof_next_ = 0;
  int prof_begin(int kind);
  void prof_end(int slot, int kind, double flops, double bytes);
  // large fronts
  struct BigHost {
    BigFront B;
    int front;
    long long sbase = 0;
    DevBuf tb, state, tasks, sll, spart, mtile, ntile, dinv, stage, trace, f32, dinv32, lsep, stg;
  };
  std::vector<std::unique_ptr<BigHost>> big_;
  // auxiliary streams for the independent large fronts of one level, with a fork event and one
  // join event per stream
  static constexpr int NAUX = 4;
  std::vector<hipStream_t> aux_;
  std::vector<hipEvent_t> auxev_;
  int naux_ = 0;
  void fork_aux(int k);
  void join_aux(int k);
  DevBuf d_bigtick_, d_bigstick_;
  DevBuf d_panflag_;  // small-front panel steps: diagonal factor published (epoch-valued)
  int epoch_ = 0;
  int dag_grid_ = 512;
  int fac_epoch_ = 0;     // factorization counter: a large front's tile flags hold it once final
  bool fac_fp32_ = false; // the current factor of the large fronts is FP32
  bool mixed_off_ = false;  // mixed precision fell back to FP64 for the rest of this solve
  std::vector<char> hist_fp32_;  // per history row: the iteration's large fronts were FP32
  double cur_mu_ = 0.0;   // μ of the iteration being factored (selects the factor precision)
  int fac_fp32_calls_ = 0;
  size_t dev_bytes_ = 0;  // device bytes allocated by device_setup
  double host_ms_ = 0.0, device_ms_ = 0.0;
  double dev_parts_[4] = {0.0, 0.0, 0.0, 0.0};  // device setup: stream/events, buffers+uploads, schedule, start
  double bd_parts_[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // build_device: ms from its start at its phases
  // front layout on the host (assembly positions)
  std::vector<i64> fofs_;
  std::vector<std::vector<int>> sepcols_;
  std::vector<int> colfront_;
  i64 front_pos(i64 i, i64 j) const;
  i64 qdiag_arc(i64 v) const;
  // a caller's H (kkt_factor_h): CSR rows, row tiles, assembly positions and values
  bool user_h_ = false;
  // the front needs no clearing before an assembly (one large front with its factor beside it,
  // no separator or children, diagonal cone blocks, every diagonal entry in the union pattern)
  bool fill_free_ = false;
  // fill_free_: the entries of the pattern that do not change with δ are in the front (written at
  // setup, and again by any full assembly); the union's first nunion_dyn_ entries are the rest
  bool static_ok_ = false;
  int64_t nunion_dyn_ = 0;
  CSRDev hsave_{};  // the CSR not in use (Q's while the caller's H is)
  DevBuf d_Hp_, d_Hc_, d_Hv_, d_htiles_, d_hpos_, d_hval_;
  int nhtiles_ = 0;
  i64 nhent_ = 0;
  void use_user_h(bool on);
  void dshadow(int kx, int ky);
  void cone_matvec_all(int kx, int ky, Coef a);
  void diff_setup(double& ftol, double& gtol);
  void check_abort();
  // gang membership
  HostLists hl_;
  Gang* gang_ = nullptr;
  int gslot_ = -1;
  int gang_prof_ = 0;          // the profile mask to restore on leaving
  hipStream_t own_st_ = nullptr;
};

// a process-wide monotone counter for the epochs of factorizations and solve launches: tile flags
// and solve hand-off words hold the epoch of the launch that wrote them, so every launch on any
// front (alone or in a gang) needs a value above all earlier ones
int next_epoch();

// the message mb_last_error() returns on the calling thread
void set_last_error(const std::string& msg);

// restatement of equilibrate! (src/IPM/src/scaling/scaling.jl:122-184) on host BSMs
void equilibrate(BSM& B, BSM& Q, std::vector<double>& f, std::vector<double>& g, int itmax, double tol,
                 std::vector<double>& pscl, std::vector<double>& yscl);

}  // namespace mb
