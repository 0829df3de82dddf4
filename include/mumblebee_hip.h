/*
 * mumblebee_hip.h — C ABI of the MI355X conic interior-point solve path.
 *
 * The entry points replace, one for one, the Julia calls of Mumblebee.jl's IPM solve path
 * (paths relative to the Mumblebee.jl repository root):
 *
 *   mb_problem_create        IPMProblem(Q, B, f, g, μ, K, s; compress)   src/IPM/src/ipm/problem.jl:64-127
 *   mb_problem_create_bsm    IPMProblem(Q::BlockSparseMatrix, B::BlockSparseMatrix, f, g, μ, K)
 *                                                                          src/IPM/src/ipm/problem.jl:49-53
 *   mb_problem_symbolic*     symbkkt → CliqueTrees.symbolic               src/IPM/src/kkt/kkt.jl:23-53
 *   mb_settings_default      IPMSettings{Float64}() defaults              src/IPM/src/ipm/settings/ipm.jl:1-18
 *   mb_solver_create         init(prob, settings; p0, d0, y0)             src/IPM/src/ipm/solver/ipm.jl:431-491
 *   mb_solver_step           step!(solver)                                src/IPM/src/ipm/solver/ipm.jl:577-774
 *   mb_solver_solve          solve!(solver)                               src/IPM/src/ipm/solver/solver.jl:119-150
 *   mb_solver_result         result(solver, status) → IPMResult           src/IPM/src/ipm/solver/ipm.jl:30-74
 *   mb_solver_history        IPMHistory rows                              src/IPM/src/ipm/history/ipm.jl:1-24
 *   mb_solver_kkt_factor     initkkt!(kkt, H; δ) → (ok, ρ)                src/IPM/src/kkt/uzawa.jl:73-115
 *   mb_solver_kkt_factor_h   initkkt!(kkt, H; δ) with the caller's H       src/IPM/src/kkt/uzawa.jl:73-76
 *   mb_solver_kkt_solve      solvekkt!(fstop, gstop, kkt, x, y, H, B, f, g; …)  src/IPM/src/kkt/uzawa.jl:265-344
 *   mb_solver_frule          frule!(Δp, Δy, Δd, s, Δf, Δg, Δμ)            src/IPM/src/ipm/differential.jl:5-75
 *   mb_solver_rrule          rrule!(Δf, Δg, s, Δp, Δy, Δd)                src/IPM/src/ipm/differential.jl:252-334
 *   mb_cone_scale            scale!(sched, K, H, Q, caches, p, d, …)      src/IPM/src/cone/cone.jl:146-247
 *   mb_cone_corr0            initpredictor! / corr0!                      src/IPM/src/cone/cone.jl:361-444
 *   mb_cone_corr             initcorrector! / corr!                       src/IPM/src/cone/cone.jl:261-352
 *   mb_cone_maxsteps         maxsteps(sched, K, p, d, Δp, Δd, …)          src/IPM/src/cone/cone.jl:455-553
 *   mb_solver_solve_in,      solve!(solver) for many solvers at once      src/IPM/src/ipm/solver/solver.jl:119-150
 *   mb_solve_batch           (lockstep gang: no counterpart; each result equals solve!'s)
 *
 * Sparse inputs are CSC with 0-based int64 indices, rows sorted within each column (stored zeros
 * are kept, as in Julia's SparseMatrixCSC). Cone kinds: 0 nonnegative, 1 free ("cofree"),
 * 2 second-order, 3 PSD (svec), 4 exponential, 5 power (cone_param = alpha).
 *
 * Errors: functions returning a pointer return NULL, functions returning int return a negative
 * value; mb_last_error() describes the last failure of the calling thread. Nothing is thrown
 * across the ABI. Handles are not re-entrant: one host thread per solver at a time.
 *
 * Status codes follow the reference's @enum order: IPMStatus CONTINUE=0, OPTIMAL=1,
 * NEAR_OPTIMAL=2, STALLED=3, NUMERICAL_FAILURE=4, ITERATION_LIMIT=5, PRIMAL_INFEASIBLE=6,
 * DUAL_INFEASIBLE=7, ILL_POSED=8, NEAR_PRIMAL_INFEASIBLE=9, NEAR_DUAL_INFEASIBLE=10,
 * NEAR_ILL_POSED=11 (src/IPM/src/ipm/ipm.jl:1); KKTStatus KKT_CONTINUE=0, KKT_SOLVED=1,
 * KKT_STAGNATED=2, KKT_ITMAX=3, KKT_NUMERICAL_FAILURE=4, KKT_IRMAX=5 (src/IPM/src/kkt/kkt.jl:1).
 */
#ifndef MUMBLEBEE_HIP_H
#define MUMBLEBEE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mb_problem mb_problem;
typedef struct mb_solver mb_solver;

/* IPMSettings (src/IPM/src/ipm/settings/ipm.jl:1-18) plus build-specific knobs. */
typedef struct mb_settings {
  int32_t verbose;
  double step_frac;
  double feas_tol;
  double gap_tol;
  int32_t max_iter;
  double near_factor;
  double stall_tol;
  double refine_stall_tol;
  int32_t scale_max_iter;
  int32_t refine_max_iter;
  int32_t newton_max_iter;
  double aug_tol;
  double infeas_abs;
  double infeas_rel;
  int32_t amalgamate;   /* relaxed supernode amalgamation (1 default) */
  int32_t precision;    /* 0: FP64 factor; 1: mixed — large fronts factored in FP32 on FP32 MFMA,
                           FP64 Uzawa/CG refinement, FP64 factor once μ < mixed_mu */
  int32_t big_front_ld; /* fronts of this order or larger take the persistent kernels (512) */
  int32_t dag_grid;     /* workgroups of the persistent factorization (0: one per CU) */
  double mixed_mu;      /* precision 1: μ below which the factor returns to FP64 (1e-5) */
} mb_settings;

/* Fill the reference defaults. */
void mb_settings_default(mb_settings* s);

/* IPMProblem(Q, B, f, g, μ, K, s; compress) — problem.jl:64-127. */
mb_problem* mb_problem_create(int64_t n, int64_t m, const int64_t* Q_colptr, const int64_t* Q_rowval,
                              const double* Q_nzval, const int64_t* B_colptr, const int64_t* B_rowval,
                              const double* B_nzval, const double* f, const double* g, double mu,
                              int64_t ncone, const int32_t* cone_kind, const double* cone_param,
                              const int64_t* cone_size, double compress);
/* IPMProblem(Q::BlockSparseMatrix, B::BlockSparseMatrix, f, g, μ, K) — problem.jl:49-53, the form
 * every reference example builds with blocksparse(...): the caller's block-sparse Q and B as they
 * are (no column partition or compression; identity P1, P2). The arrays are those of
 * block_sparse_matrix.jl:10-109: vertex v owns columns xcol[v] .. xcol[v+1]−1, out vertex u rows
 * xrow[u] .. xrow[u+1]−1; arc e of vertex v (xsrc[v] .. xsrc[v+1]−1) targets out vertex tgt[e] with
 * the dense column-major nrows(tgt[e]) × ncols(v) block val[xblk[e] .. xblk[e+1]−1]; targets
 * strictly increasing within a vertex. Q is nvtx × nvtx over the same partition xcol (its out
 * vertices are the vertices). index_base = 1 takes Julia's 1-based arrays unchanged (every offset
 * and index is read minus one); 0 takes 0-based arrays. cone_kind[v] / cone_param[v]: the cone
 * of vertex v. f (length n = xcol[nvtx] − base) and g (m = xrow[nout] − base) in the same column
 * and row order. Returns NULL on an inconsistent layout, with the reason in mb_last_error(). */
mb_problem* mb_problem_create_bsm(int64_t nvtx, int64_t nout, const int64_t* xcol, const int64_t* xrow,
                                  const int64_t* B_xsrc, const int64_t* B_tgt, const int64_t* B_xblk,
                                  const double* B_val, const int64_t* Q_xsrc, const int64_t* Q_tgt,
                                  const int64_t* Q_xblk, const double* Q_val, const double* f, const double* g,
                                  double mu, const int32_t* cone_kind, const double* cone_param,
                                  int32_t index_base);
void mb_problem_destroy(mb_problem* p);

/* dims[0..6] = nvtx, nout, B narc, B nbnz, Q narc, Q nbnz, n */
int mb_problem_dims(const mb_problem* p, int64_t* dims);

/* Block-sparse Q and B in the layout of block_sparse_matrix.jl:10-158 (0-based), permuted f, g,
 * row / column permutations P1, P2 (new -> old), per-vertex cone kinds. Any pointer may be NULL. */
int mb_problem_export(const mb_problem* p, int64_t* P1, int64_t* P2, int32_t* vertex_kind, int64_t* B_xsrc,
                      int64_t* B_xcol, int64_t* B_xrow, int64_t* B_xblk, int64_t* B_tgt, double* B_val,
                      int64_t* Q_xsrc, int64_t* Q_xblk, int64_t* Q_tgt, double* Q_val, double* f, double* g);

/* Symbolic analysis of δH + BᵀB (kkt.jl:23-53 with the build's own ordering).
 * dims[0..2] = nvtx, nfront, total separator length. */
int mb_problem_symbolic_dims(mb_problem* p, int32_t amalgamate, int64_t* dims);
/* vperm[nvtx] (new -> old vertex), front_vtx[nfront+1], sep_ptr[nfront+1], sep[...] (new vertex
 * indices), parent[nfront], flops[2] = {exact, amalgamated}. */
int mb_problem_symbolic(mb_problem* p, int32_t amalgamate, int64_t* vperm, int64_t* front_vtx, int64_t* sep_ptr,
                        int64_t* sep, int64_t* parent, double* flops);

/* init(prob, settings; p0, d0, y0): symbolic analysis, equilibration, device allocation,
 * starting point. p0 / d0 / y0 may be NULL (user space, length n / n / m). */
mb_solver* mb_solver_create(const mb_problem* p, const mb_settings* s, const double* p0, const double* d0,
                            const double* y0);
void mb_solver_destroy(mb_solver* s);

/* One IPM iteration; returns the IPMStatus (CONTINUE = 0 to keep going). */
int mb_solver_step(mb_solver* s);
/* Iterate until the status leaves CONTINUE; returns it. */
int mb_solver_solve(mb_solver* s);

/* IPMResult: p, d (length n), y (length m) in user space; info[0..7] = mu, pres, dres, pobj,
 * dobj, niter, nsolve, status. Any pointer may be NULL. */
int mb_solver_result(mb_solver* s, double* p, double* d, double* y, double* info);

int64_t mb_solver_niter(const mb_solver* s);
/* rows[niter][17] = pres dres pobj dobj step piter ppass pstat citer cpass cstat dmin dmax rho
 * delta mu chi (history/ipm.jl:1-24) */
int mb_solver_history(const mb_solver* s, double* rows);

/* stats[0..18] = exact factor flops, front flops (after amalgamation), nfront, nlevel,
 * total factorization ms, factorization calls, last factorization ms, factorizations whose
 * large fronts were FP32 (precision 1), init ms on the host (symbolic analysis, equilibration,
 * BᵀB), init ms of the device setup (buffers, uploads, starting point), and that device setup's
 * parts: stream and events, buffers and uploads, launch schedule, starting point, and ms from the
 * start of the buffers part to the end of its phases: front layout, assembly map, cone tables,
 * allocations, sparse matrices */
int mb_solver_stats(const mb_solver* s, double* stats);

/* reinit!(solver; p0, d0, y0) (solver/ipm.jl:371-429): new starting point, empty history;
 * the symbolic factor and device buffers stay resident. */
int mb_solver_reinit(mb_solver* s, const double* p0, const double* d0, const double* y0);

/* Live kernel timing with HIP events on the solver's stream. mask bit k enables kind k:
 * 0 factorization, 1 trailing update, 2 panel, 3 forward solve, 4 backward solve, 5 extend-add,
 * 6 persistent large-front factorization.
 * out[4k..4k+3] = total ms, launches, algorithmic flops, algorithmic bytes of kind k; at most
 * len doubles are written; returns the number of kinds. Mask bit 16 records the task timeline
 * of the persistent large-front factorization. */
int mb_solver_set_profile(mb_solver* s, int32_t mask);
int mb_solver_profile(mb_solver* s, double* out, int32_t len);

/* Task timeline of the last factorization of large front `front` (in schedule order): eight
 * 100 MHz wall-clock stamps per task (claimed, updates done, last input ready, done; a diagonal
 * task also: factor panels done, diagonal-block inverses done, inverse complete, inverse
 * published; 0 where a task has no such stamp). Copies at most len values; returns 8 × the task
 * count (0 if there is no such front).
 * Diagnostic only; no counterpart in the reference. */
int mb_solver_dag_trace(mb_solver* s, int32_t front, int64_t* out, int32_t len);

/* initkkt!(kkt, H; δ) on the solver's current H (uzawa.jl:73-115): returns 1 on success,
 * 0 on a failed factorization, negative on error; *rho receives the shift used. */
int mb_solver_kkt_factor(mb_solver* s, double delta, double* rho);

/* initkkt!(kkt, H; δ) with the caller's H (uzawa.jl:73-76): H block-sparse over the solver's
 * internal vertex partition (mb_solver_layout's xcol, for rows and columns alike; arcs H_xsrc /
 * H_tgt / H_xblk / H_val as in mb_problem_create_bsm, both triangles, index_base 0 or 1) in its
 * internal coordinates. Factors δH + BᵀB (retry with ρ as mb_solver_kkt_factor); the following
 * mb_solver_kkt_solve calls use this H in their residuals, until the next mb_cone_scale,
 * mb_solver_step or mb_solver_kkt_factor brings back the solver's own H (Q and the cone blocks).
 * H's lower entries must lie in the factor's structure (the pattern of Q, BᵀB and the diagonal
 * blocks): else an error. Returns 1 on success, 0 on a failed factorization, negative on error. */
int mb_solver_kkt_factor_h(mb_solver* s, const int64_t* H_xsrc, const int64_t* H_tgt, const int64_t* H_xblk,
                           const double* H_val, int32_t index_base, double delta, double* rho);

/* Differentiation of the solution map at the solver's last iterate (after mb_solver_solve or
 * steps), in the caller's space (the same as mb_solver_result): a fresh factor of δH + BᵀB at
 * getaug's δ, one KKT solve each, the KKT tolerances of differential.jl:5-16.
 * mb_solver_frule: frule!(Δp, Δy, Δd, s, Δf, Δg, Δμ) (differential.jl:5-75): the tangent of
 * (p, y, d) for perturbations Δf (n), Δg (m) of f and g and Δμ of the barrier target.
 * mb_solver_rrule: rrule!(Δf, Δg, s, Δp, Δy, Δd) (differential.jl:252-334): the cotangent; *dmu
 * receives the μ component. Return 0, or negative on error (a solver with no iterations, a
 * failed factorization). */
int mb_solver_frule(mb_solver* s, const double* df, const double* dg, double dmu, double* dp, double* dy,
                    double* dd);
int mb_solver_rrule(mb_solver* s, const double* dp, const double* dy, const double* dd, double* df, double* dg,
                    double* dmu);

/* Cone fan-outs on the solver's iterate in its internal coordinates (columns permuted by the
 * problem construction and scaled by the equilibration, as the reference's IPMSolver holds p and
 * d). All vectors have length n.
 * mb_cone_scale: sets the iterate to (p, d), forms every cone's scaling block and cache; returns
 * 1 if every cone accepted the point (the reference's all(flags)), 0 if not, −1 on error; spsd ←
 * Σ⟨p*, d*⟩. mb_cone_blocks: the blocks D_v in cone order — n_v entries for a nonnegative cone
 * (the diagonal), none for a free cone, n_v² column-major for the others; mb_cone_blocks_size
 * gives the total. mb_cone_corr0: r = the predictor right-hand side −d + σμ(−f′(p)) per cone.
 * mb_cone_corr: r = the Mehrotra corrector for directions (dp, dd). mb_cone_maxsteps: step ←
 * the largest τ ≤ 1 keeping both iterates interior, × step_frac except on free cones. */
int64_t mb_solver_ncone(const mb_solver* s);
/* The internal layout: P1[i] = the caller's row of internal row i and rowscale[i] its
 * equilibration factor (m entries; internal y = y_caller[P1[i]] / rowscale[i]); P2[j] = the
 * caller's column of internal column j and colscale[j] its factor (n entries; internal p =
 * p_caller[P2[j]] / colscale[j], internal d = d_caller[P2[j]] · colscale[j]); cone v owns internal
 * columns xcol[v]..xcol[v+1]−1 and has kind[v], param[v] (ncone entries, xcol ncone+1). */
int mb_solver_layout(const mb_solver* s, int64_t* P1, int64_t* P2, int64_t* xcol, int32_t* kind, double* param,
                     double* rowscale, double* colscale);
/* solvekkt! on the solver's current H (from the last scale or step) and factor (from the last
 * mb_solver_kkt_factor or step): H x − Bᵀ y = f, B x = g by Uzawa refinement with the factor of
 * δH + BᵀB, then CG on the dual (uzawa.jl:319-507). x (n), y (m): in/out, read when warm ≠ 0;
 * f (n), g (m). ftol, gtol: absolute residual tolerances; irmin: minimum refinement passes;
 * use_stop ≠ 0 adds the scaled feasibility stops of the IPM's closures (solver/ipm.jl:301-307).
 * counts[3] = ncg, nir, KKTStatus; window[2] = δmin, δmax (kktwindow, uzawa.jl:199-227). */
int mb_solver_kkt_solve(mb_solver* s, double* x, double* y, const double* f, const double* g, int warm, double ftol,
                        double gtol, int irmin, int use_stop, int32_t* counts, double* window);
int mb_cone_scale(mb_solver* s, const double* p, const double* d, double* spsd);
int64_t mb_cone_blocks_size(const mb_solver* s);
int mb_cone_blocks(mb_solver* s, double* out);
int mb_cone_corr0(mb_solver* s, double sigma_mu, double* r);
int mb_cone_corr(mb_solver* s, const double* dp, const double* dd, double sigma_mu, double* r);
int mb_cone_maxsteps(mb_solver* s, const double* dp, const double* dd, double step_frac, double* step);

/* The frontal POTRF of the supernodal factorization on one dense front the caller holds in HBM:
 * the dense kernel CliqueTrees' cholesky! runs per front (called at src/IPM/src/kkt/uzawa.jl:93
 * and :111), here the persistent FP64-MFMA tile kernel of the large fronts. A: device pointer to
 * an n×n column-major matrix with leading dimension lda ≥ max(n, 1); its lower triangle is
 * replaced by L with A = L Lᵀ, the strict upper triangle is not touched. grid: workgroups of the
 * persistent kernel (≤ 0: one per CU). stream: a hipStream_t (NULL: the null stream); the call
 * returns after the factorization has finished on it. Returns 0 on success, 1 if a pivot was not
 * positive (A is not positive definite: the reference's issuccess(F) is false and the retry with
 * ρ follows), negative on error. The workspace (≈ 48·T² + 32 KiB·T bytes for T = ⌈n/64⌉ tiles) is
 * allocated per call from the library's cache of device memory. */
int mb_dense_cholesky(double* A, int64_t n, int64_t lda, int32_t grid, void* stream);

/* mb_dense_cholesky with no device allocation: no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc). Same arguments, results and return codes as mb_dense_cholesky, and the
 * same kernels; the workspace is one file-scope device array per GPU (4.4 MB, kept until the
 * process ends), so the order is limited to n ≤ 8192 (larger n: an error, nothing launched). The
 * call allocates nothing: no hipMalloc, hipFree or allocating library routine. Every call on one
 * GPU shares that GPU's array and each call resets what it reads: calls that overlap on one GPU
 * (several host threads or streams at once) race and must be serialised by the caller. */
int mb_dense_cholesky_noalloc(double* A, int64_t n, int64_t lda, int32_t grid, void* stream);

/* Debug: the current factor as a dense n×n column-major lower triangle (small n only). */
int mb_solver_dense_factor(mb_solver* s, double* L);

/* Lane-layout probes of v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 (256 doubles out:
 * element t of lane l at out[4l + t]). */
int mb_probe_mfma_f64(double* out);
int mb_probe_mfma_f32(double* out);

/* ------------------------------------------------------------------------------------------
 * Serving: a continuous-batching request loop on one GPU. No counterpart in the reference; it
 * serves many independent solve(prob, settings) calls (solver/solver.jl:119-150) at once.
 * nslots slots (host threads, one HIP stream each) take requests in arrival order; a request is
 * admitted once a slot is free and its device-memory bound fits mem_budget_bytes (≤ 0: 90 % of
 * the free device memory) beside the running ones; its kernels then overlap on the device with
 * the running solves. dag_grid ≤ 0 gives each slot 1/nslots of the CUs for the persistent
 * factorization. Results are copied to host memory when a solve ends and freed on release.
 * ------------------------------------------------------------------------------------------ */
typedef struct mb_server mb_server;
mb_server* mb_server_create(int32_t nslots, double mem_budget_bytes, int32_t dag_grid);
/* Finishes the queued and running requests, then frees the server. */
void mb_server_destroy(mb_server* s);
/* Queue solve(prob, settings); the problem must stay alive until its result is collected.
 * Returns a ticket ≥ 0, or −1. */
int64_t mb_server_submit(mb_server* s, const mb_problem* p, const mb_settings* st);
/* Wait up to timeout_s (< 0: no limit) for a ticket. Returns its IPMStatus, −2 on timeout, −1 on
 * error. p, d (length n), y (length m), info[8] as in mb_solver_result; times[5] = submitted,
 * admitted and finished (mb_clock seconds), device bytes allocated, device bytes reserved at
 * admission (the request's upper bound). Any pointer may be NULL. */
int mb_server_wait(mb_server* s, int64_t ticket, double timeout_s, double* p, double* d, double* y, double* info,
                   double* times);
/* Forget a collected ticket. */
int mb_server_release(mb_server* s, int64_t ticket);
/* out[5] = queued, running, completed, device bytes reserved now, peak bytes reserved. */
int mb_server_stats(mb_server* s, double* out);
/* The same loop with its slots as members of one gang (below): the solves of the admitted
 * requests advance in lockstep and share every launch; a finished request's slot takes the next
 * admitted one between two launches. */
mb_server* mb_server_create_gang(int32_t nslots, double mem_budget_bytes);

/* ------------------------------------------------------------------------------------------
 * Lockstep batching ("gang"). No counterpart in the reference: several independent
 * solve!(solver) calls (solver/solver.jl:119-150) advance together, each taking the reference's
 * decisions for its own instance, while the device work of each phase of step! is one launch
 * for all of them (vector kernels over concatenated tile lists, the cone kernels over one cone
 * table, the small fronts of one elimination-tree level of all instances per launch, the large
 * fronts of all instances in one persistent factorization and one pipelined solve launch, one
 * scalar read-back per phase). Each result is bit for bit the result of the same solver run
 * alone. capacity ≤ 32 members at a time; dag_grid ≤ 0: one workgroup per CU for the combined
 * factorization; profile ≠ 0: time the combined factorizations with events.
 * ------------------------------------------------------------------------------------------ */
typedef struct mb_gang mb_gang;
mb_gang* mb_gang_create(int32_t capacity, int32_t dag_grid, int32_t profile);
void mb_gang_destroy(mb_gang* g);
/* solve!(solver) as a member of the gang: joins (waiting for a free slot), runs the IPM loop and
 * leaves; returns the IPMStatus or −1. Call it from one host thread per solver; read the result
 * with mb_solver_result afterwards. */
int mb_solver_solve_in(mb_solver* s, mb_gang* g);
/* n solvers through the gang from min(n, capacity) host threads of the library, a finished
 * solver's slot taken by the next one; status[i] = IPMStatus of solver i (−1: failed). Returns 0,
 * or −1 if any failed (mb_last_error: the first message). */
int mb_solve_batch(mb_solver* const* s, int32_t n, mb_gang* g, int32_t* status);
/* out[18] = ticks, launches, ops, ops in combined launches, read-backs, joins, combined
 * factorizations, their large fronts, their algorithmic flops, their time in ms (profile ≠ 0),
 * persistent factorization launches, flops in them, their time in ms, members now; then the
 * scheduler thread's ms admitting members, running their fibers, issuing ticks (read-backs
 * included) and waiting for the device in read-backs. */
int mb_gang_stats(mb_gang* g, double* out);

/* Seconds on the steady clock the server's timestamps use. */
double mb_clock(void);
/* out[3] = compute units, free and total device bytes of the current device. */
int mb_device_info(double* out);
/* Device memory of destroyed solvers stays cached for later solvers (freeing it would wait for
 * every kernel on the device). mb_cached_bytes: bytes cached now; mb_release_cached_memory: hand
 * them back to the HIP runtime (waits for the device) and return the bytes released. */
int64_t mb_cached_bytes(void);
int64_t mb_release_cached_memory(void);

const char* mb_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MUMBLEBEE_HIP_H */
