// extern "C" boundary (include/mumblebee_hip.h). Catches every exception; errors are
// reported through return codes and mb_last_error().
#include <cstring>
#include <memory>
#include <string>

#include "../../include/mumblebee_hip.h"
#include "gang.h"
#include "solver.h"
#include "server.h"

#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& fn) {
  try {
    return fn();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}
}  // namespace

namespace mb {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace mb

struct mb_problem {
  mb::Problem P;
  int64_t n = 0, m = 0;
  int sym_amalg = -1;
  mb::Symbolic sym;
};

struct mb_solver {
  std::unique_ptr<mb::Solver> S;
};

static mb::CSC make_csc(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowval, const double* nzval) {
  mb::CSC A;
  A.m = m;
  A.n = n;
  A.colptr.assign(colptr, colptr + n + 1);
  int64_t nz = A.colptr[n];
  A.rowval.assign(rowval, rowval + nz);
  A.nzval.assign(nzval, nzval + nz);
  for (int64_t j = 0; j < n; ++j) {
    if (A.colptr[j + 1] < A.colptr[j]) throw std::invalid_argument("CSC colptr not monotone");
    for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      if (A.rowval[p] < 0 || A.rowval[p] >= m) throw std::invalid_argument("CSC row index out of range");
      if (p > A.colptr[j] && A.rowval[p] <= A.rowval[p - 1]) throw std::invalid_argument("CSC rows not sorted");
    }
  }
  return A;
}

static void ensure_symbolic(mb_problem* p, int amalg) {
  if (p->sym_amalg == amalg) return;
  p->sym = mb::analyse_kkt(p->P.B, p->P.Q, amalg != 0);
  p->sym_amalg = amalg;
}

static mb::Settings to_settings(const mb_settings* s) {
  mb::Settings st;
  if (!s) return st;
  st.verbose = s->verbose;
  st.step_frac = s->step_frac;
  st.feas_tol = s->feas_tol;
  st.gap_tol = s->gap_tol;
  st.max_iter = s->max_iter;
  st.near_factor = s->near_factor;
  st.stall_tol = s->stall_tol;
  st.refine_stall_tol = s->refine_stall_tol;
  st.scale_max_iter = s->scale_max_iter;
  st.refine_max_iter = s->refine_max_iter;
  st.newton_max_iter = s->newton_max_iter;
  st.aug_tol = s->aug_tol;
  st.infeas_abs = s->infeas_abs;
  st.infeas_rel = s->infeas_rel;
  st.amalgamate = s->amalgamate;
  st.big_front_ld = s->big_front_ld > 0 ? s->big_front_ld : mb::BIG_LD;
  st.dag_grid = s->dag_grid;
  if (s->precision != 0 && s->precision != 1) throw std::invalid_argument("precision must be 0 or 1");
  st.precision = s->precision;
  st.mixed_mu = s->mixed_mu > 0 ? s->mixed_mu : mb::Settings().mixed_mu;
  return st;
}

extern "C" {

void mb_settings_default(mb_settings* s) {
  mb::Settings d;
  s->verbose = d.verbose;
  s->step_frac = d.step_frac;
  s->feas_tol = d.feas_tol;
  s->gap_tol = d.gap_tol;
  s->max_iter = d.max_iter;
  s->near_factor = d.near_factor;
  s->stall_tol = d.stall_tol;
  s->refine_stall_tol = d.refine_stall_tol;
  s->scale_max_iter = d.scale_max_iter;
  s->refine_max_iter = d.refine_max_iter;
  s->newton_max_iter = d.newton_max_iter;
  s->aug_tol = d.aug_tol;
  s->infeas_abs = d.infeas_abs;
  s->infeas_rel = d.infeas_rel;
  s->amalgamate = d.amalgamate;
  s->precision = d.precision;
  s->big_front_ld = d.big_front_ld;
  s->dag_grid = d.dag_grid;
  s->mixed_mu = d.mixed_mu;
}

mb_problem* mb_problem_create(int64_t n, int64_t m, const int64_t* Q_colptr, const int64_t* Q_rowval,
                              const double* Q_nzval, const int64_t* B_colptr, const int64_t* B_rowval,
                              const double* B_nzval, const double* f, const double* g, double mu, int64_t ncone,
                              const int32_t* cone_kind, const double* cone_param, const int64_t* cone_size,
                              double compress) {
  mb_problem* out = nullptr;
  guard([&]() {
    if (n < 0 || m < 0 || ncone < 0) throw std::invalid_argument("negative dimension");
    mb::CSC Q = make_csc(n, n, Q_colptr, Q_rowval, Q_nzval);
    mb::CSC B = make_csc(m, n, B_colptr, B_rowval, B_nzval);
    std::vector<mb::Cone> K(ncone);
    std::vector<int64_t> s(cone_size, cone_size + ncone);
    for (int64_t k = 0; k < ncone; ++k) {
      if (cone_kind[k] < 0 || cone_kind[k] > 5) throw std::invalid_argument("unknown cone kind");
      if (cone_size[k] < 0) throw std::invalid_argument("negative cone size");
      K[k].kind = cone_kind[k];
      K[k].param = cone_param ? cone_param[k] : 0.0;
    }
    std::vector<double> fv(f, f + n), gv(g, g + m);
    auto p = std::make_unique<mb_problem>();
    p->P = mb::make_problem(Q, B, fv, gv, mu, K, s, compress);
    p->n = n;
    p->m = m;
    out = p.release();
    return 0;
  });
  return out;
}

// One block-sparse matrix from the caller's arrays in the layout of block_sparse_matrix.jl:10-109
// (index_base 1: Julia's 1-based arrays as they are; 0: 0-based), checked: monotone offsets,
// targets in range and strictly increasing within each vertex (block_sparse_matrix.jl:703-711),
// block lengths nrows(tgt) × ncols(v).
static mb::BSM make_bsm(int64_t nout, int64_t nvtx, const int64_t* xcol, const int64_t* xrow, const int64_t* xsrc,
                        const int64_t* tgt, const int64_t* xblk, const double* val, int64_t base, const char* name) {
  auto bad = [&](const std::string& what) { throw std::invalid_argument(std::string(name) + ": " + what); };
  if (nout < 0 || nvtx < 0) bad("negative size");
  if (!xcol || !xrow || !xsrc) bad("null offset array");
  mb::BSM A;
  A.nout = nout;
  A.nvtx = nvtx;
  A.xcol.resize(nvtx + 1);
  A.xrow.resize(nout + 1);
  A.xsrc.resize(nvtx + 1);
  for (int64_t v = 0; v <= nvtx; ++v) {
    A.xcol[v] = xcol[v] - base;
    A.xsrc[v] = xsrc[v] - base;
    if (v > 0 && (A.xcol[v] < A.xcol[v - 1] || A.xsrc[v] < A.xsrc[v - 1])) bad("xcol / xsrc not monotone");
  }
  for (int64_t u = 0; u <= nout; ++u) {
    A.xrow[u] = xrow[u] - base;
    if (u > 0 && A.xrow[u] < A.xrow[u - 1]) bad("xrow not monotone");
  }
  if (A.xcol[0] != 0 || A.xrow[0] != 0 || A.xsrc[0] != 0) bad("offsets must start at the index base");
  const int64_t narc = A.xsrc[nvtx];
  if (narc > 0 && (!tgt || !xblk)) bad("null arc array");
  A.tgt.resize(narc);
  A.xblk.resize(narc + 1);
  A.xblk[0] = narc >= 0 && xblk ? xblk[0] - base : 0;
  if (A.xblk[0] != 0) bad("xblk must start at the index base");
  for (int64_t v = 0; v < nvtx; ++v)
    for (int64_t e = A.xsrc[v]; e < A.xsrc[v + 1]; ++e) {
      A.tgt[e] = tgt[e] - base;
      A.xblk[e + 1] = xblk[e + 1] - base;
      if (A.tgt[e] < 0 || A.tgt[e] >= nout) bad("arc target out of range");
      if (e > A.xsrc[v] && A.tgt[e] <= A.tgt[e - 1]) bad("arc targets not strictly increasing within a vertex");
      if (A.xblk[e + 1] - A.xblk[e] != A.nrows(A.tgt[e]) * A.ncols(v)) bad("block length is not nrows × ncols");
    }
  const int64_t nbnz = A.xblk[narc];
  if (nbnz > 0 && !val) bad("null values");
  A.val.assign(val, val + nbnz);
  return A;
}

mb_problem* mb_problem_create_bsm(int64_t nvtx, int64_t nout, const int64_t* xcol, const int64_t* xrow,
                                  const int64_t* B_xsrc, const int64_t* B_tgt, const int64_t* B_xblk,
                                  const double* B_val, const int64_t* Q_xsrc, const int64_t* Q_tgt,
                                  const int64_t* Q_xblk, const double* Q_val, const double* f, const double* g,
                                  double mu, const int32_t* cone_kind, const double* cone_param,
                                  int32_t index_base) {
  mb_problem* out = nullptr;
  guard([&]() {
    if (index_base != 0 && index_base != 1) throw std::invalid_argument("index_base must be 0 or 1");
    auto p = std::make_unique<mb_problem>();
    mb::Problem& P = p->P;
    P.B = make_bsm(nout, nvtx, xcol, xrow, B_xsrc, B_tgt, B_xblk, B_val, index_base, "B");
    P.Q = make_bsm(nvtx, nvtx, xcol, xcol, Q_xsrc, Q_tgt, Q_xblk, Q_val, index_base, "Q");
    const int64_t n = P.B.ncol(), m = P.B.nrow();
    P.f.assign(f, f + n);
    P.g.assign(g, g + m);
    P.mu = mu;
    P.K.resize(nvtx);
    for (int64_t v = 0; v < nvtx; ++v) {
      if (cone_kind[v] < 0 || cone_kind[v] > 5) throw std::invalid_argument("unknown cone kind");
      P.K[v].kind = cone_kind[v];
      P.K[v].param = cone_param ? cone_param[v] : 0.0;
    }
    P.P1.resize(m);
    P.P2.resize(n);
    for (int64_t i = 0; i < m; ++i) P.P1[i] = i;
    for (int64_t j = 0; j < n; ++j) P.P2[j] = j;
    p->n = n;
    p->m = m;
    out = p.release();
    return 0;
  });
  return out;
}

void mb_problem_destroy(mb_problem* p) { delete p; }

int mb_problem_dims(const mb_problem* p, int64_t* dims) {
  return guard([&]() {
    dims[0] = p->P.B.nvtx;
    dims[1] = p->P.B.nout;
    dims[2] = p->P.B.narc();
    dims[3] = p->P.B.nbnz();
    dims[4] = p->P.Q.narc();
    dims[5] = p->P.Q.nbnz();
    dims[6] = p->n;
    return 0;
  });
}

int mb_problem_export(const mb_problem* p, int64_t* P1, int64_t* P2, int32_t* vertex_kind, int64_t* B_xsrc,
                      int64_t* B_xcol, int64_t* B_xrow, int64_t* B_xblk, int64_t* B_tgt, double* B_val,
                      int64_t* Q_xsrc, int64_t* Q_xblk, int64_t* Q_tgt, double* Q_val, double* f, double* g) {
  return guard([&]() {
    const mb::Problem& P = p->P;
    auto cp = [](auto* dst, const auto& v) {
      if (dst) std::copy(v.begin(), v.end(), dst);
    };
    cp(P1, P.P1);
    cp(P2, P.P2);
    if (vertex_kind)
      for (size_t v = 0; v < P.K.size(); ++v) vertex_kind[v] = P.K[v].kind;
    cp(B_xsrc, P.B.xsrc);
    cp(B_xcol, P.B.xcol);
    cp(B_xrow, P.B.xrow);
    cp(B_xblk, P.B.xblk);
    cp(B_tgt, P.B.tgt);
    cp(B_val, P.B.val);
    cp(Q_xsrc, P.Q.xsrc);
    cp(Q_xblk, P.Q.xblk);
    cp(Q_tgt, P.Q.tgt);
    cp(Q_val, P.Q.val);
    cp(f, P.f);
    cp(g, P.g);
    return 0;
  });
}

int mb_problem_symbolic_dims(mb_problem* p, int32_t amalgamate, int64_t* dims) {
  return guard([&]() {
    ensure_symbolic(p, amalgamate);
    int64_t tot = 0;
    for (auto& s : p->sym.front_sep) tot += (int64_t)s.size();
    dims[0] = (int64_t)p->sym.vperm.size();
    dims[1] = p->sym.nfront();
    dims[2] = tot;
    return 0;
  });
}

int mb_problem_symbolic(mb_problem* p, int32_t amalgamate, int64_t* vperm, int64_t* front_vtx, int64_t* sep_ptr,
                        int64_t* sep, int64_t* parent, double* flops) {
  return guard([&]() {
    ensure_symbolic(p, amalgamate);
    const mb::Symbolic& S = p->sym;
    if (vperm) std::copy(S.vperm.begin(), S.vperm.end(), vperm);
    if (front_vtx) std::copy(S.front_vtx.begin(), S.front_vtx.end(), front_vtx);
    int64_t k = 0;
    for (int64_t f = 0; f < S.nfront(); ++f) {
      if (sep_ptr) sep_ptr[f] = k;
      for (int64_t u : S.front_sep[f]) {
        if (sep) sep[k] = u;
        ++k;
      }
    }
    if (sep_ptr) sep_ptr[S.nfront()] = k;
    if (parent) std::copy(S.front_parent.begin(), S.front_parent.end(), parent);
    if (flops) {
      flops[0] = S.exact_flops;
      double t = 0.0;
      for (int64_t f = 0; f < S.nfront(); ++f) {
        double r = (double)(S.xcol[S.front_vtx[f + 1]] - S.xcol[S.front_vtx[f]]);
        double s = 0.0;
        for (int64_t u : S.front_sep[f]) s += (double)(S.xcol[u + 1] - S.xcol[u]);
        t += r * r * r / 3.0 + s * r * r + s * s * r;
      }
      flops[1] = t;
    }
    return 0;
  });
}

mb_solver* mb_solver_create(const mb_problem* p, const mb_settings* s, const double* p0, const double* d0,
                            const double* y0) {
  mb_solver* out = nullptr;
  guard([&]() {
    mb::Settings st = to_settings(s);
    auto h = std::make_unique<mb_solver>();
    h->S = std::make_unique<mb::Solver>(p->P, st, p0, d0, y0);
    out = h.release();
    return 0;
  });
  return out;
}

void mb_solver_destroy(mb_solver* s) { delete s; }

int mb_solver_step(mb_solver* s) { return guard([&]() { return s->S->step(); }); }
int mb_solver_solve(mb_solver* s) { return guard([&]() { return s->S->solve(); }); }

int mb_solver_result(mb_solver* s, double* p, double* d, double* y, double* info) {
  return guard([&]() {
    s->S->result(p, d, y, info);
    return 0;
  });
}

int64_t mb_solver_niter(const mb_solver* s) { return (int64_t)s->S->history().size(); }

int mb_solver_history(const mb_solver* s, double* rows) {
  return guard([&]() {
    size_t k = 0;
    for (const mb::HistRow& h : s->S->history()) {
      double r[17] = {h.pres, h.dres, h.pobj, h.dobj, h.step, (double)h.piter, (double)h.ppass, (double)h.pstat,
                      (double)h.citer, (double)h.cpass, (double)h.cstat, h.dmin, h.dmax, h.rho, h.delta, h.mu,
                      h.chi};
      std::memcpy(rows + 17 * k, r, sizeof(r));
      ++k;
    }
    return 0;
  });
}

int mb_solver_stats(const mb_solver* s, double* st) {
  return guard([&]() {
    st[0] = s->S->exact_flops();
    st[1] = s->S->front_flops();
    st[2] = s->S->nfront();
    st[3] = s->S->nlevel();
    st[4] = s->S->total_factor_ms();
    st[5] = s->S->factor_calls();
    st[6] = s->S->last_factor_ms();
    st[7] = s->S->factor_calls_fp32();
    st[8] = s->S->host_setup_ms();
    st[9] = s->S->device_setup_ms();
    for (int k = 0; k < 4; ++k) st[10 + k] = s->S->device_setup_part(k);
    for (int k = 0; k < 5; ++k) st[14 + k] = s->S->build_device_mark(k);
    return 0;
  });
}

int mb_solver_reinit(mb_solver* s, const double* p0, const double* d0, const double* y0) {
  return guard([&]() {
    s->S->reinit_public(p0, d0, y0);
    return 0;
  });
}

int mb_solver_set_profile(mb_solver* s, int32_t mask) {
  return guard([&]() {
    s->S->set_profile(mask);
    return 0;
  });
}

int mb_solver_profile(mb_solver* s, double* out, int32_t len) {
  return guard([&]() {
    s->S->prof_flush();
    double tmp[4 * mb::PROF_NKIND];
    s->S->prof_get(tmp);
    for (int k = 0; k < std::min<int>(len, 4 * mb::PROF_NKIND); ++k) out[k] = tmp[k];
    return (int)mb::PROF_NKIND;
  });
}

int mb_solver_dag_trace(mb_solver* s, int32_t front, int64_t* out, int32_t len) {
  return guard([&]() { return s->S->get_trace(front, (long long*)out, len); });
}

int mb_solver_kkt_factor(mb_solver* s, double delta, double* rho) {
  return guard([&]() { return s->S->kkt_factor(delta, rho); });
}

int mb_solver_kkt_factor_h(mb_solver* s, const int64_t* H_xsrc, const int64_t* H_tgt, const int64_t* H_xblk,
                           const double* H_val, int32_t index_base, double delta, double* rho) {
  return guard([&]() {
    if (index_base != 0 && index_base != 1) throw std::invalid_argument("index_base must be 0 or 1");
    const mb::BSM& B = s->S->internal().B;
    std::vector<int64_t> xc(B.xcol.begin(), B.xcol.end());
    for (auto& x : xc) x += index_base;  // the solver's partition in the caller's index base
    mb::BSM H = make_bsm(B.nvtx, B.nvtx, xc.data(), xc.data(), H_xsrc, H_tgt, H_xblk, H_val, index_base, "H");
    return s->S->kkt_factor_h(H, delta, rho);
  });
}

int mb_solver_frule(mb_solver* s, const double* df, const double* dg, double dmu, double* dp, double* dy,
                    double* dd) {
  return guard([&]() {
    s->S->frule(df, dg, dmu, dp, dy, dd);
    return 0;
  });
}

int mb_solver_rrule(mb_solver* s, const double* dp, const double* dy, const double* dd, double* df, double* dg,
                    double* dmu) {
  return guard([&]() {
    const double v = s->S->rrule(dp, dy, dd, df, dg);
    if (dmu) *dmu = v;
    return 0;
  });
}

int64_t mb_solver_ncone(const mb_solver* s) { return (int64_t)s->S->internal().K.size(); }

int mb_solver_layout(const mb_solver* s, int64_t* P1, int64_t* P2, int64_t* xcol, int32_t* kind, double* param,
                     double* rowscale, double* colscale) {
  return guard([&]() {
    const mb::Problem& P = s->S->internal();
    for (size_t i = 0; i < P.P1.size(); ++i) P1[i] = P.P1[i];
    const std::vector<double>& rs = s->S->row_scale();
    for (size_t i = 0; i < rs.size(); ++i) rowscale[i] = rs[i];
    for (size_t j = 0; j < P.P2.size(); ++j) P2[j] = P.P2[j];
    for (size_t v = 0; v < P.K.size(); ++v) {
      xcol[v] = P.B.xcol[v];
      kind[v] = P.K[v].kind;
      param[v] = P.K[v].param;
    }
    xcol[P.K.size()] = P.B.xcol[P.K.size()];
    const std::vector<double>& sc = s->S->col_scale();
    for (size_t j = 0; j < sc.size(); ++j) colscale[j] = sc[j];
    return 0;
  });
}

int mb_solver_kkt_solve(mb_solver* s, double* x, double* y, const double* f, const double* g, int warm, double ftol,
                        double gtol, int irmin, int use_stop, int32_t* counts, double* window) {
  return guard([&]() {
    mb::Solver::KKTOut o = s->S->kkt_solve(x, y, f, g, warm != 0, ftol, gtol, irmin, use_stop != 0);
    counts[0] = o.ncg;
    counts[1] = o.nir;
    counts[2] = o.status;
    window[0] = o.dmin;
    window[1] = o.dmax;
    return 0;
  });
}

int mb_cone_scale(mb_solver* s, const double* p, const double* d, double* spsd) {
  return guard([&]() { return s->S->cone_scale(p, d, spsd) ? 1 : 0; });
}

int64_t mb_cone_blocks_size(const mb_solver* s) { return (int64_t)s->S->cone_blocks_size(); }

int mb_cone_blocks(mb_solver* s, double* out) {
  return guard([&]() {
    s->S->cone_blocks(out);
    return 0;
  });
}

int mb_cone_corr0(mb_solver* s, double sigma_mu, double* r) {
  return guard([&]() {
    s->S->cone_corr0(sigma_mu, r);
    return 0;
  });
}

int mb_cone_corr(mb_solver* s, const double* dp, const double* dd, double sigma_mu, double* r) {
  return guard([&]() {
    s->S->cone_corr(dp, dd, sigma_mu, r);
    return 0;
  });
}

int mb_cone_maxsteps(mb_solver* s, const double* dp, const double* dd, double step_frac, double* step) {
  return guard([&]() {
    *step = s->S->cone_maxsteps(dp, dd, step_frac);
    return 0;
  });
}

int mb_solver_dense_factor(mb_solver* s, double* L) {
  return guard([&]() {
    s->S->dense_factor(L);
    return 0;
  });
}

static int probe_mfma(double* out, bool fp32) {
  return guard([&]() {
    double* d = nullptr;
    if (hipMalloc(&d, 256 * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    mb::launch_probe_mfma(nullptr, d, fp32);
    hipError_t e = hipMemcpy(out, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    return 0;
  });
}

int mb_probe_mfma_f64(double* out) { return probe_mfma(out, false); }
int mb_probe_mfma_f32(double* out) { return probe_mfma(out, true); }

const char* mb_last_error(void) { return g_err.c_str(); }


// ------------------------------------------------------------------------------------------
// serving
// ------------------------------------------------------------------------------------------

struct mb_server {
  std::unique_ptr<mb::Server> S;
};

mb_server* mb_server_create(int32_t nslots, double mem_budget_bytes, int32_t dag_grid) {
  mb_server* out = nullptr;
  guard([&]() {
    auto h = std::make_unique<mb_server>();
    h->S = std::make_unique<mb::Server>(nslots, mem_budget_bytes, dag_grid);
    out = h.release();
    return 0;
  });
  return out;
}

void mb_server_destroy(mb_server* s) { delete s; }

int64_t mb_server_submit(mb_server* s, const mb_problem* p, const mb_settings* st) {
  int64_t id = -1;
  guard([&]() {
    if (!p) throw std::invalid_argument("null problem");
    id = s->S->submit(&p->P, to_settings(st));
    return 0;
  });
  return id;
}

int mb_server_wait(mb_server* s, int64_t ticket, double timeout_s, double* p, double* d, double* y, double* info,
                   double* times) {
  return guard([&]() {
    auto job = s->S->wait(ticket, timeout_s);
    if (!job) return -2;
    if (job->status < 0) throw std::runtime_error(job->error.empty() ? "solve failed" : job->error);
    if (p) std::copy(job->p.begin(), job->p.end(), p);
    if (d) std::copy(job->d.begin(), job->d.end(), d);
    if (y) std::copy(job->y.begin(), job->y.end(), y);
    if (info) std::copy(job->info, job->info + 8, info);
    if (times) {
      times[0] = job->t_submit;
      times[1] = job->t_admit;
      times[2] = job->t_end;
      times[3] = job->bytes;
      times[4] = job->reserve;
    }
    return job->status;
  });
}

int mb_server_release(mb_server* s, int64_t ticket) {
  return guard([&]() {
    s->S->release(ticket);
    return 0;
  });
}

int mb_server_stats(mb_server* s, double* out) {
  return guard([&]() {
    s->S->stats(out);
    return 0;
  });
}

mb_server* mb_server_create_gang(int32_t nslots, double mem_budget_bytes) {
  mb_server* out = nullptr;
  guard([&]() {
    auto h = std::make_unique<mb_server>();
    h->S = std::make_unique<mb::Server>(nslots, mem_budget_bytes, 0, 0, true);
    out = h.release();
    return 0;
  });
  return out;
}

// ------------------------------------------------------------------------------------------
// lockstep batching
// ------------------------------------------------------------------------------------------

struct mb_gang {
  std::unique_ptr<mb::Gang> G;
};

mb_gang* mb_gang_create(int32_t capacity, int32_t dag_grid, int32_t profile) {
  mb_gang* out = nullptr;
  guard([&]() {
    if (capacity < 1 || capacity > mb::GANG_MAX) throw std::invalid_argument("gang capacity must be 1..32");
    auto h = std::make_unique<mb_gang>();
    h->G = std::make_unique<mb::Gang>(capacity, dag_grid, profile != 0);
    out = h.release();
    return 0;
  });
  return out;
}

void mb_gang_destroy(mb_gang* g) { delete g; }

int mb_solver_solve_in(mb_solver* s, mb_gang* g) {
  return guard([&]() { return s->S->solve_in(*g->G); });
}

int mb_solve_batch(mb_solver* const* s, int32_t n, mb_gang* g, int32_t* status) {
  return guard([&]() {
    if (n <= 0) return 0;
    std::atomic<int> next{0};
    std::string first;
    std::mutex emu;
    const int nt = std::min<int>(n, g->G->capacity());
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&]() {
        for (;;) {
          const int i = next++;
          if (i >= n) return;
          try {
            status[i] = s[i]->S->solve_in(*g->G);
          } catch (const std::exception& e) {
            status[i] = -1;
            std::lock_guard<std::mutex> lk(emu);
            if (first.empty()) first = e.what();
          }
        }
      });
    for (auto& t : th) t.join();
    if (!first.empty()) throw std::runtime_error(first);
    return 0;
  });
}

int mb_gang_stats(mb_gang* g, double* out) {
  return guard([&]() {
    mb::GangStats st;
    g->G->stats(st);
    const double v[18] = {st.ticks, st.launches, st.ops, st.combined_ops, st.readbacks, st.joins,
                          st.factor_launches, st.fronts_per_factor, st.factor_flops, st.factor_ms,
                          st.dag_launches, st.dag_flops, st.dag_ms, (double)g->G->members(),
                          st.admit_ms, st.fiber_ms, st.tick_ms, st.wait_ms};
    std::memcpy(out, v, sizeof(v));
    return 0;
  });
}

double mb_clock(void) { return mb::Server::now(); }

int64_t mb_cached_bytes(void) { return (int64_t)mb::cached_memory(); }

int64_t mb_release_cached_memory(void) { return (int64_t)mb::release_cached_memory(); }

int mb_device_info(double* out) {
  return guard([&]() {
    int dev = 0, ncu = 0;
    size_t free_b = 0, total_b = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipMemGetInfo(&free_b, &total_b) != hipSuccess)
      throw std::runtime_error("no HIP device");
    out[0] = ncu;
    out[1] = (double)free_b;
    out[2] = (double)total_b;
    return 0;
  });
}

}  // extern "C"
