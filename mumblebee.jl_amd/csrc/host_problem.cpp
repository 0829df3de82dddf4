// Problem construction on the host: the integer preprocessing of
// src/IPM/src/ipm/problem.jl:84-349 and src/BlockSparseArrays/src/compress.jl, plus the
// block-sparse vertex selections of block_sparse_matrix.jl and the product BᵀB.
#include <algorithm>
#include <cassert>
#include <map>
#include <stdexcept>

#include "host.h"

namespace mb {

CSC csc_transpose(const CSC& A) {
  CSC T;
  T.m = A.n;
  T.n = A.m;
  T.colptr.assign(A.m + 1, 0);
  i64 nz = A.nnz();
  for (i64 p = 0; p < nz; ++p) T.colptr[A.rowval[p] + 1]++;
  for (i64 i = 0; i < A.m; ++i) T.colptr[i + 1] += T.colptr[i];
  T.rowval.resize(nz);
  T.nzval.resize(nz);
  std::vector<i64> nxt(T.colptr.begin(), T.colptr.end() - 1);
  for (i64 j = 0; j < A.n; ++j)
    for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      i64 q = nxt[A.rowval[p]]++;
      T.rowval[q] = j;
      T.nzval[q] = A.nzval[p];
    }
  return T;
}

static CSC dropzeros(const CSC& A) {
  CSC R;
  R.m = A.m;
  R.n = A.n;
  R.colptr.push_back(0);
  for (i64 j = 0; j < A.n; ++j) {
    for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p)
      if (A.nzval[p] != 0.0) {
        R.rowval.push_back(A.rowval[p]);
        R.nzval.push_back(A.nzval[p]);
      }
    R.colptr.push_back((i64)R.rowval.size());
  }
  return R;
}

// problem.jl:133-186 — keep a structural diagonal, drop stored off-diagonal zeros
CSC dropoffdz(const CSC& Q) {
  CSC R;
  R.m = Q.m;
  R.n = Q.n;
  R.colptr.push_back(0);
  for (i64 j = 0; j < Q.n; ++j) {
    i64 iprv = -1;
    for (i64 p = Q.colptr[j]; p < Q.colptr[j + 1]; ++p) {
      i64 i = Q.rowval[p];
      double x = Q.nzval[p];
      if (x != 0.0 && iprv < j && j < i) {
        R.rowval.push_back(j);
        R.nzval.push_back(0.0);
        iprv = j;
      }
      if (x != 0.0 || i == j) {
        R.rowval.push_back(i);
        R.nzval.push_back(x);
        iprv = i;
      }
    }
    if (iprv < j) {
      R.rowval.push_back(j);
      R.nzval.push_back(0.0);
    }
    R.colptr.push_back((i64)R.rowval.size());
  }
  return R;
}

static CSC select_cols(const CSC& A, const std::vector<i64>& cols) {
  CSC R;
  R.m = A.m;
  R.n = (i64)cols.size();
  R.colptr.push_back(0);
  for (i64 j : cols) {
    for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      R.rowval.push_back(A.rowval[p]);
      R.nzval.push_back(A.nzval[p]);
    }
    R.colptr.push_back((i64)R.rowval.size());
  }
  return R;
}

static CSC vcat(const CSC& A, const CSC& B) {
  CSC R;
  R.m = A.m + B.m;
  R.n = A.n;
  R.colptr.push_back(0);
  for (i64 j = 0; j < A.n; ++j) {
    for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      R.rowval.push_back(A.rowval[p]);
      R.nzval.push_back(A.nzval[p]);
    }
    for (i64 p = B.colptr[j]; p < B.colptr[j + 1]; ++p) {
      R.rowval.push_back(B.rowval[p] + A.m);
      R.nzval.push_back(B.nzval[p]);
    }
    R.colptr.push_back((i64)R.rowval.size());
  }
  return R;
}

// A[rowperm, colperm]
static CSC permute(const CSC& A, const std::vector<i64>& rowperm, const std::vector<i64>& colperm) {
  std::vector<i64> inv(A.m);
  for (i64 i = 0; i < A.m; ++i) inv[rowperm[i]] = i;
  CSC R;
  R.m = A.m;
  R.n = (i64)colperm.size();
  R.colptr.push_back(0);
  std::vector<std::pair<i64, double>> tmp;
  for (i64 j : colperm) {
    tmp.clear();
    for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) tmp.emplace_back(inv[A.rowval[p]], A.nzval[p]);
    std::sort(tmp.begin(), tmp.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (auto& t : tmp) {
      R.rowval.push_back(t.first);
      R.nzval.push_back(t.second);
    }
    R.colptr.push_back((i64)R.rowval.size());
  }
  return R;
}

// compress.jl:9-116. Partition refinement with a free list (singly linked, next = var)
// and one doubly linked list per set (head/prev/next), 1-based like the reference so that
// the produced order is the reference's order.
i64 twins_exact(i64 nout, i64 nvtx, const i64* xsrc, const i64* tgt, std::vector<i64>& ptr,
                std::vector<i64>& perm) {
  const i64 N = nout;
  std::vector<i64> var(N + 2, 0), svar(N + 1, 0), flag(N + 1, 0), card(N + 1, 0),
      head(N + 1, 0), prev(N + 1, 0), next(N + 1, 0);
  i64 freeh = 0, nsv = 0;
  if (N > 0) {
    freeh = 1;
    for (i64 i = 1; i < N; ++i) var[i] = i + 1;
    var[N] = 0;
  }
  auto popfree = [&]() { i64 i = freeh; freeh = var[i]; return i; };
  auto pushfree = [&](i64 i) { var[i] = freeh; freeh = i; };
  auto pushfirst = [&](i64 s, i64 i) {
    i64 n = head[s];
    next[i] = n;
    head[s] = i;
    if (n > 0) prev[n] = i;
  };
  auto del = [&](i64 s, i64 i) {
    i64 h = head[s], n = next[i];
    if (i == h) {
      head[s] = n;
    } else {
      i64 p = prev[i];
      next[p] = n;
      if (n > 0) prev[n] = p;
    }
  };
  if (N > 0) {
    i64 s = popfree();
    nsv += 1;
    i64 h = head[s];
    head[s] = 1;
    i64 i = 1;
    for (i64 j = 2; j <= N; ++j) {
      prev[j] = i;
      next[i] = j;
      i = j;
    }
    next[i] = h;
    card[s] = N;
    for (i64 k = 1; k <= N; ++k) svar[k] = s;
  }
  for (i64 j = 1; j <= nvtx; ++j) {
    for (i64 e = xsrc[j - 1]; e < xsrc[j]; ++e) {
      i64 i = tgt[e] + 1;
      i64 s = svar[i];
      if (flag[s] < j) {
        flag[s] = j;
        if (card[s] > 1) {
          del(s, i);
          card[s] -= 1;
          var[s] = i;
          i64 ns = popfree();
          svar[i] = ns;
          nsv += 1;
          pushfirst(ns, i);
          card[ns] += 1;
        }
      } else {
        del(s, i);
        card[s] -= 1;
        i64 k = var[s];
        i64 ns = svar[k];
        svar[i] = ns;
        pushfirst(ns, i);
        card[ns] += 1;
        if (head[s] == 0) {
          pushfree(s);
          nsv -= 1;
        }
      }
    }
  }
  ptr.assign(nsv + 1, 0);
  perm.assign(N, 0);
  i64 t = 0, p = 0;
  ptr[0] = 0;
  for (i64 s = 1; s <= N; ++s) {
    if (head[s] == 0) continue;
    for (i64 i = head[s]; i > 0; i = next[i]) perm[p++] = i - 1;
    ptr[++t] = p;
  }
  return nsv;
}

// compress.jl:147-265 — cosine-similarity grouping (tau < 1)
i64 twins_tau(i64 nout, i64 nvtx, const i64* xrow, const i64* cadj, const i64* xcol,
              const i64* radj, double tau, std::vector<i64>& ptr, std::vector<i64>& perm) {
  (void)nvtx;
  std::vector<i64> adjmap(nout, 0), cosine(nout, 0), degree(nout);
  for (i64 i = 0; i < nout; ++i) degree[i] = xrow[i + 1] - xrow[i];
  // doubly linked list with pushfirst / delete semantics; iteration follows next links
  std::vector<i64> lprev(nout, -1), lnext(nout, -1);
  i64 lhead = -1;
  i64 nb = 0;
  for (i64 i = 0; i < nout; ++i) {
    if (adjmap[i] != 0) continue;
    adjmap[i] = ++nb;
    for (i64 e = xrow[i]; e < xrow[i + 1]; ++e) {
      i64 j = cadj[e];
      for (i64 f = xcol[j + 1] - 1; f >= xcol[j]; --f) {
        i64 k = radj[f];
        if (!(k > i)) break;
        if (adjmap[k] == 0) {
          if (cosine[k] == 0) {
            lnext[k] = lhead;
            if (lhead >= 0) lprev[lhead] = k;
            lhead = k;
          }
          cosine[k] += 1;
        }
      }
    }
    for (i64 k = lhead; k >= 0;) {
      i64 kn = lnext[k];
      double cs = (double)cosine[k];
      if (cs * cs >= tau * tau * (double)degree[i] * (double)degree[k]) adjmap[k] = nb;
      cosine[k] = 0;
      k = kn;
    }
    lhead = -1;
  }
  ptr.assign(nb + 1, 0);
  std::vector<i64> cnt(nb + 2, 0);
  for (i64 i = 0; i < nout; ++i) cnt[adjmap[i] + 1]++;
  cnt[1] = 0;
  for (i64 t = 1; t <= nb; ++t) cnt[t + 1] += cnt[t];
  for (i64 t = 0; t <= nb; ++t) ptr[t] = cnt[t + 1];
  perm.assign(nout, 0);
  std::vector<i64> nxt(ptr.begin(), ptr.end());
  for (i64 i = 0; i < nout; ++i) perm[nxt[adjmap[i] - 1]++] = i;
  return nb;
}

// twins(transpose(X), X, tau): group columns of X by row pattern
static i64 twins_of_cols(const CSC& X, double tau, std::vector<i64>& ptr, std::vector<i64>& perm) {
  CSC Xt = csc_transpose(X);
  if (tau == 1.0) return twins_exact(X.n, X.m, Xt.colptr.data(), Xt.rowval.data(), ptr, perm);
  return twins_tau(X.n, X.m, X.colptr.data(), X.rowval.data(), Xt.colptr.data(), Xt.rowval.data(),
                   tau, ptr, perm);
}

// twins(A, transpose(A), tau): group rows of A by column pattern
static i64 twins_of_rows(const CSC& A, double tau, std::vector<i64>& ptr, std::vector<i64>& perm) {
  if (tau == 1.0) return twins_exact(A.m, A.n, A.colptr.data(), A.rowval.data(), ptr, perm);
  CSC At = csc_transpose(A);
  return twins_tau(A.m, A.n, At.colptr.data(), At.rowval.data(), A.colptr.data(), A.rowval.data(),
                   tau, ptr, perm);
}

// compress.jl:359-483
static BSM compress2(const CSC& A, i64 nvtx, i64 nout, const std::vector<i64>& xcol,
                     const std::vector<i64>& xrow) {
  std::vector<i64> row(A.m);
  for (i64 u = 0; u < nout; ++u)
    for (i64 i = xrow[u]; i < xrow[u + 1]; ++i) row[i] = u;
  BSM R;
  R.nout = nout;
  R.nvtx = nvtx;
  R.xcol = xcol;
  R.xrow = xrow;
  R.xsrc.push_back(0);
  R.xblk.push_back(0);
  std::vector<i64> mark(nout, -1), outs;
  for (i64 v = 0; v < nvtx; ++v) {
    i64 vn = xcol[v + 1] - xcol[v];
    outs.clear();
    for (i64 j = xcol[v]; j < xcol[v + 1]; ++j)
      for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
        i64 u = row[A.rowval[p]];
        if (mark[u] != v) {
          mark[u] = v;
          outs.push_back(u);
        }
      }
    std::sort(outs.begin(), outs.end());
    for (i64 u : outs) {
      R.tgt.push_back(u);
      R.xblk.push_back(R.xblk.back() + (xrow[u + 1] - xrow[u]) * vn);
    }
    R.xsrc.push_back((i64)R.tgt.size());
  }
  R.val.assign(R.xblk.back(), 0.0);
  for (i64 v = 0; v < nvtx; ++v) {
    for (i64 e = R.xsrc[v]; e < R.xsrc[v + 1]; ++e) {
      i64 u = R.tgt[e];
      i64 nr = xrow[u + 1] - xrow[u];
      for (i64 j = xcol[v]; j < xcol[v + 1]; ++j)
        for (i64 p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
          i64 i = A.rowval[p];
          if (i >= xrow[u] && i < xrow[u + 1])
            R.val[R.xblk[e] + (j - xcol[v]) * nr + (i - xrow[u])] = A.nzval[p];
        }
    }
  }
  return R;
}

Problem make_problem(const CSC& Q0, const CSC& B0, const std::vector<double>& f,
                     const std::vector<double>& g, double mu, const std::vector<Cone>& K,
                     const std::vector<i64>& s, double tau) {
  const i64 m = B0.m, n = B0.n;
  if (Q0.m != n || Q0.n != n || (i64)f.size() != n || (i64)g.size() != m || s.size() != K.size())
    throw std::invalid_argument("IPMProblem: inconsistent dimensions");
  i64 tot = 0;
  for (i64 x : s) tot += x;
  if (tot != n) throw std::invalid_argument("IPMProblem: cone sizes do not sum to n");
  if (!(tau > 0.0 && tau <= 1.0)) throw std::invalid_argument("IPMProblem: compress must be in (0, 1]");
  CSC Q = dropoffdz(Q0);
  CSC B = dropzeros(B0);
  // colpartition (problem.jl:247-349)
  i64 ncolP = 0, ncolC = 0, nvtxR = 0;
  for (size_t k = 0; k < K.size(); ++k) {
    if (K[k].kind == CONE_POS) ncolP += s[k];
    else if (K[k].kind == CONE_FREE) ncolC += s[k];
    else nvtxR += 1;
  }
  std::vector<i64> work(n);
  {
    i64 p = 0, c = 0, r = 0, j = 0;
    for (size_t k = 0; k < K.size(); ++k)
      for (i64 t = 0; t < s[k]; ++t, ++j) {
        if (K[k].kind == CONE_POS) work[p++] = j;
        else if (K[k].kind == CONE_FREE) work[ncolP + c++] = j;
        else work[ncolP + ncolC + r++] = j;
      }
  }
  std::vector<i64> indP(work.begin(), work.begin() + ncolP);
  std::vector<i64> indC(work.begin() + ncolP, work.begin() + ncolP + ncolC);
  CSC XP = vcat(select_cols(Q, indP), select_cols(B, indP));
  CSC XC = vcat(select_cols(Q, indC), select_cols(B, indC));
  std::vector<i64> xcolP, permP, xcolC, permC;
  i64 nvtxP = twins_of_cols(XP, tau, xcolP, permP);
  i64 nvtxC = twins_of_cols(XC, tau, xcolC, permC);
  i64 nvtx = nvtxP + nvtxC + nvtxR;
  std::vector<i64> colperm(n), xcol(nvtx + 1);
  std::vector<Cone> cones(nvtx);
  for (i64 j = 0; j < ncolP; ++j) colperm[j] = work[permP[j]];
  for (i64 v = 0; v < nvtxP; ++v) {
    xcol[v] = xcolP[v];
    cones[v].kind = CONE_POS;
  }
  for (i64 j = 0; j < ncolC; ++j) colperm[ncolP + j] = work[ncolP + permC[j]];
  for (i64 v = 0; v < nvtxC; ++v) {
    xcol[nvtxP + v] = ncolP + xcolC[v];
    cones[nvtxP + v].kind = CONE_FREE;
  }
  for (i64 j = ncolP + ncolC; j < n; ++j) colperm[j] = work[j];
  {
    i64 r = 0, j = 0;
    for (size_t k = 0; k < K.size(); ++k)
      if (K[k].kind != CONE_POS && K[k].kind != CONE_FREE) {
        xcol[nvtxP + nvtxC + r] = ncolP + ncolC + j;
        cones[nvtxP + nvtxC + r] = K[k];
        ++r;
        j += s[k];
      }
  }
  xcol[nvtx] = n;
  // colcompress (problem.jl:188-245)
  CSC Bc;
  Bc.m = m;
  Bc.n = nvtx;
  Bc.colptr.push_back(0);
  {
    std::vector<i64> mark(m, -1), rows;
    for (i64 v = 0; v < nvtx; ++v) {
      rows.clear();
      for (i64 j = xcol[v]; j < xcol[v + 1]; ++j) {
        i64 k = colperm[j];
        for (i64 p = B.colptr[k]; p < B.colptr[k + 1]; ++p) {
          i64 i = B.rowval[p];
          if (mark[i] != v) {
            mark[i] = v;
            rows.push_back(i);
          }
        }
      }
      std::sort(rows.begin(), rows.end());
      for (i64 i : rows) {
        Bc.rowval.push_back(i);
        Bc.nzval.push_back(1.0);
      }
      Bc.colptr.push_back((i64)Bc.rowval.size());
    }
  }
  std::vector<i64> xrow, rowperm;
  i64 nout = twins_of_rows(Bc, tau, xrow, rowperm);
  CSC Bp = permute(B, rowperm, colperm);
  CSC Qp = permute(Q, colperm, colperm);
  Problem P;
  P.B = compress2(Bp, nvtx, nout, xcol, xrow);
  P.Q = compress2(Qp, nvtx, nvtx, xcol, xcol);
  P.f.resize(n);
  P.g.resize(m);
  for (i64 j = 0; j < n; ++j) P.f[j] = f[colperm[j]];
  for (i64 i = 0; i < m; ++i) P.g[i] = g[rowperm[i]];
  P.mu = mu;
  P.K = cones;
  P.P1 = rowperm;
  P.P2 = colperm;
  return P;
}

// block_sparse_matrix.jl:1111-1169
BSM selectvtxs(const BSM& A, const std::vector<i64>& sel) {
  BSM R;
  R.nout = A.nout;
  R.nvtx = (i64)sel.size();
  R.xrow = A.xrow;
  R.xcol.push_back(0);
  R.xsrc.push_back(0);
  R.xblk.push_back(0);
  R.xcol.reserve(sel.size() + 1);
  R.xsrc.reserve(sel.size() + 1);
  R.tgt.reserve(A.tgt.size());
  R.xblk.reserve(A.xblk.size());
  R.val.resize(A.val.size());
  i64 nv = 0;
  for (i64 av : sel) {
    R.xcol.push_back(R.xcol.back() + A.ncols(av));
    for (i64 e = A.xsrc[av]; e < A.xsrc[av + 1]; ++e) {
      R.tgt.push_back(A.tgt[e]);
      std::copy(A.val.begin() + A.xblk[e], A.val.begin() + A.xblk[e + 1], R.val.begin() + nv);
      nv += A.xblk[e + 1] - A.xblk[e];
      R.xblk.push_back(nv);
    }
    R.xsrc.push_back((i64)R.tgt.size());
  }
  R.val.resize(nv);
  return R;
}

// halfselectvtxs ∘ halfselectvtxs (block_sparse_matrix.jl:787-874; kkt.jl:38)
BSM symselectvtxs(const BSM& A, const std::vector<i64>& sel) {
  const i64 N = (i64)sel.size();
  std::vector<i64> inv(N);
  for (i64 k = 0; k < N; ++k) inv[sel[k]] = k;
  BSM R;
  R.nout = R.nvtx = N;
  R.xcol.push_back(0);
  for (i64 av : sel) R.xcol.push_back(R.xcol.back() + A.ncols(av));
  R.xrow = R.xcol;
  R.xsrc.push_back(0);
  R.xblk.push_back(0);
  R.xsrc.reserve(N + 1);
  R.tgt.reserve(A.tgt.size());
  R.xblk.reserve(A.xblk.size());
  R.val.resize(A.val.size());
  i64 nv = 0;
  std::vector<std::pair<i64, i64>> items;
  for (i64 av : sel) {
    items.clear();
    for (i64 e = A.xsrc[av]; e < A.xsrc[av + 1]; ++e) items.emplace_back(inv[A.tgt[e]], e);
    std::sort(items.begin(), items.end());
    for (auto& it : items) {
      i64 e = it.second;
      R.tgt.push_back(it.first);
      std::copy(A.val.begin() + A.xblk[e], A.val.begin() + A.xblk[e + 1], R.val.begin() + nv);
      nv += A.xblk[e + 1] - A.xblk[e];
      R.xblk.push_back(nv);
    }
    R.xsrc.push_back((i64)R.tgt.size());
  }
  R.val.resize(nv);
  return R;
}

// Bᵀ as a block-sparse matrix (blocks transposed), arcs sorted
BSM transpose_bsm(const BSM& A) {
  BSM R;
  R.nout = A.nvtx;
  R.nvtx = A.nout;
  R.xcol = A.xrow;
  R.xrow = A.xcol;
  std::vector<std::vector<std::pair<i64, i64>>> lists(A.nout);
  for (i64 v = 0; v < A.nvtx; ++v)
    for (i64 e = A.xsrc[v]; e < A.xsrc[v + 1]; ++e) lists[A.tgt[e]].emplace_back(v, e);
  R.xsrc.push_back(0);
  R.xblk.push_back(0);
  for (i64 u = 0; u < A.nout; ++u) {
    i64 nr = A.nrows(u);
    for (auto& it : lists[u]) {
      i64 v = it.first, e = it.second;
      i64 nc = A.ncols(v);
      R.tgt.push_back(v);
      // block of Aᵀ at (v, u): nc × nr column-major
      for (i64 i = 0; i < nr; ++i)
        for (i64 j = 0; j < nc; ++j) R.val.push_back(A.val[A.xblk[e] + j * nr + i]);
      R.xblk.push_back((i64)R.val.size());
    }
    R.xsrc.push_back((i64)R.tgt.size());
  }
  return R;
}

// uzawa.jl:85 (`B' * B`), pattern as matprod_dest (block_sparse_matrix.jl:1171-1252). For each
// vertex v: its neighbours w (vertices sharing an out vertex with v) in increasing order, each
// block Σ_rows B_wᵀ B_v accumulated in arc order straight into the result.
void gram_lower(i64 n, const std::vector<int>& rp, const std::vector<int>& rc, const std::vector<double>& rv,
                const std::vector<int>& cp, const std::vector<int>& cr, const std::vector<double>& cv,
                std::vector<i64>& colp, std::vector<int>& rows, std::vector<double>& vals) {
  colp.assign(n + 1, 0);
  rows.clear();
  vals.clear();
  std::vector<i64> mark(n, -1);
  std::vector<double> acc(n, 0.0);
  std::vector<int> pat;
  // the pattern of a column comes out sorted by a bitset scan where n is small, else by a sort
  const bool bits = n <= (1 << 14);
  std::vector<uint64_t> row(bits ? (n + 63) / 64 : 0, 0);
  for (i64 j = 0; j < n; ++j) {
    pat.clear();
    for (int t = cp[j]; t < cp[j + 1]; ++t) {
      const int r = cr[t];
      const double bj = cv[t];
      // the row's columns are sorted: start at the first one ≥ j
      const int* c0 = rc.data() + rp[r];
      const int* c1 = rc.data() + rp[r + 1];
      for (const int* q = std::lower_bound(c0, c1, (int)j); q < c1; ++q) {
        const int i = *q;
        if (mark[i] != j) {
          mark[i] = j;
          acc[i] = 0.0;
          pat.push_back(i);
        }
        acc[i] += rv[q - rc.data()] * bj;
      }
    }
    if (bits) {
      for (int i : pat) row[i >> 6] |= 1ull << (i & 63);
      const size_t np = pat.size();
      pat.clear();
      for (size_t w = (size_t)(j >> 6); pat.size() < np; ++w)
        for (uint64_t x = row[w]; x; x &= x - 1) pat.push_back((int)(w * 64 + __builtin_ctzll(x)));
      for (int i : pat) row[i >> 6] = 0;
    } else {
      std::sort(pat.begin(), pat.end());
    }
    for (int i : pat) {
      rows.push_back(i);
      vals.push_back(acc[i]);
    }
    colp[j + 1] = (i64)rows.size();
  }
}

}  // namespace mb
