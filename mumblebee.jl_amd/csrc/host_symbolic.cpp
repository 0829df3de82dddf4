// Symbolic analysis of δH + BᵀB on the cone-block vertex graph.
//
// The vertex graph restates `weightedgraph` (src/IPM/src/utils.jl:135-145): vertex weights
// are block widths, edges are pattern(Q) ∪ linegraph(B). The reference obtains the ordering,
// elimination tree and fronts from CliqueTrees.jl `symbolic` (kkt.jl:36), whose source is not
// available; this file defines the build's own: weighted exact-external-degree minimum degree
// on the quotient graph (ties to the smallest vertex), elimination-tree postorder, fundamental
// supernodes, then relaxed amalgamation so that fronts are dense enough for MFMA tiles.
#include <algorithm>
#include <cstdint>
#include <queue>
#include <set>
#include <stdexcept>

#include "host.h"

namespace mb {

void vertex_graph(const BSM& B, const BSM& Q, std::vector<i64>& weights,
                  std::vector<std::vector<i64>>& adj) {
  const i64 N = B.nvtx, U = B.nout;
  weights.resize(N);
  for (i64 v = 0; v < N; ++v) weights[v] = B.ncols(v);
  // the vertices of each out vertex (row block) of B, and the transposed pattern of Q (so that
  // the graph is the symmetric closure whatever Q stores), as compressed lists
  std::vector<i64> rp(U + 1, 0), rv(B.narc());
  for (i64 e = 0; e < B.narc(); ++e) rp[B.tgt[e] + 1]++;
  for (i64 u = 0; u < U; ++u) rp[u + 1] += rp[u];
  {
    std::vector<i64> nx(rp.begin(), rp.end() - 1);
    for (i64 v = 0; v < N; ++v)
      for (i64 e = B.xsrc[v]; e < B.xsrc[v + 1]; ++e) rv[nx[B.tgt[e]]++] = v;
  }
  std::vector<i64> qp(N + 1, 0), qv(Q.narc());
  for (i64 e = 0; e < Q.narc(); ++e) qp[Q.tgt[e] + 1]++;
  for (i64 u = 0; u < N; ++u) qp[u + 1] += qp[u];
  {
    std::vector<i64> nx(qp.begin(), qp.end() - 1);
    for (i64 v = 0; v < Q.nvtx; ++v)
      for (i64 e = Q.xsrc[v]; e < Q.xsrc[v + 1]; ++e) qv[nx[Q.tgt[e]]++] = v;
  }
  std::vector<i64> mark(N, -1);
  adj.assign(N, {});
  for (i64 v = 0; v < N; ++v) {
    std::vector<i64>& a = adj[v];
    mark[v] = v;  // no self loop
    auto add = [&](i64 u) {
      if (mark[u] != v) {
        mark[u] = v;
        a.push_back(u);
      }
    };
    if (v < Q.nvtx)
      for (i64 e = Q.xsrc[v]; e < Q.xsrc[v + 1]; ++e) add(Q.tgt[e]);
    for (i64 k = qp[v]; k < qp[v + 1]; ++k) add(qv[k]);
    for (i64 e = B.xsrc[v]; e < B.xsrc[v + 1]; ++e)
      for (i64 k = rp[B.tgt[e]]; k < rp[B.tgt[e] + 1]; ++k) add(rv[k]);
    std::sort(a.begin(), a.end());
  }
}

namespace {

struct MinDegree {
  i64 N;
  const std::vector<i64>& w;
  std::vector<std::vector<i64>> A, E, L;  // L indexed by element (= eliminated vertex)
  std::vector<i64> mark, emark;
  i64 stamp = 0, estamp = 0;
  std::vector<i64> deg;
  std::vector<char> alive;

  MinDegree(const std::vector<i64>& weights, const std::vector<std::vector<i64>>& adj)
      : N((i64)weights.size()), w(weights), A(adj), E(weights.size()), L(weights.size()),
        mark(weights.size(), 0), emark(weights.size(), 0), deg(weights.size(), 0),
        alive(weights.size(), 1) {}

  // reach(i) collected into out (excluding i); returns weighted size
  i64 reach(i64 i, std::vector<i64>* out) {
    ++stamp;
    mark[i] = stamp;
    i64 d = 0;
    for (i64 j : A[i])
      if (mark[j] != stamp) {
        mark[j] = stamp;
        d += w[j];
        if (out) out->push_back(j);
      }
    for (i64 e : E[i])
      for (i64 j : L[e])
        if (mark[j] != stamp) {
          mark[j] = stamp;
          d += w[j];
          if (out) out->push_back(j);
        }
    return d;
  }

  void run(std::vector<i64>& order, std::vector<std::vector<i64>>& strct) {
    order.clear();
    strct.assign(N, {});
    i64 W = 0;
    for (i64 x : w) W += x;
    std::set<std::pair<i64, i64>> pq;
    for (i64 i = 0; i < N; ++i) {
      deg[i] = reach(i, nullptr);
      pq.insert({deg[i], i});
    }
    std::vector<i64> Lp;
    while (!pq.empty()) {
      i64 p = pq.begin()->second;
      if (deg[p] + w[p] == W) {
        // p reaches every remaining variable: after it the rest is a clique, where minimum
        // degree picks by (-weight, index)
        pq.erase(pq.begin());
        std::vector<i64> rest;
        for (auto& it : pq) rest.push_back(it.second);
        std::sort(rest.begin(), rest.end());
        order.push_back(p);
        strct[p] = rest;
        std::stable_sort(rest.begin(), rest.end(), [&](i64 a, i64 b) {
          if (w[a] != w[b]) return w[a] > w[b];
          return a < b;
        });
        for (size_t k = 0; k < rest.size(); ++k) {
          order.push_back(rest[k]);
          std::vector<i64> s(rest.begin() + k + 1, rest.end());
          std::sort(s.begin(), s.end());
          strct[rest[k]] = s;
        }
        break;
      }
      pq.erase(pq.begin());
      Lp.clear();
      reach(p, &Lp);
      std::sort(Lp.begin(), Lp.end());
      strct[p] = Lp;
      // mark L_p members and E_p elements
      ++stamp;
      for (i64 j : Lp) mark[j] = stamp;
      mark[p] = stamp;
      ++estamp;
      for (i64 e : E[p]) emark[e] = estamp;
      for (i64 i : Lp) {
        auto& Ai = A[i];
        Ai.erase(std::remove_if(Ai.begin(), Ai.end(), [&](i64 j) { return mark[j] == stamp; }), Ai.end());
        auto& Ei = E[i];
        Ei.erase(std::remove_if(Ei.begin(), Ei.end(), [&](i64 e) { return emark[e] == estamp; }), Ei.end());
        Ei.push_back(p);
      }
      for (i64 e : E[p]) std::vector<i64>().swap(L[e]);
      L[p] = Lp;
      std::vector<i64>().swap(A[p]);
      std::vector<i64>().swap(E[p]);
      alive[p] = 0;
      W -= w[p];
      order.push_back(p);
      for (i64 i : Lp) {
        pq.erase({deg[i], i});
        deg[i] = reach(i, nullptr);
        pq.insert({deg[i], i});
      }
    }
  }
};

// The same ordering with the reach sets kept explicitly as bitsets. Eliminating p changes the
// reach of exactly the vertices i ∈ L_p = reach(p), to (reach(i) ∪ L_p) \ {i, p}; every other
// reach is unchanged (a vertex outside L_p shares no element or edge with p). So each exact
// external degree is updated by the weights of the few new members instead of recomputing a
// union over the quotient graph, and the elimination decisions are identical.
// r |= lp over nw words; returns the number of bits lp adds to r. The popcount instruction is
// used where the host CPU has it (checked once at run time), the portable sequence elsewhere.
template <int>
static inline __attribute__((always_inline)) i64 merge_count_body(uint64_t* r, const uint64_t* lp, i64 nw) {
  i64 c = 0;
  for (i64 k = 0; k < nw; ++k) {
    c += __builtin_popcountll(lp[k] & ~r[k]);
    r[k] |= lp[k];
  }
  return c;
}
__attribute__((target("popcnt"))) static i64 merge_count_hw(uint64_t* r, const uint64_t* lp, i64 nw) {
  return merge_count_body<1>(r, lp, nw);
}
static i64 merge_count_sw(uint64_t* r, const uint64_t* lp, i64 nw) { return merge_count_body<0>(r, lp, nw); }
static i64 merge_count(uint64_t* r, const uint64_t* lp, i64 nw) {
  static const bool hw = __builtin_cpu_supports("popcnt");
  return hw ? merge_count_hw(r, lp, nw) : merge_count_sw(r, lp, nw);
}

struct MinDegreeBits {
  i64 N, nw;
  const std::vector<i64>& w;
  std::vector<uint64_t> R;  // N rows of nw words
  std::vector<i64> deg;

  i64 w0 = 0;  // the common weight when every vertex has it (else 0): degrees by popcount

  MinDegreeBits(const std::vector<i64>& weights, const std::vector<std::vector<i64>>& adj)
      : N((i64)weights.size()), nw(((i64)weights.size() + 63) / 64), w(weights),
        R((size_t)((i64)weights.size() * (((i64)weights.size() + 63) / 64)), 0), deg(weights.size(), 0) {
    for (i64 i = 0; i < N; ++i)
      for (i64 j : adj[i]) {
        R[(size_t)(i * nw + j / 64)] |= 1ull << (j % 64);
        deg[i] += w[j];
      }
    set_uniform();
  }

  // the vertex graph of vertex_graph (pattern(Q) ∪ pattern(Qᵀ) ∪ linegraph(B), no self loops)
  // set straight into the reach bitsets
  MinDegreeBits(const std::vector<i64>& weights, const BSM& B, const BSM& Q)
      : N((i64)weights.size()), nw(((i64)weights.size() + 63) / 64), w(weights),
        R((size_t)((i64)weights.size() * (((i64)weights.size() + 63) / 64)), 0), deg(weights.size(), 0) {
    auto set = [&](i64 i, i64 j) { R[(size_t)(i * nw + j / 64)] |= 1ull << (j % 64); };
    for (i64 v = 0; v < Q.nvtx; ++v)
      for (i64 e = Q.xsrc[v]; e < Q.xsrc[v + 1]; ++e) {
        set(v, Q.tgt[e]);
        set(Q.tgt[e], v);
      }
    std::vector<i64> rp(B.nout + 1, 0), rv(B.narc());
    for (i64 e = 0; e < B.narc(); ++e) rp[B.tgt[e] + 1]++;
    for (i64 u = 0; u < B.nout; ++u) rp[u + 1] += rp[u];
    std::vector<i64> nx(rp.begin(), rp.end() - 1);
    for (i64 v = 0; v < B.nvtx; ++v)
      for (i64 e = B.xsrc[v]; e < B.xsrc[v + 1]; ++e) rv[nx[B.tgt[e]]++] = v;
    for (i64 u = 0; u < B.nout; ++u)
      for (i64 a = rp[u]; a < rp[u + 1]; ++a)
        for (i64 b = rp[u]; b < rp[u + 1]; ++b) set(rv[a], rv[b]);
    for (i64 i = 0; i < N; ++i) R[(size_t)(i * nw + i / 64)] &= ~(1ull << (i % 64));
    set_uniform();
    for (i64 i = 0; i < N; ++i) {
      const uint64_t* r = &R[(size_t)(i * nw)];
      i64 d = 0;
      for (i64 k = 0; k < nw; ++k) {
        if (w0) d += w0 * __builtin_popcountll(r[k]);
        else
          for (uint64_t b = r[k]; b; b &= b - 1) d += w[k * 64 + __builtin_ctzll(b)];
      }
      deg[i] = d;
    }
  }

  void set_uniform() {
    w0 = N > 0 ? w[0] : 0;
    for (i64 x : w)
      if (x != w0) w0 = 0;
  }

  void members(i64 i, std::vector<i64>& out) const {
    out.clear();
    const uint64_t* r = &R[(size_t)(i * nw)];
    for (i64 k = 0; k < nw; ++k)
      for (uint64_t b = r[k]; b; b &= b - 1) out.push_back(k * 64 + __builtin_ctzll(b));
  }

  // tail: when not null, the vertices from the first one that reaches every remaining vertex
  // onwards form a clique eliminated in order; their structures (each one's later vertices) are
  // left empty and their count is returned here instead
  void run(std::vector<i64>& order, std::vector<std::vector<i64>>& strct, i64* tail = nullptr) {
    order.clear();
    strct.assign(N, {});
    if (tail) *tail = 0;
    i64 W = 0;
    for (i64 x : w) W += x;
    // lazy min-heap of (degree, vertex): an entry is current while its degree is the vertex's
    // and the vertex is not eliminated, so the first current entry is the set minimum that an
    // ordered set would give (ties to the smallest vertex)
    typedef std::pair<i64, i64> Item;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    std::vector<char> gone(N, 0);
    for (i64 i = 0; i < N; ++i) pq.push({deg[i], i});
    std::vector<i64> Lp;
    std::vector<uint64_t> lp(nw);
    for (;;) {
      while (!pq.empty() && (gone[pq.top().second] || pq.top().first != deg[pq.top().second])) pq.pop();
      if (pq.empty()) break;
      i64 p = pq.top().second;
      pq.pop();
      gone[p] = 1;
      members(p, Lp);  // sorted
      if (deg[p] + w[p] == W) {
        // p reaches every remaining variable: after it the rest is a clique, where minimum
        // degree picks by (-weight, index)
        order.push_back(p);
        std::vector<i64> rest = Lp;
        std::stable_sort(rest.begin(), rest.end(), [&](i64 a, i64 b) {
          if (w[a] != w[b]) return w[a] > w[b];
          return a < b;
        });
        for (i64 v : rest) order.push_back(v);
        if (tail) {
          *tail = (i64)rest.size() + 1;
          break;
        }
        strct[p] = Lp;
        const bool sorted = std::is_sorted(rest.begin(), rest.end());
        for (size_t k = 0; k < rest.size(); ++k) {
          std::vector<i64> s(rest.begin() + k + 1, rest.end());
          if (!sorted) std::sort(s.begin(), s.end());
          strct[rest[k]] = std::move(s);
        }
        break;
      }
      strct[p] = Lp;
      order.push_back(p);
      W -= w[p];
      std::copy(&R[(size_t)(p * nw)], &R[(size_t)(p * nw)] + nw, lp.begin());
      for (i64 i : Lp) {
        uint64_t* r = &R[(size_t)(i * nw)];
        // i joins p's reach set: every vertex of lp that i did not reach yet adds its weight. i
        // itself is in lp and not in r (no self loops), so it is counted once and taken off
        i64 d = deg[i] - w[p];
        if (w0) {
          d += w0 * (merge_count(r, lp.data(), nw) - 1);
        } else {
          for (i64 k = 0; k < nw; ++k) {
            for (uint64_t b = lp[k] & ~r[k]; b; b &= b - 1) d += w[k * 64 + __builtin_ctzll(b)];
            r[k] |= lp[k];
          }
          d -= w[i];
        }
        r[i / 64] &= ~(1ull << (i % 64));
        r[p / 64] &= ~(1ull << (p % 64));
        if (d != deg[i]) {
          deg[i] = d;
          pq.push({d, i});
        }
      }
    }
  }
};

}  // namespace

static Symbolic finish(const std::vector<i64>& weights, std::vector<i64>& order,
                       std::vector<std::vector<i64>>& strct, bool amalgamate, i64 tail = 0);

Symbolic analyse(const std::vector<i64>& weights, const std::vector<std::vector<i64>>& adj,
                 bool amalgamate) {
  const i64 N = (i64)weights.size();
  std::vector<i64> order;
  std::vector<std::vector<i64>> strct;
  if (N <= 16384) {  // N² bits stay within 32 MB
    MinDegreeBits md(weights, adj);
    md.run(order, strct);
  } else {
    MinDegree md(weights, adj);
    md.run(order, strct);
  }
  return finish(weights, order, strct, amalgamate);
}

Symbolic analyse_kkt(const BSM& B, const BSM& Q, bool amalgamate) {
  const i64 N = B.nvtx;
  if (N > 16384) {
    std::vector<i64> w;
    std::vector<std::vector<i64>> adj;
    vertex_graph(B, Q, w, adj);
    return analyse(w, adj, amalgamate);
  }
  std::vector<i64> weights(N);
  for (i64 v = 0; v < N; ++v) weights[v] = B.ncols(v);
  std::vector<i64> order;
  std::vector<std::vector<i64>> strct;
  MinDegreeBits md(weights, B, Q);
  i64 tail = 0;
  md.run(order, strct, &tail);
  return finish(weights, order, strct, amalgamate, tail);
}

// etree, postorder, fundamental supernodes and relaxed amalgamation of an elimination order with
// its structures (strct[v]: the later neighbours of v, sorted). The last `tail` vertices of the
// order form a clique eliminated in that order: their structures are implicit (each one's later
// vertices) and strct holds nothing for them.
static Symbolic finish(const std::vector<i64>& weights, std::vector<i64>& order,
                       std::vector<std::vector<i64>>& strct, bool amalgamate, i64 tail) {
  const i64 N = (i64)weights.size();
  const i64 T0 = N - tail;  // first position of the clique tail
  std::vector<i64> pos(N);
  for (i64 k = 0; k < N; ++k) pos[order[k]] = k;
  std::vector<i64> parent(N, -1);
  for (i64 k = 0; k < N; ++k) {
    const i64 v = order[k];
    if (k >= T0) {
      parent[v] = k + 1 < N ? order[k + 1] : -1;
      continue;
    }
    i64 best = -1;
    for (i64 u : strct[v])
      if (best < 0 || pos[u] < pos[best]) best = u;
    parent[v] = best;
  }
  std::vector<std::vector<i64>> children(N);
  std::vector<i64> roots;
  for (i64 k = 0; k < N; ++k) {
    i64 v = order[k];
    if (parent[v] >= 0) children[parent[v]].push_back(v);
    else roots.push_back(v);
  }
  std::vector<i64> post;
  post.reserve(N);
  for (i64 r : roots) {
    std::vector<std::pair<i64, size_t>> st{{r, 0}};
    while (!st.empty()) {
      auto& top = st.back();
      i64 v = top.first;
      if (top.second < children[v].size()) {
        i64 c = children[v][top.second++];
        st.push_back({c, 0});
      } else {
        post.push_back(v);
        st.pop_back();
      }
    }
  }
  std::vector<i64> newpos(N);
  for (i64 k = 0; k < N; ++k) newpos[post[k]] = k;
  // the clique tail is a chain whose every vertex is its parent's last child, so the postorder
  // ends with it, in elimination order: its structures are the position ranges (k, N)
  for (i64 k = T0; k < N; ++k)
    if (newpos[order[k]] != k) throw std::logic_error("symbolic: clique tail out of place in the postorder");
  std::vector<i64> npar(N), nchild(N, 0);
  std::vector<std::vector<i64>> nstr(N);
  for (i64 k = 0; k < N; ++k) {
    i64 v = post[k];
    npar[k] = parent[v] >= 0 ? newpos[parent[v]] : -1;
    if (k >= T0) continue;
    nstr[k].reserve(strct[v].size());
    for (i64 u : strct[v]) nstr[k].push_back(newpos[u]);
    if (!std::is_sorted(nstr[k].begin(), nstr[k].end())) std::sort(nstr[k].begin(), nstr[k].end());
  }
  auto nsize = [&](i64 k) { return k >= T0 ? N - 1 - k : (i64)nstr[k].size(); };
  for (i64 k = 0; k < N; ++k)
    if (npar[k] >= 0) nchild[npar[k]]++;
  // fundamental supernodes
  std::vector<i64> fv{0};
  for (i64 k = 0; k + 1 < N; ++k) {
    bool merge = npar[k] == k + 1 && nchild[k + 1] == 1 && nsize(k) == nsize(k + 1) + 1;
    if (!merge) fv.push_back(k + 1);
  }
  if (N > 0) fv.push_back(N);
  const i64 F = (i64)fv.size() - 1;
  std::vector<i64> vfront(N);
  for (i64 f = 0; f < F; ++f)
    for (i64 k = fv[f]; k < fv[f + 1]; ++k) vfront[k] = f;
  std::vector<i64> wpost(N), xc(N + 1, 0);
  for (i64 k = 0; k < N; ++k) {
    wpost[k] = weights[post[k]];
    xc[k + 1] = xc[k] + wpost[k];
  }
  std::vector<i64> fpar(F), r(F), s(F);
  std::vector<std::vector<i64>> fsep(F);
  double flops = 0.0;
  for (i64 f = 0; f < F; ++f) {
    const i64 last = fv[f + 1] - 1;
    if (last >= T0) {
      fsep[f].resize(N - 1 - last);
      for (i64 u = last + 1; u < N; ++u) fsep[f][u - last - 1] = u;
    } else {
      fsep[f] = std::move(nstr[last]);
    }
    fpar[f] = fsep[f].empty() ? -1 : vfront[fsep[f][0]];
    r[f] = xc[fv[f + 1]] - xc[fv[f]];
    s[f] = 0;
    for (i64 u : fsep[f]) s[f] += wpost[u];
    double rr = (double)r[f], ss = (double)s[f];
    flops += rr * rr * rr / 3.0 + ss * rr * rr + ss * ss * rr;
  }
  // relaxed amalgamation (merge a child front into its parent when the explicit zeros this
  // adds stay below a size-dependent fraction)
  std::vector<i64> rep(F);
  for (i64 f = 0; f < F; ++f) rep[f] = f;
  std::vector<std::vector<i64>> members(F);
  for (i64 f = 0; f < F; ++f) members[f] = {f};
  if (amalgamate) {
    std::vector<std::vector<i64>> fch(F);
    for (i64 f = 0; f < F; ++f)
      if (fpar[f] >= 0) fch[fpar[f]].push_back(f);
    std::vector<double> z(F, 0.0);
    std::vector<i64> rr(r);
    std::vector<std::vector<i64>> kept(F);  // children that stay separate fronts
    for (i64 f = 0; f < F; ++f) {
      std::vector<i64> mem;
      // worklist: original children, then the kept children of every child merged into f
      std::vector<i64> queue(fch[f].begin(), fch[f].end());
      for (size_t qi = 0; qi < queue.size(); ++qi) {
        i64 c = queue[qi];
        double rc = (double)rr[c], rf = (double)rr[f];
        double rt = rc + rf, st = (double)s[f];
        double zz = z[c] + z[f] + rc * (rf + st - (double)s[c]);
        double tot = rt * (rt + 1.0) / 2.0 + rt * st;
        bool ok = rt <= 4 || (rt <= 16 && zz <= 0.8 * tot) || (rt <= 48 && zz <= 0.1 * tot) ||
                  zz <= 0.05 * tot;
        // GPU rule: a child of at most one 64-column panel whose separator is at least half the
        // parent's order merges: its own update is memory bound, and merging saves its frontal
        // matrix, its extend-add and a launch, for at most ~4x the child's own flops
        ok = ok || (rc <= 64 && 2.0 * (double)s[c] >= rf + st);
        // and any merge that costs at most 2 % more flops than the two fronts apart
        auto fl = [](double r, double s) { return r * r * r / 3.0 + s * r * r + s * s * r; };
        ok = ok || fl(rt, st) <= 1.02 * (fl(rc, (double)s[c]) + fl(rf, st));
        if (ok) {
          rr[f] = (i64)rt;
          z[f] = zz;
          rep[c] = f;
          mem.insert(mem.end(), members[c].begin(), members[c].end());
          members[c].clear();
          queue.insert(queue.end(), kept[c].begin(), kept[c].end());
        } else {
          kept[f].push_back(c);
        }
      }
      mem.push_back(f);
      members[f] = mem;
    }
  }
  auto find = [&](i64 f) {
    while (rep[f] != f) f = rep[f];
    return f;
  };
  // new vertex order: representatives in increasing index, each listing its members' vertices
  std::vector<i64> reps;
  for (i64 f = 0; f < F; ++f)
    if (rep[f] == f) reps.push_back(f);
  const i64 NF = (i64)reps.size();
  std::vector<i64> repidx(F, -1);
  for (i64 k = 0; k < NF; ++k) repidx[reps[k]] = k;
  Symbolic S;
  S.front_vtx.push_back(0);
  std::vector<i64> vnew(N);  // postorder vertex -> final vertex index
  i64 cnt = 0;
  for (i64 R : reps) {
    for (i64 f : members[R])
      for (i64 k = fv[f]; k < fv[f + 1]; ++k) {
        vnew[k] = cnt++;
        S.vperm.push_back(post[k]);
      }
    S.front_vtx.push_back(cnt);
  }
  S.xcol.assign(N + 1, 0);
  for (i64 k = 0; k < N; ++k) S.xcol[k + 1] = S.xcol[k] + weights[S.vperm[k]];
  for (i64 R : reps) {
    std::vector<i64> sep;
    for (i64 u : fsep[R]) sep.push_back(vnew[u]);
    std::sort(sep.begin(), sep.end());
    S.front_sep.push_back(sep);
    S.front_parent.push_back(fpar[R] >= 0 ? repidx[find(fpar[R])] : -1);
  }
  S.exact_flops = flops;
  return S;
}

}  // namespace mb
