// Host-side data structures of the MI355X conic IPM path: sparse inputs, block-sparse
// matrices in the reference's storage contract, and the symbolic factor structure.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "kinds.h"

namespace mb {

using i64 = int64_t;
using i32 = int32_t;

struct Cone {
  i32 kind = CONE_POS;
  double param = 0.0;  // power-cone alpha
};

// Compressed sparse column, 0-based, rows sorted within each column, stored zeros kept.
struct CSC {
  i64 m = 0, n = 0;
  std::vector<i64> colptr, rowval;
  std::vector<double> nzval;
  i64 nnz() const { return colptr.empty() ? 0 : colptr[n]; }
};

// Block-sparse matrix: the diagram U <-tgt- E -src-> V of
// src/BlockSparseArrays/src/block_sparse_matrix.jl:10-158, 0-based.
struct BSM {
  i64 nout = 0, nvtx = 0;
  std::vector<i64> xsrc, xcol, xrow, xblk, tgt;
  std::vector<double> val;
  i64 narc() const { return xsrc.empty() ? 0 : xsrc[nvtx]; }
  i64 ncol() const { return xcol.empty() ? 0 : xcol[nvtx]; }
  i64 nrow() const { return xrow.empty() ? 0 : xrow[nout]; }
  i64 nbnz() const { return xblk.empty() ? 0 : xblk[narc()]; }
  i64 ncols(i64 v) const { return xcol[v + 1] - xcol[v]; }
  i64 nrows(i64 u) const { return xrow[u + 1] - xrow[u]; }
};

// IPMProblem (src/IPM/src/ipm/problem.jl:21-42) after construction.
struct Problem {
  BSM Q, B;
  std::vector<double> f, g;
  double mu = 0.0;
  std::vector<Cone> K;          // one per vertex
  std::vector<i64> P1, P2;      // new -> old row / column indices
};

// problem.jl:84-127
Problem make_problem(const CSC& Q, const CSC& B, const std::vector<double>& f,
                     const std::vector<double>& g, double mu, const std::vector<Cone>& K,
                     const std::vector<i64>& s, double compress);

// compress.jl:9-116 (exact) and :147-265 (cosine tau < 1)
i64 twins_exact(i64 nout, i64 nvtx, const i64* xsrc, const i64* tgt, std::vector<i64>& ptr,
                std::vector<i64>& perm);
i64 twins_tau(i64 nout, i64 nvtx, const i64* xrow, const i64* cadj, const i64* xcol,
              const i64* radj, double tau, std::vector<i64>& ptr, std::vector<i64>& perm);

CSC csc_transpose(const CSC& A);
CSC dropoffdz(const CSC& Q);

BSM selectvtxs(const BSM& A, const std::vector<i64>& sel);
BSM symselectvtxs(const BSM& A, const std::vector<i64>& sel);
BSM transpose_bsm(const BSM& A);
// The lower triangle (i ≥ j) of BᵀB as scalar CSC (colp: n + 1 offsets; rows sorted within each
// column), from the stored nonzeros of B given by rows (rp, rc, rv) and by columns (cp, cr, cv).
// Each entry sums its products in the order of B's rows.
void gram_lower(i64 n, const std::vector<int>& rp, const std::vector<int>& rc, const std::vector<double>& rv,
                const std::vector<int>& cp, const std::vector<int>& cr, const std::vector<double>& cv,
                std::vector<i64>& colp, std::vector<int>& rows, std::vector<double>& vals);

// Symbolic factor structure in the final vertex order.
struct Symbolic {
  std::vector<i64> vperm;       // new vertex -> old vertex
  std::vector<i64> xcol;        // column offsets of new vertices (nvtx+1)
  std::vector<i64> front_vtx;   // front f owns new vertices [front_vtx[f], front_vtx[f+1])
  std::vector<std::vector<i64>> front_sep;  // sorted separator (new vertices)
  std::vector<i64> front_parent;            // -1 for roots
  double exact_flops = 0.0;     // Σ r³/3 + s r² + s² r over fundamental supernodes
  i64 nfront() const { return (i64)front_sep.size(); }
};

// Ordering, etree, postorder, fundamental supernodes, relaxed amalgamation.
Symbolic analyse(const std::vector<i64>& weights, const std::vector<std::vector<i64>>& adj,
                 bool amalgamate);
// analyse() of vertex_graph(B, Q), with the graph set straight into the ordering's bitsets when
// the vertex count allows; the same result
Symbolic analyse_kkt(const BSM& B, const BSM& Q, bool amalgamate);
void vertex_graph(const BSM& B, const BSM& Q, std::vector<i64>& weights,
                  std::vector<std::vector<i64>>& adj);

}  // namespace mb
