"""CPU oracle: an independent numpy restatement of Mumblebee.jl's conic IPM solve path.

This package is TEST INFRASTRUCTURE. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import it, and only as the checker (or as the timed
CPU restatement in the baseline leg). The product path (the HIP library and its Python
mirror in `mumblebee.jl_amd/mumblebee_hip`) never imports, links or calls anything here.

Each function cites the reference file:line it restates (paths relative to the
Mumblebee.jl repository root). Indices are 0-based here; the reference is 1-based.

Parity status of each part (what pins it) is listed in the module docstrings:
  bsm.py        block-sparse storage, gemv, BᵀB, vertex selections
  problem.py    dropoffdz / colpartition / twins / colcompress / compress2
  symbolic.py   weighted vertex graph, elimination ordering, etree, supernodal fronts
  scaling.py    block-aware Ruiz equilibration
  cones.py      nonnegative / free / second-order cone kernels
  kkt.py        Uzawa iterative refinement, dual CG, kktwindow, factorization with ρ retry
  ipm.py        IPMSolver, step!, solve, result, history
"""
