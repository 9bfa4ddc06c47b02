"""ORACLE — test infrastructure only.

A CPU restatement of TensorKit's hot path used exclusively as the checker:
only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` may import it.  The product (tensorkit.jl_amd/tkmi + libtkmi.so)
never calls into this package.

Contents
  abelian.py     independent restatement of the block/sub-block layout of
                 abelian (U1, fZ2⊠U1) HomSpaces and the dense embedding
                 `convert(Array, t)`; U1 permute = numpy transpose of the dense
                 embedding (pure data movement); fermionic sign = Koszul sign
                 of the linearised permutation.
  su2.py         Clebsch-Gordan (Condon-Shortley) fusion tensors and the dense
                 embedding of SU(2)-symmetric tensors (`fusiontensor`,
                 `convert(Array, t)`) used to pin the recoupling matrices.
  ref_kernels.c  the table-driven per-sub-block permute loop of
                 `add_transform_kernel!` (abelian and generic) in plain C
                 (OpenMP over entries) — the CPU baseline and a kernel-level
                 checker.
  kernels.py     ctypes loader of the compiled ref_kernels.c.

How each part is pinned (there are no golden vectors in the reference; see
the test files for the checks):
  layout         tree counts and block dimensions vs closed-form counts,
                 total dim = Σ blocks, sector order vs the documented
                 `isless` (docs/src/man/sectors.md:262-264)
  U1 permute     permute∘inverse = id, norm invariance, dense transpose
  fermion sign   closed-form Koszul sign; involution / inner-product checks
  SU(2)          CG orthogonality, F-symbol pentagon / unitarity, dense
                 embedding commutes with permutedims
  GEMM           exact integer inputs (exact in FP64) and numpy float128
                 comparisons on small blocks
"""
