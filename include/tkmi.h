/*
 * tkmi.h — C ABI of the MI355X (gfx950) backend for TensorKit's symmetric-tensor
 * hot path: grouped permute/braid of block-sparse TensorMaps and grouped
 * per-coupled-sector GEMM.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns
 * TKMI_OK (0) on success; on failure the message is available from
 * tkmi_last_error() (thread-local) and nothing has been launched.
 *
 * Conventions
 *  - All indices / offsets are 0-based element offsets into flat data vectors.
 *  - A "strided structure" is the reference's StridedStructure
 *    (sizes[N], strides[N], offset) (src/spaces/structure.jl:2), stored as
 *    2N+1 consecutive int64 values.
 *  - `perm` is the linearised index permutation p = (p1..., p2...) of
 *    TO.tensoradd!: dst index k takes source index perm[k] (0-based).
 *  - Streams are hipStream_t passed as void*; NULL is the default stream.
 *  - Concurrency.  Abelian and generic transform plans without a gather are
 *    immutable after creation and may be executed concurrently from several host
 *    threads on different streams.  The other plan kinds own device state that
 *    every execution writes, so each serves ONE stream at a time (calls on one plan
 *    from different streams, or overlapping in time, race silently):
 *      - a GEMM plan: its split-K workspace, dot-product partials and 3×BF16 panels
 *        (tkmi_gemm_exec, tkmi_gemm_exec_parts, tkmi_gemm_plan_autotune);
 *      - a factorisation plan: its lanes' streams, handles and workspaces
 *        (tkmi_svd_exec, tkmi_eigh_exec);
 *      - a transform plan with a gather: its coefficient matrix, rewritten by every
 *        tkmi_transform_exec_gather.
 *    A binding that shares plans between tasks keeps one such plan per stream.
 */
#ifndef TKMI_H
#define TKMI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tkmi_status {
    TKMI_OK = 0,
    TKMI_ERR_ARGUMENT = 1,       /* ArgumentError in the reference          */
    TKMI_ERR_SPACE_MISMATCH = 2, /* SpaceMismatch (sizes/lengths disagree)   */
    TKMI_ERR_INDEX = 3,          /* IndexError (bad permutation)            */
    TKMI_ERR_HIP = 4,            /* HIP runtime failure                     */
    TKMI_ERR_ALLOC = 5,          /* device allocation failed                */
    TKMI_ERR_UNSUPPORTED = 6     /* dtype / rank not supported              */
};

/* Element types: Float64, Float32, ComplexF64 and ComplexF32 (interleaved (re, im) pairs,
 * Julia's layout of Complex{T} arrays). */
enum tkmi_dtype { TKMI_F64 = 0, TKMI_F32 = 1, TKMI_C128 = 2, TKMI_C64 = 3 };

/* Arithmetic forms of the grouped GEMM. */
enum tkmi_precision {
    TKMI_PREC_F64 = 0,     /* FP64 MFMA, FP64 accumulate (drop-in path)               */
    TKMI_PREC_F32 = 1,     /* FP32 MFMA on down-converted inputs, FP32 accumulate       */
    TKMI_PREC_BF16X3 = 2   /* 3-way BF16 split products on MFMA, FP32 panels, FP64 fold */
};

typedef struct tkmi_transform_plan_s *tkmi_transform_plan;
typedef struct tkmi_gemm_plan_s *tkmi_gemm_plan;
typedef struct tkmi_graph_s *tkmi_graph;
typedef struct tkmi_svd_plan_s *tkmi_svd_plan;

/* ---- runtime ------------------------------------------------------------ */
int tkmi_init(int device);
const char *tkmi_last_error(void);
int tkmi_version(void);
int tkmi_alloc(void **ptr, size_t bytes);
int tkmi_free(void *ptr);
int tkmi_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int tkmi_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int tkmi_stream_synchronize(void *stream);

/* ---- permute / braid (add_transform_kernel!) ------------------------------
 * Abelian tree transformer upload.  Replaces the per-entry loop of
 *   add_transform_kernel!(data_dst, data_src, p, ::AbelianTreeTransformer, α, β, …)
 *   (src/tensors/indexmanipulations.jl:658-669)
 * over the table built by AbelianTreeTransformer
 *   (src/tensors/treetransformers.jl:16-42): one entry per source tree pair,
 *   (coeff, dst StridedStructure, src StridedStructure).
 * coeffs[nentries]; dst_structs / src_structs: nentries × (2N+1) int64.      */
int tkmi_abelian_plan_create(tkmi_transform_plan *plan, int32_t numind, const int32_t *perm,
                             int64_t nentries, const double *coeffs,
                             const int64_t *dst_structs, const int64_t *src_structs,
                             int64_t dst_len, int64_t src_len);

/* Generic tree transformer upload.  Replaces
 *   add_transform_kernel!(data_dst, data_src, p, ::GenericTreeTransformer, α, β, …)
 *   (src/tensors/indexmanipulations.jl:670-726)
 * over the table built by GenericTreeTransformer (treetransformers.jl:54-120):
 * per fusion block b: a (rows_b × cols_b) recoupling matrix U_b (column-major,
 * concatenated in `U`), the common sub-block sizes (N int64, sz_dst / sz_src)
 * and per tree (strides[N], offset) (N+1 int64, concatenated per block).
 * The kernel computes dst_j = β dst_j + α Σ_i U_b[j,i] permute(src_i, perm)
 * for every block, i.e. pack → src·Uᵀ → unpack fused in one launch.
 * w_dst / w_src: per-tree weights dim(c) of the tree's coupled sector (NULL = 1);
 * they only enter the adjoint transform, which is the adjoint with respect to
 * TensorKit's inner product Σ_c dim(c)⟨·,·⟩ (= the inverse permutation).       */
int tkmi_generic_plan_create(tkmi_transform_plan *plan, int32_t numind, const int32_t *perm,
                             int64_t nblocks, const int64_t *rows, const int64_t *cols,
                             const double *U, const int64_t *sz_dst, const int64_t *sz_src,
                             const int64_t *strides_offsets_dst,
                             const int64_t *strides_offsets_src,
                             const double *w_dst, const double *w_src,
                             int64_t dst_len, int64_t src_len);

/* dst = β·dst + α·transform(src).  dtype: TKMI_F64 or TKMI_F32 (both buffers).
 * Rounding follows the reference's tensoradd! (Julia evaluates
 * scale(C, β) + scale(A, α) without a fused multiply-add):
 *   - a block with one source tree (every abelian entry): dst = β·dst + (α·coeff)·src,
 *     with α·coeff, the product with src and β·dst each rounded, then added;
 *   - a block with several source trees: t = Σ_i U[j,i]·src_i (the recoupling GEMM),
 *     then dst = β·dst + α·t with β·dst and α·t rounded separately.
 * α and β are Float64; FP32 data is promoted (as Julia promotes Float32 data by
 * Float64 scalars) and rounded to FP32 once, when stored.  So abelian permutes are
 * bit-exact for every α and β.
 * src and dst must not overlap (TKMI_ERR_ARGUMENT), except src == dst on an
 * elementwise plan (identity permutation, equal tables: add!/scale! in place) and
 * α = 0.  BLAS convention: β = 0 writes dst without reading it; α = 0 reads no
 * source.  A plan that writes fewer than 2^25 elements and has no fusion block wider
 * than 4 trees runs as one launch; larger plans run one launch per tree class
 * (1, 2, 4, 8 trees per block).                                               */
int tkmi_transform_exec(tkmi_transform_plan plan, int32_t dtype, const void *src, void *dst,
                        double alpha, double beta, void *stream);

/* tkmi_transform_exec for every element type, with complex α = alpha_re + i·alpha_im and
 * β = beta_re + i·beta_im, and the source read conjugated when conj_src ≠ 0 (the data of
 * an AdjointTensorMap of complex tensors, src/tensors/adjoint.jl:30-48, as
 * add_transform! reads it).  The recoupling coefficients are real, so a complex element
 * is transformed as (re, im) with the same coefficients; a complex α·coeff multiplies a
 * one-tree block's source as Julia's complex product (four rounded products, then their
 * difference and sum), a complex β multiplies dst likewise.  Real dtypes need zero
 * imaginary parts (TKMI_ERR_ARGUMENT) and ignore conj_src.  Plans with gathered
 * coefficients take real data only (TKMI_ERR_UNSUPPORTED).  Accumulation is FP64 for
 * every element type; TKMI_C64 results are rounded to FP32 pairs when stored.      */
int tkmi_transform_exec_ex(tkmi_transform_plan plan, int32_t dtype, int32_t conj_src, const void *src,
                           void *dst, double alpha_re, double alpha_im, double beta_re,
                           double beta_im, void *stream);

/* Adjoint transform: dsrc = β·dsrc + α·transform†(ddst), the adjoint with
 * respect to Σ_c dim(c)⟨·,·⟩, i.e. the pullback of tkmi_transform_exec and
 * the inverse permutation (W⁻¹-weighted transposed recoupling matrices), used
 * by the permute/braid rrules (src/pullbacks/indexmanipulations.jl:1-31).    */
int tkmi_transform_adjoint_exec(tkmi_transform_plan plan, int32_t dtype, const void *ddst,
                                void *dsrc, double alpha, double beta, void *stream);

/* Coefficients gathered from a small device tensor g at execution time:
 * U[e] = coef[e]·g[idx[e]] (0 where idx[e] < 0), one entry per recoupling coefficient
 * of a generic plan (n = Σ rows·cols).  tkmi_transform_exec_gather refreshes U from g
 * and runs the transform: a contraction C = β·C + α·A·B with a small operand B,
 * folded with the permutes of blas_contract! (src/tensors/tensoroperations.jl:392-462)
 * into one pass over A and C.  U belongs to the plan: a plan with a gather serves one
 * stream at a time (calls on one plan must not overlap).  Once a plan has a gather, tkmi_transform_exec and
 * tkmi_transform_adjoint_exec on it return TKMI_ERR_ARGUMENT (its U has no meaning
 * without the current g). */
int tkmi_transform_plan_set_gather(tkmi_transform_plan plan, int64_t n, const double *coef, const int64_t *idx);
int tkmi_transform_exec_gather(tkmi_transform_plan plan, int32_t dtype, const void *src, void *dst,
                               const void *gsrc, double alpha, double beta, void *stream);
int tkmi_transform_plan_info(tkmi_transform_plan plan, int64_t *ntiles, int64_t *nentries,
                             int64_t *elements);
int tkmi_transform_plan_destroy(tkmi_transform_plan plan);

/* ---- grouped GEMM (mul! per coupled sector) --------------------------------
 * Replaces the merge-walk of LinearAlgebra.mul!(tC, tA, tB, α, β)
 * (src/tensors/linalg.jl:330-370): one C_g = α·op(A_g)·op(B_g) + β·C_g per
 * group g, all groups in one launch.  Blocks are column-major
 * (reshape(view(data, r), d₁, d₂), src/tensors/tensor.jl:474-479).
 * op(A_g) is m×k: transa=0 → A[i + l·lda]; transa=1 → A[l + i·lda]
 * (the adjoint block of an AdjointTensorMap, src/tensors/adjoint.jl:30).
 * op(B_g) is k×n likewise.  k = 0 gives C_g = β·C_g (the rmul! branch).
 * Groups on the matrix cores need lda·(transa ? 128 : 32) and
 * ldb·(transb ? 32 : 128) below 2^29 (32-bit byte offsets inside a K slab);
 * larger leading dimensions return TKMI_ERR_UNSUPPORTED.                      */
int tkmi_gemm_plan_create(tkmi_gemm_plan *plan, int64_t ngroups, const int64_t *m,
                          const int64_t *n, const int64_t *k, const int64_t *a_off,
                          const int64_t *lda, const int32_t *transa, const int64_t *b_off,
                          const int64_t *ldb, const int32_t *transb, const int64_t *c_off,
                          const int64_t *ldc);
/* C_g = α·op(A_g)·op(B_g) + β·C_g for every group, in precision TKMI_PREC_F64,
 * TKMI_PREC_F32 (inputs rounded to FP32, FP32 accumulation) or TKMI_PREC_BF16X3 (each
 * input split into three bf16 terms, six products, FP32/FP64 accumulation).  BLAS
 * convention: β = 0 writes C without reading it; α = 0 reads neither A nor B
 * (C_g = β·C_g, zeros for β = 0), as mul! reduces to rmul!.  The bandwidth-bound
 * groups (min(m, n) ≤ 32 with k ≤ 64, and 1×1 results with longer k, i.e. full
 * contractions to a scalar) compute in FP64 for every precision form; a 1×1 result
 * is a dot product of fixed summation order (partials over k chunks, then their sum),
 * with partial sums in a plan-owned buffer.  The plan's workspaces (split-K partials,
 * dot partials, 3×BF16 panels) are written by every call: one stream at a time.  */
int tkmi_gemm_exec(tkmi_gemm_plan plan, int32_t precision, double alpha, double beta,
                   const double *A, const double *B, double *C, void *stream);
/* Tile shape for `precision`: 0 = automatic, 64 / 128 = square tiles with a
 * 16-deep K slab, 1000 + i = configuration i (see tkmi_gemm_config_info).      */
int tkmi_gemm_plan_set_tile(tkmi_gemm_plan plan, int32_t precision, int32_t tile);
/* Tile configuration i: BM×BN output tile, BK-deep K slab and the FP64/FP32 kernel's
 * staging (2: double-buffered LDS, gemm_k4_kernel; 1: one LDS buffer and one register
 * stage, gemm_k4s_kernel; 3: one LDS buffer and two register stages, gemm_k4p_kernel;
 * 5: the persistent gemm_pks_kernel; 6: gemm_k4d_kernel, direct global → LDS loads into
 * two LDS stages, FP64 arithmetic with untransposed A and B only).  Configurations 0–9
 * and 13 take the split-K factor of the schedule model; 10, 11 and 12 are configuration 5
 * with the factor fixed at 1, 2 and 3, and 14, 15 and 16 configuration 13 likewise.
 * In the FP32 form on FP32 data they run gemm_k4d32_kernel (the same direct loads), in
 * the 3×BF16 form gemm_bf16x3d_kernel; where no direct-load kernel applies (a transposed
 * operand, FP32 arithmetic on FP64 data) configurations 13–16 run configuration 5's
 * kernel with their own split factor (tkmi_gemm_plan_splitk reports the factor of each).
 * Configurations 17, 18 and 19 are the direct-load kernel on 128×128, 128×64 and 64×128
 * tiles (gemm_k4dw_kernel; model split factor; FP64 with untransposed operands, else the
 * one-buffer kernel of the same tiles).  Untuned FP64 plans of untransposed products
 * run 13, the rest 5.                                                                 */
int tkmi_gemm_config_info(int32_t config, int32_t *bm, int32_t *bn, int32_t *bk, int32_t *nbuf);
/* Same as tkmi_gemm_exec restricted to a subset of the plan's launches:
 * parts bit 0 = MFMA tiles (groups with min(m,n) > 32 or k > 64) with their split-K
 * reduction, bit 1 = the bandwidth-bound groups (skinny ones, and 1×1 results with
 * k > 64, which run as two-pass dot products); parts = 3 is tkmi_gemm_exec.
 * For timing the kernels one by one: bit 2 (with bit 0) leaves out the split-K
 * reduction launch, bit 3 (without bit 0) runs that reduction alone; either leaves
 * the split groups of C without their full sums.                              */
int tkmi_gemm_exec_parts(tkmi_gemm_plan plan, int32_t precision, int32_t parts, double alpha,
                         double beta, const double *A, const double *B, double *C, void *stream);
/* tkmi_gemm_exec_parts for the element type `dtype` of A, B and C (all three):
 * TKMI_F64, or TKMI_F32 for the blocks of Float32 TensorMaps.  FP32 data runs in
 * TKMI_PREC_F32 (FP32 MFMA: the sgemm of a Float32 mul!, src/tensors/linalg.jl:330-370)
 * or TKMI_PREC_BF16X3 (a float splits exactly into its three bf16 terms); each element of
 * C is formed in FP64, α·(A·B) + β·C, and rounded once to FP32.  Split-K partials and dot
 * partials stay FP64.  TKMI_PREC_F64 with TKMI_F32 returns TKMI_ERR_UNSUPPORTED.      */
int tkmi_gemm_exec_dtype(tkmi_gemm_plan plan, int32_t precision, int32_t dtype, int32_t parts,
                         double alpha, double beta, const void *A, const void *B, void *C,
                         void *stream);
/* C_g = α·op(A_g)·op(B_g) + β·C_g for complex data (dtype TKMI_C128 or TKMI_C64, A, B and
 * C alike), complex α = alpha_re + i·alpha_im, β likewise: the zgemm/cgemm of mul! on
 * ComplexF64 / ComplexF32 blocks (src/tensors/linalg.jl:330-370).  conj_a / conj_b
 * conjugate op(A) / op(B) — with transa / transb set, op is the conjugate transpose,
 * the adjoint block of an AdjointTensorMap.  precision TKMI_PREC_F64 multiplies on FP64
 * MFMA, TKMI_PREC_F32 on FP32 MFMA (inputs rounded to FP32), four real products per
 * complex one; results are formed in FP64 and rounded once to C's type.  Every group of
 * the plan runs on one 64×64-tile complex kernel (split-K partials in FP64 pairs).  The
 * first complex call on a plan builds its complex tiles (device allocation: not during
 * graph capture).  TKMI_PREC_BF16X3 returns TKMI_ERR_UNSUPPORTED.                 */
int tkmi_gemm_exec_z(tkmi_gemm_plan plan, int32_t precision, int32_t dtype, int32_t conj_a,
                     int32_t conj_b, double alpha_re, double alpha_im, double beta_re,
                     double beta_im, const void *A, const void *B, void *C, void *stream);
/* Algorithmic flops routed to the MFMA kernel and to the skinny kernel.       */
int tkmi_gemm_plan_split(tkmi_gemm_plan plan, double *flops_mfma, double *flops_skinny,
                         int64_t *skinny_groups);
int tkmi_gemm_plan_info(tkmi_gemm_plan plan, int64_t *ntiles, double *flops);
/* Plan-wide split-K factor chosen for tile configuration `config` (1 = no split) and
 * the workgroup slots (occupancy × compute units) of its FP64 kernel that the choice
 * assumed.  The factor minimises a model of the launch: the heaviest-first schedule of
 * the tiles over the slots plus the partial products' extra HBM traffic.          */
int tkmi_gemm_plan_splitk(tkmi_gemm_plan plan, int32_t config, int64_t *splitk, int64_t *slots);
/* Autotuner: times every tile shape of the plan on the given operands (α=1,
 * β=0: C is overwritten) and keeps the fastest for `precision`.  Synchronises
 * `stream`; meant for warm-up, not for capture.                               */
int tkmi_gemm_plan_autotune(tkmi_gemm_plan plan, int32_t precision, const double *A, const double *B,
                            double *C, int32_t reps, void *stream);
/* The autotuner on operands of element type `dtype` (see tkmi_gemm_exec_dtype).  The
 * choice is kept per precision form, for every element type: tune with the data the
 * plan runs on.                                                                 */
int tkmi_gemm_plan_autotune_dtype(tkmi_gemm_plan plan, int32_t precision, int32_t dtype,
                                  const void *A, const void *B, void *C, int32_t reps,
                                  void *stream);
/* 3×BF16 operands that do not change between calls (an environment tensor of a sweep,
 * constant across a Lanczos or Davidson run): operands = 1 (A), 2 (B), 3 (both), 0 (none).
 * The split of a constant operand into its three bf16 planes is made by the first
 * TKMI_PREC_BF16X3 call and reused while later calls pass the same pointer; a call with
 * another pointer splits it again.  After changing a constant operand's data in place,
 * call tkmi_gemm_plan_invalidate, or the next call multiplies the old split.  A captured
 * graph holds only the splits its capture made: make one eager 3×BF16 call first.     */
int tkmi_gemm_plan_set_constant(tkmi_gemm_plan plan, int32_t operands);
int tkmi_gemm_plan_invalidate(tkmi_gemm_plan plan);
/* Launch timing: the next MFMA-tile kernel this plan launches (any precision; not its
 * split pass, split-K reduction or skinny launch) records its own start and end in the
 * two events (hipEvent_t, created with timing enabled) through hipExtLaunchKernel, so
 * hipEventElapsedTime gives that kernel's execution time with no event packets of
 * their own around it on the stream.  One-shot: cleared by that launch; NULL, NULL
 * clears it.  No counterpart in the reference: a measurement hook for timing a launch
 * where it runs, between the other launches of a chain; not for graph capture.      */
int tkmi_gemm_plan_time_next(tkmi_gemm_plan plan, void *start_event, void *stop_event);
/* Calibration of the FP64 matrix-core rate (no counterpart in the reference: it serves
 * the roofline of the grouped GEMM): `blocks` workgroups of 4 waves, each wave `iters`
 * rounds of 16 v_mfma_f64_16x16x4_f64 (2 048 flops each) with no memory traffic;
 * out holds blocks × 256 doubles.  flops = blocks · 4 · iters · 16 · 2048.           */
int tkmi_calib_mfma_f64(int64_t blocks, int32_t iters, double *out, void *stream);
/* Configuration index and tile count the next exec with `precision` uses.      */
int tkmi_gemm_plan_query(tkmi_gemm_plan plan, int32_t precision, int32_t *config, int64_t *ntiles);
int tkmi_gemm_plan_destroy(tkmi_gemm_plan plan);

/* ---- svd_trunc (FP64 block factorisation + truncation on the device) -------
 * Per coupled block A_c (m_c × n_c, column-major at a_off[c] with leading
 * dimension lda[c]): A_c = U_c·diag(S_c)·Vt_c, r_c = min(m_c, n_c).  Outputs are
 * packed block after block: U_c (m_c×r_c, ld m_c), S_c (r_c), Vt_c (r_c×n_c,
 * ld r_c) — the layouts of U: codomain←V_r, Vᴴ: V_r←domain as TensorMaps.
 * The blocks run concurrently on streams owned by the plan, ordered after the
 * work already on `stream`; `stream` waits for all of them before continuing.
 * The plan's lanes (streams, handles, workspaces) serve one call at a time.
 * Replaces MAK.svd_compact! per block (src/factorizations/matrixalgebrakit.jl:34-51). */
int tkmi_svd_plan_create(tkmi_svd_plan *plan, int64_t nblocks, const int64_t *m, const int64_t *n,
                         const int64_t *a_off, const int64_t *lda);
int tkmi_svd_plan_sizes(tkmi_svd_plan plan, int64_t *u_len, int64_t *s_len, int64_t *vt_len);
int tkmi_svd_exec(tkmi_svd_plan plan, const double *A, double *U, double *S, double *Vt, void *stream);
/* Block SVD driver, the `alg` of MAK.svd_compact!: MatrixAlgebraKit's
 * ROCSOLVER_QRIteration / ROCSOLVER_DivideAndConquer / ROCSOLVER_Jacobi
 * (TensorKit forwards `alg` per block, src/factorizations/matrixalgebrakit.jl:34-51). */
#define TKMI_SVD_QRITERATION 0
#define TKMI_SVD_DIVIDEANDCONQUER 1
#define TKMI_SVD_JACOBI 2
/* This library's own block one-sided Jacobi SVD: every block of a call in the same
 * launches (one workgroup per pair of 16-column blocks per round, sweeps until the columns
 * are orthogonal to √m·ε), σ accurate to ε·‖A‖ as with the drivers above.  It has no
 * MatrixAlgebraKit counterpart; SVD plans only (tkmi_eigh_* refuse it). */
#define TKMI_SVD_BLOCKJACOBI 3
int tkmi_svd_plan_set_algorithm(tkmi_svd_plan plan, int32_t alg);
int tkmi_svd_plan_destroy(tkmi_svd_plan plan);

/* ---- eigh_full / eigh_trunc (symmetric blocks, FP64 on the device) ----------
 * Per square block A_c (n_c × n_c at a_off[c], leading dimension lda[c], lower
 * triangle read): A_c = V_c·diag(λ_c)·V_cᵀ, λ_c ascending.  Outputs packed block after
 * block: λ_c (n_c values) and V_c (n_c×n_c, ld n_c), the layout of
 * V: codomain ← V_r as a TensorMap.  The plan is a factorisation plan: sizes
 * (λ length in s_len, V length in u_len), algorithm (divide and conquer = dsyevd,
 * Jacobi = dsyevj, QR iteration = dsyev) and destruction go through the tkmi_svd_plan_*
 * calls.  Replaces MAK.eigh_full! per block (src/factorizations/matrixalgebrakit.jl:34-51);
 * truncation (eigh_trunc, by |λ|) uses tkmi_truncate_select. */
int tkmi_eigh_plan_create(tkmi_svd_plan *plan, int64_t nblocks, const int64_t *n, const int64_t *a_off,
                          const int64_t *lda);
int tkmi_eigh_exec(tkmi_svd_plan plan, const double *A, double *W, double *V, void *stream);

/* Keep-mask for truncrank(howmany) ∩ trunctol(atol) over all singular values
 * (weights = quantum dimension of each value's sector) and the truncation
 * error ε = sqrt(Σ_discarded w·σ²) (findtruncated / truncation_error!,
 * src/factorizations/truncation.jl:171-322).  S, weights, keep, disc_work
 * (≥ nvals doubles) and err (1 double) are device pointers.                   */
int tkmi_truncate_select(int64_t nvals, const double *S, const double *weights, double howmany,
                         double atol, unsigned char *keep, double *disc_work, double *err,
                         void *stream);

/* ---- reductions ------------------------------------------------------------
 * work[0] = Σ_i a[i]·b[i] over n elements (device result, fixed summation order).
 * `work` is a caller-owned device buffer of at least 1025 doubles.  Used for
 * TensorKit's inner(A, B) = Σ_c dim(c)·⟨A_c, B_c⟩ (src/tensors/linalg.jl) in the
 * reverse-mode rules.                                                          */
int tkmi_dot(int32_t dtype, const void *a, const void *b, int64_t n, double *work, void *stream);

/* result[0] = Σ_i a[i]·b[i], the same kernels and summation order as tkmi_dot (so the
 * same bits), with no workspace argument and no allocation of any kind.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The partial
 * sums live in a static device array of 1025 doubles per GPU, kept until the process
 * ends; any n fits (at most 1024 partials).  Every call on one GPU shares that array:
 * calls that overlap in time on one GPU (different streams, or host threads) race, so
 * order them (one stream, or events).  `result` is a caller-owned device double.     */
int tkmi_dot_noalloc(int32_t dtype, const void *a, const void *b, int64_t n, double *result, void *stream);

/* ---- hipGraph capture of a chain of launches (one sweep step) ------------ */
int tkmi_graph_begin(void *stream);
int tkmi_graph_end(void *stream, tkmi_graph *graph);
int tkmi_graph_launch(tkmi_graph graph, void *stream);
int tkmi_graph_destroy(tkmi_graph graph);

#ifdef __cplusplus
}
#endif
#endif /* TKMI_H */
