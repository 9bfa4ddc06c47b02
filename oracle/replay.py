"""Oracle (test infrastructure): host replay of a recorded device kernel sequence.

`tensors.start_trace()` records every launch of a high-level call (transforms, grouped
GEMMs, fused small-operand contractions) with its plan and buffers.  `cpu_replay` runs
the same sequence on the host: the permute loops of `add_transform_kernel!`
(src/tensors/indexmanipulations.jl:658-726) in the oracle's C code over the same tables,
and one numpy/OpenBLAS GEMM per coupled sector (the `mul!` merge-walk,
src/tensors/linalg.jl:330-370).  A fused small-operand contraction is replayed as the
reference's unfused `blas_contract!` (src/tensors/tensoroperations.jl:392-462): operand
permutes, per-sector GEMMs, result permute.  Used by the tests (full-size parity of the
two-site H_eff chain) and by the benchmark's CPU baseline; nothing in the product
imports it.
"""


def prepare_cpu_replay(trace):
    """Host-side tables of the replay, built once outside the timed region (the reference
    caches its transformer tables; table construction is excluded on both sides): the
    recorded transforms and GEMMs as they are, and for each fused small-operand
    contraction the reference's blas_contract! steps — the two operand permute tables,
    the per-sector GEMM list and the result permute table — with their temporaries."""
    import numpy as np
    from tkmi import treetransformers as TT
    steps = []
    for op in trace:
        if op[0] != "fused":
            steps.append(op)
            continue
        _, plan, (pA, pB, pC, twA, twB, WA, WB, WC), ta, tb, tc, a, b = op
        WA2, WB2 = WA.permute(*pA), WB.permute(*pB)
        WC2 = WA2.compose(WB2)
        LA = {c: (d, o) for c, d, o in WA2.degeneracystructure()[1]}
        LB = {c: (d, o) for c, d, o in WB2.degeneracystructure()[1]}
        sectors = []
        for c, (m, n), off in WC2.degeneracystructure()[1]:
            if c in LA and c in LB:
                sectors.append((m, n, LA[c][0][1], LA[c][1], LB[c][1], off))
        steps.append(("fused", TT.abelian_table(WA2, WA, pA[0], pA[1], None, twA), pA[0] + pA[1],
                      TT.abelian_table(WB2, WB, pB[0], pB[1], None, twB), pB[0] + pB[1],
                      TT.abelian_table(WC, WC2, pC[0], pC[1]), pC[0] + pC[1], sectors,
                      np.zeros(WA2.dim()), np.zeros(WB2.dim()), np.zeros(WC2.dim()), ta, tb, tc, a, b))
    return steps


def cpu_replay(steps, init, nthreads):
    """Replay the recorded kernel sequence on the host with the oracle's C permute loop
    and numpy (OpenBLAS) GEMMs per coupled sector (`steps` from prepare_cpu_replay)."""
    import numpy as np
    from oracle import kernels as OK
    bufs = {}

    def get(tn):
        t, n = tn
        p = t.data_ptr()
        if p not in bufs:
            bufs[p] = init[p].copy() if p in init else np.zeros(n)
        return bufs[p]

    for op in steps:
        if op[0] == "transform":
            _, plan, perm, src, dst, a, b = op
            s, d = get(src), get(dst)
            if plan.kind == "abelian":
                OK.abelian_permute(plan.table, perm, s, d, a, b, nthreads)
            else:
                OK.generic_permute(plan.table, perm, s, d, a, b, nthreads)
        elif op[0] == "fused":
            # the reference's blas_contract!: both operand permutes, one dgemm per sector,
            # then the result permute
            _, tabA, permA, tabB, permB, tabC, permC, sectors, A2, B2, C2, ta, tb, tc, a, b = op
            A, B, Cm = get(ta), get(tb), get(tc)
            OK.abelian_permute(tabA, permA, A, A2, 1.0, 0.0, nthreads)
            OK.abelian_permute(tabB, permB, B, B2, 1.0, 0.0, nthreads)
            for m, n, k, oa, ob, off in sectors:
                np.matmul(A2[oa:oa + m * k].reshape(k, m).T, B2[ob:ob + k * n].reshape(n, k).T,
                          out=C2[off:off + m * n].reshape(n, m).T)
            OK.abelian_permute(tabC, permC, C2, Cm, a, b, nthreads)
        else:
            _, plan, prec, pa, pb, pc, a, b = op
            A, B, Cm = get(pa), get(pb), get(pc)
            for (m, n, k, oa, lda, ta, ob, ldb, tb, oc, ldc) in plan.table:
                if m == 0 or n == 0:
                    continue
                c = Cm[oc:oc + ldc * n].reshape(n, ldc).T[:m, :]
                if k == 0:
                    c *= b
                    continue
                am = A[oa:oa + lda * m].reshape(m, lda)[:, :k] if ta else A[oa:oa + lda * k].reshape(k, lda).T[:m]
                bm = B[ob:ob + ldb * k].reshape(k, ldb)[:, :n] if tb else B[ob:ob + ldb * n].reshape(n, ldb).T[:k]
                if b == 0:
                    np.matmul(am, bm, out=c)  # straight into C, no temporaries
                    if a != 1:
                        c *= a
                else:
                    c *= b
                    c += a * (am @ bm)
    return bufs
