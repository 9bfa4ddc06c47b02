// Block one-sided Jacobi SVD of all the coupled blocks of a tensor at once (FP64).
//
// A block A (m×n) is brought to its tall orientation W (M×N, M ≥ N: W = A, or Aᵀ when
// m < n) and W is rotated from the right, W ← W·J, until its columns are mutually
// orthogonal; the product of the rotations is V (N×N, from the identity).  Then
// W = U·Σ with σ_j = ‖w_j‖ and u_j = w_j/σ_j, so A = U·Σ·Vᵀ (or, transposed, A = V·Σ·Uᵀ).
//
// The N columns are cut into blocks of JB = 16.  A sweep pairs every two blocks once
// (a round-robin tournament: NB − 1 rounds of NB/2 disjoint pairs, NB rounded up to even
// with an empty block); a round is one launch in which one workgroup takes one pair of
// one matrix, and every matrix of the plan has its pairs in the same launch.  For a pair,
// the workgroup forms the 32×32 Gram matrix G = PᵀP of its columns P (rows staged through
// LDS in chunks of 64), records the largest cosine |g_ij|/√(g_ii·g_jj) of the pair, and,
// when it exceeds the tolerance √M·ε, diagonalises G with INNER_SWEEPS cyclic Jacobi sweeps
// in LDS (Golub & Van Loan, Matrix Computations, §8.5: the symmetric Schur rotation) into
// Q, then applies P ← P·Q and V ← V·Q (Hari & Veselić's block one-sided method; the outer
// iteration is the one that converges, so the inner problem needs no exact solution).
// After each sweep the host reads the pairs' largest cosines: a matrix whose columns are
// orthogonal to √M·ε relative to their norms is done; σ are the column norms, so they are
// accurate to ε·‖A‖ like a bidiagonalising SVD, and U is orthonormal to √M·ε per pair.
//
// Columns that end at exactly zero (an exactly rank-deficient block) get an orthonormal
// completion (Gram–Schmidt applied twice to unit vectors), as a thin SVD must return
// orthonormal U and Vᴴ whatever the rank.
#include "tkmi_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tkmi {
namespace {

constexpr int JT = 256;  // threads per workgroup
constexpr int JB = 16;   // columns per Jacobi block
constexpr int JP = 2 * JB;
constexpr int JR = 128;  // rows per staged chunk
constexpr int INNER_SWEEPS = 1;
constexpr int MAX_SWEEPS = 40;

struct JMat {
    int64_t M, N;                   // oriented form: M ≥ N
    int64_t w_off, v_off;           // W (M×N, ld M) and V (N×N, ld N) in the workspaces
    int64_t a_off, lda;             // the input block (m×n, column-major)
    int64_t u_off, s_off, vt_off;   // outputs: U (m×r, ld m), S (r), Vt (r×n, ld r)
    int64_t col0;                   // first column of this matrix in the plan-wide numbering
    int32_t trans;                  // 1: W = Aᵀ (m < n)
    int32_t nb, nbe;                // Jacobi blocks, and rounded up to even
    int32_t pad_;
    double tol;                     // √M·ε
};

// Pair u of round k of a round-robin over n (even) players: player n − 1 stays, the
// others rotate, so over n − 1 rounds every two players meet exactly once.
__device__ __forceinline__ void pair_of(int u, int k, int n, int &a, int &b) {
    const int r = n - 1;
    if (u == 0) {
        a = k;
        b = r;
    } else {
        a = (k + u) % r;
        b = (k - u + r) % r;
    }
}

__device__ __forceinline__ int find_mat(const JMat *__restrict__ mats, int64_t nmats, int64_t c) {
    int64_t lo = 0, hi = nmats - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (mats[mid].col0 <= c) lo = mid; else hi = mid - 1;
    }
    return (int)lo;
}

// W = A (or Aᵀ) column by column, V = I: one workgroup per plan-wide column.
__global__ __launch_bounds__(JT) void jac_init(const JMat *__restrict__ mats, int64_t nmats,
                                               const double *__restrict__ A, double *__restrict__ W,
                                               double *__restrict__ V) {
    const int64_t c = blockIdx.x;
    const JMat &J = mats[find_mat(mats, nmats, c)];
    const int64_t j = c - J.col0;
    for (int64_t i = threadIdx.x; i < J.M; i += JT)
        W[J.w_off + j * J.M + i] = J.trans ? A[J.a_off + j + i * J.lda] : A[J.a_off + i + j * J.lda];
    for (int64_t i = threadIdx.x; i < J.N; i += JT) V[J.v_off + j * J.N + i] = (i == j) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(JT) void jac_reset(double *__restrict__ slots, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * JT + threadIdx.x; i < n; i += (int64_t)gridDim.x * JT) slots[i] = 0.0;
}

// One round: workgroup w rotates pair wg_t[w] of round (g mod (nbe − 1)) of matrix wg_mat[w].
// RT = 1024 threads (four waves per SIMD to hide the latency of the staged loads).  Rows of
// the pair's columns are staged JR at a time through two LDS buffers; the global loads of
// the next chunk are issued before the current chunk is used.
constexpr int RT = 1024;
constexpr int RPT = JR * JP / RT;  // staged elements per thread per chunk
constexpr int GG = RT / 64 / 4;    // row groups of the Gram phase: 16 waves over 4 tiles of G
typedef double d4_t __attribute__((ext_vector_type(4)));
static_assert(JP == 32 && JR == 128 && RT == 1024, "the matrix-core tilings assume these sizes");

__global__ __launch_bounds__(RT) void jac_round(const JMat *__restrict__ mats, const int32_t *__restrict__ wg_mat,
                                                const int32_t *__restrict__ wg_t, int64_t g,
                                                double *__restrict__ W, double *__restrict__ V,
                                                double *__restrict__ slots) {
    __shared__ double Ls[2][JR][JP + 1];
    __shared__ double Gp[GG][JP][JP + 1];  // partial Gram matrices; Gp[0] is then G
    __shared__ double Q[JP][JP + 1];
    __shared__ double red[RT];
    __shared__ double rc[JB], rs[JB];
    __shared__ int cols[JP];
    __shared__ int rotated;
    const int tid = threadIdx.x;
    const JMat &J = mats[wg_mat[blockIdx.x]];
    const int t = wg_t[blockIdx.x];
    int p1, p2;
    pair_of(t, (int)(g % (J.nbe - 1)), J.nbe, p1, p2);
    const int64_t M = J.M, N = J.N;
    if (tid < JP) {
        const int p = tid < JB ? p1 : p2;
        const int64_t c = (int64_t)p * JB + (tid % JB);
        cols[tid] = (p < J.nb && c < N) ? (int)c : -1;
    }
    __syncthreads();

    // element k of this thread in a chunk: row (tid + k·RT) % JR of local column (tid + k·RT) / JR
    double pre[RPT];
    auto fetch = [&](const double *X, int64_t rows, int64_t r0) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int e = tid + k * RT, row = e % JR, c = e / JR;
            const int gc = cols[c];
            pre[k] = (gc >= 0 && r0 + row < rows) ? X[(int64_t)gc * rows + r0 + row] : 0.0;
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int e = tid + k * RT;
            Ls[buf][e % JR][e / JR] = pre[k];
        }
    };

    // Gram matrix on the matrix cores (v_mfma_f64_16x16x4_f64: A[i][k] from lane i + 16k,
    // B[k][j] from lane j + 16k, D[i][j] in lane j + 16(i % 4), value i / 4): wave w
    // accumulates 16×16 tile (TI, TJ) = (w & 1, (w >> 1) & 1) of G over rows kg·32 ... + 31 of
    // every chunk, kg = w / 4; the four row groups are added afterwards.
    const int lane = tid & 63, wave = tid >> 6;
    const int TI = wave & 1, TJ = (wave >> 1) & 1, kg = wave >> 2;
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    const double *Wm = W + J.w_off;
    const int64_t nch = (M + JR - 1) / JR;
    fetch(Wm, M, 0);
    for (int64_t ch = 0; ch < nch; ++ch) {
        const int buf = (int)(ch & 1);
        put(buf);
        __syncthreads();
        if (ch + 1 < nch) fetch(Wm, M, (ch + 1) * JR);
#pragma unroll
        for (int st = 0; st < JR / (GG * 4); ++st) {  // rows past M are zero
            const int k = kg * (JR / GG) + st * 4 + (lane >> 4);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[buf][k][16 * TI + (lane & 15)],
                                                       Ls[buf][k][16 * TJ + (lane & 15)], acc, 0, 0, 0);
        }
        // no barrier here: the next put writes the other buffer, and the put after that
        // comes after the next barrier, which every thread reaches only once done here
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) Gp[kg][16 * TI + (lane >> 4) + 4 * v][16 * TJ + (lane & 15)] = acc[v];
    __syncthreads();
    double(*G)[JP + 1] = Gp[0];
    if (tid < JP * JP) {
        const int i = tid % JP, j = tid / JP;
        double s = Gp[0][i][j];
        for (int q = 1; q < GG; ++q) s += Gp[q][i][j];
        G[i][j] = s;
    }
    __syncthreads();

    // largest cosine between two columns of the pair
    double off = 0.0;
    if (tid < JP * JP) {
        const int i = tid % JP, j = tid / JP;
        if (i < j && cols[i] >= 0 && cols[j] >= 0) {
            const double d = G[i][i] * G[j][j];
            if (d > 0.0) off = fabs(G[i][j]) / sqrt(d);
        }
    }
    red[tid] = off;
    __syncthreads();
    for (int s = RT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = max(red[tid], red[tid + s]);
        __syncthreads();
    }
    const double offmax = red[0];
    if (tid == 0) slots[blockIdx.x] = max(slots[blockIdx.x], offmax);
    if (!(offmax > J.tol)) return;  // uniform over the workgroup: the pair is orthogonal

    // cyclic Jacobi on G until a sweep rotates nothing (at most INNER_SWEEPS), Q = Π J.  The
    // 16 rotations J_u of a round act on disjoint index pairs (a_u, b_u), so G′ = Jᵀ·G·J
    // splits into the 2×2 blocks of rows (a_u, b_u) and columns (a_w, b_w), each of which
    // needs only its own four entries and J_u, J_w: one step per round (256 threads for the
    // blocks, 512 for Q's row pairs) between two barriers.
    if (tid < JP * JP) Q[tid % JP][tid / JP] = (tid % JP == tid / JP) ? 1.0 : 0.0;
    for (int sw = 0; sw < INNER_SWEEPS; ++sw) {
        if (tid == 0) rotated = 0;
        for (int k = 0; k < JP - 1; ++k) {
            __syncthreads();
            if (tid < JB) {
                int a, b;
                pair_of(tid, k, JP, a, b);
                double c = 1.0, s = 0.0;
                const double app = G[a][a], aqq = G[b][b], apq = G[a][b];
                if (apq != 0.0 && fabs(apq) > J.tol * sqrt(app * aqq)) {
                    const double theta = (aqq - app) / (2.0 * apq);
                    const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                    rotated = 1;
                }
                rc[tid] = c;
                rs[tid] = s;
            }
            __syncthreads();
            if (tid < JB * JB) {  // block (u, w) of G ← J_uᵀ·G·J_w
                const int u = tid % JB, w = tid / JB;
                const double cu = rc[u], su = rs[u], cw = rc[w], sw_ = rs[w];
                if (su != 0.0 || sw_ != 0.0) {
                    int a, b, a2, b2;
                    pair_of(u, k, JP, a, b);
                    pair_of(w, k, JP, a2, b2);
                    const double g11 = G[a][a2], g12 = G[a][b2], g21 = G[b][a2], g22 = G[b][b2];
                    const double t11 = cu * g11 - su * g21, t12 = cu * g12 - su * g22;
                    const double t21 = su * g11 + cu * g21, t22 = su * g12 + cu * g22;
                    G[a][a2] = cw * t11 - sw_ * t12;
                    G[a][b2] = sw_ * t11 + cw * t12;
                    G[b][a2] = cw * t21 - sw_ * t22;
                    G[b][b2] = sw_ * t21 + cw * t22;
                }
            } else if (tid < JB * JB + JB * JP) {  // Q ← Q·J: row x, pair w
                const int e = tid - JB * JB, x = e % JP, w = e / JP;
                const double c = rc[w], s = rs[w];
                if (s != 0.0) {
                    int a, b;
                    pair_of(w, k, JP, a, b);
                    const double qa = Q[x][a], qb = Q[x][b];
                    Q[x][a] = c * qa - s * qb;
                    Q[x][b] = s * qa + c * qb;
                }
            }
        }
        __syncthreads();
        const int any = rotated;
        __syncthreads();  // every thread has read the flag before thread 0 resets it
        if (!any) break;
    }

    // P ← P·Q on W (M rows) and on V (N rows), on the matrix cores: wave w computes the
    // 16×16 tile (UI, UJ) = (w % 8, w / 8) of chunk·Q
    const int UI = wave & 7, UJ = wave >> 3;
    for (int pass = 0; pass < 2; ++pass) {
        double *X = pass == 0 ? W + J.w_off : V + J.v_off;
        const int64_t rows = pass == 0 ? M : N;
        const int64_t nc = (rows + JR - 1) / JR;
        __syncthreads();
        fetch(X, rows, 0);
        for (int64_t ch = 0; ch < nc; ++ch) {
            const int buf = (int)(ch & 1);
            put(buf);
            __syncthreads();
            if (ch + 1 < nc) fetch(X, rows, (ch + 1) * JR);
            d4_t o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int st = 0; st < JP / 4; ++st) {
                const int k = st * 4 + (lane >> 4);
                o = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[buf][16 * UI + (lane & 15)][k], Q[k][16 * UJ + (lane & 15)],
                                                         o, 0, 0, 0);
            }
            const int gc = cols[16 * UJ + (lane & 15)];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = ch * JR + 16 * UI + (lane >> 4) + 4 * v;
                if (gc >= 0 && r < rows) X[(int64_t)gc * rows + r] = o[v];
            }
        }
    }
}

// σ_j = ‖w_j‖ for every plan-wide column (one wave per column).
__global__ __launch_bounds__(64) void jac_norm(const JMat *__restrict__ mats, int64_t nmats,
                                               const double *__restrict__ W, double *__restrict__ sig) {
    __shared__ double red[64];
    const int64_t c = blockIdx.x;
    const JMat &J = mats[find_mat(mats, nmats, c)];
    const double *w = W + J.w_off + (c - J.col0) * J.M;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < J.M; i += 64) s = fma(w[i], w[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 32; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) sig[c] = sqrt(red[0]);
}

// Column j of a matrix goes to position rank(j) (σ descending, ties by index): S, the
// U column or Vt row from w_j/σ_j, and the other factor from v_j.
__global__ __launch_bounds__(JT) void jac_write(const JMat *__restrict__ mats, int64_t nmats,
                                                const double *__restrict__ W, const double *__restrict__ V,
                                                const double *__restrict__ sig, double *__restrict__ U,
                                                double *__restrict__ S, double *__restrict__ Vt) {
    __shared__ int64_t cnt[JT];
    const int64_t c = blockIdx.x;
    const JMat &J = mats[find_mat(mats, nmats, c)];
    const int64_t j = c - J.col0, N = J.N, M = J.M;
    const double sj = sig[c];
    int64_t before = 0;
    for (int64_t q = threadIdx.x; q < N; q += JT) {
        const double sq = sig[J.col0 + q];
        before += (sq > sj || (sq == sj && q < j)) ? 1 : 0;
    }
    cnt[threadIdx.x] = before;
    __syncthreads();
    for (int k = JT / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) cnt[threadIdx.x] += cnt[threadIdx.x + k];
        __syncthreads();
    }
    const int64_t rk = cnt[0];
    const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
    const double *w = W + J.w_off + j * M;
    const double *v = V + J.v_off + j * N;
    if (threadIdx.x == 0) S[J.s_off + rk] = sj;
    if (!J.trans) {
        // A (m×n, m ≥ n) = U·Σ·Vᵀ: U[:, rk] = w/σ (m = M rows), Vt[rk, :] = vᵀ (n = N, ld r = N)
        for (int64_t i = threadIdx.x; i < M; i += JT) U[J.u_off + rk * M + i] = w[i] * inv;
        for (int64_t i = threadIdx.x; i < N; i += JT) Vt[J.vt_off + rk + i * N] = v[i];
    } else {
        // A (m×n, m < n) = V·Σ·(W/σ)ᵀ: U[:, rk] = v (m = N rows), Vt[rk, :] = (w/σ)ᵀ (n = M, ld r = N)
        for (int64_t i = threadIdx.x; i < N; i += JT) U[J.u_off + rk * N + i] = v[i];
        for (int64_t i = threadIdx.x; i < M; i += JT) Vt[J.vt_off + rk + i * N] = w[i] * inv;
    }
}

// Exactly zero σ (they sort last): the corresponding vectors of the w-side factor (U
// columns, or Vt rows when transposed) are replaced, one after another, by unit vectors
// made orthogonal to all earlier ones (classical Gram–Schmidt, twice).  One workgroup
// per matrix; matrices without a zero σ return at once.
__global__ __launch_bounds__(JT) void jac_complete(const JMat *__restrict__ mats, const double *__restrict__ S,
                                                   double *__restrict__ U, double *__restrict__ Vt,
                                                   double *__restrict__ scratch) {
    __shared__ double red[JT];
    const JMat &J = mats[blockIdx.x];
    const int64_t N = J.N, L = J.M;  // N vectors of length L
    int64_t q0 = N;
    while (q0 > 0 && S[J.s_off + q0 - 1] == 0.0) --q0;
    if (q0 == N) return;
    // vector p: U column (ld M) or Vt row (stride r = N)
    auto at = [&](int64_t p, int64_t i) -> double & {
        return J.trans ? Vt[J.vt_off + p + i * N] : U[J.u_off + p * L + i];
    };
    auto reduce = [&](double v) {
        red[threadIdx.x] = v;
        __syncthreads();
        for (int k = JT / 2; k > 0; k >>= 1) {
            if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        return r;
    };
    double *x = scratch + J.w_off;  // W's space of this matrix is free now (M ≥ L entries)
    int64_t cand = 0;
    for (int64_t q = q0; q < N; ++q) {
        for (; cand < L; ++cand) {
            for (int64_t i = threadIdx.x; i < L; i += JT) x[i] = (i == cand) ? 1.0 : 0.0;
            __syncthreads();
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t p = 0; p < q; ++p) {
                    double d = 0.0;
                    for (int64_t i = threadIdx.x; i < L; i += JT) d = fma(at(p, i), x[i], d);
                    d = reduce(d);
                    for (int64_t i = threadIdx.x; i < L; i += JT) x[i] -= d * at(p, i);
                    __syncthreads();
                }
            double nn = 0.0;
            for (int64_t i = threadIdx.x; i < L; i += JT) nn = fma(x[i], x[i], nn);
            nn = sqrt(reduce(nn));
            // some unit vector keeps at least 1/√L of its length (the complement is not
            // empty), so a candidate above half of that always exists
            if (nn > 0.5 / sqrt((double)L)) {
                for (int64_t i = threadIdx.x; i < L; i += JT) at(q, i) = x[i] / nn;
                __syncthreads();
                ++cand;
                break;
            }
        }
    }
}

}  // namespace

struct JacobiWork {
    std::vector<JMat> mats;
    JMat *d_mats = nullptr;
    double *W = nullptr, *V = nullptr, *sig = nullptr, *slots = nullptr;
    int32_t *wg_mat = nullptr, *wg_t = nullptr;
    int64_t ncols = 0, max_wg = 0;
};

int jacobi_create(JacobiWork **out, int64_t nblocks, const int64_t *m, const int64_t *n, const int64_t *a_off,
                  const int64_t *lda, const int64_t *u_off, const int64_t *s_off, const int64_t *vt_off) {
    auto *w = new JacobiWork();
    int64_t wlen = 0, vlen = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t r = std::min(m[b], n[b]);
        if (r <= 0) continue;
        JMat J{};
        J.trans = m[b] < n[b] ? 1 : 0;
        J.M = std::max(m[b], n[b]);
        J.N = r;
        J.w_off = wlen;
        J.v_off = vlen;
        J.a_off = a_off[b];
        J.lda = lda[b];
        J.u_off = u_off[b];
        J.s_off = s_off[b];
        J.vt_off = vt_off[b];
        J.col0 = w->ncols;
        J.nb = (int32_t)((r + JB - 1) / JB);
        J.nbe = J.nb + (J.nb & 1);
        J.tol = std::sqrt((double)J.M) * 2.220446049250313e-16;
        wlen += J.M * J.N;
        vlen += J.N * J.N;
        w->ncols += r;
        w->max_wg += J.nbe / 2;
        w->mats.push_back(J);
    }
    int rc = TKMI_OK;
    if (!w->mats.empty()) {
        if ((rc = upload((void **)&w->d_mats, w->mats.data(), w->mats.size() * sizeof(JMat))) ||
            (rc = upload((void **)&w->W, nullptr, wlen * sizeof(double))) ||
            (rc = upload((void **)&w->V, nullptr, vlen * sizeof(double))) ||
            (rc = upload((void **)&w->sig, nullptr, w->ncols * sizeof(double))) ||
            (rc = upload((void **)&w->slots, nullptr, w->max_wg * sizeof(double))) ||
            (rc = upload((void **)&w->wg_mat, nullptr, w->max_wg * sizeof(int32_t))) ||
            (rc = upload((void **)&w->wg_t, nullptr, w->max_wg * sizeof(int32_t)))) {
            jacobi_destroy(w);
            return rc;
        }
    }
    *out = w;
    return TKMI_OK;
}

void jacobi_destroy(JacobiWork *w) {
    if (!w) return;
    (void)hipFree(w->d_mats);
    (void)hipFree(w->W);
    (void)hipFree(w->V);
    (void)hipFree(w->sig);
    (void)hipFree(w->slots);
    (void)hipFree(w->wg_mat);
    (void)hipFree(w->wg_t);
    delete w;
}

int jacobi_svd(JacobiWork *w, const double *A, double *U, double *S, double *Vt, hipStream_t s, int *sweeps) {
    if (sweeps) *sweeps = 0;
    if (w->mats.empty()) return TKMI_OK;
    const int64_t nm = (int64_t)w->mats.size();
    hipLaunchKernelGGL(jac_init, dim3((unsigned)w->ncols), dim3(JT), 0, s, w->d_mats, nm, A, w->W, w->V);
    TKMI_HIP(hipGetLastError());
    std::vector<char> active(nm, 1);
    std::vector<int32_t> hm, ht;
    std::vector<double> hs;
    int64_t g = 0;
    bool done = false;
    int sw = 0;
    for (; sw < MAX_SWEEPS && !done; ++sw) {
        hm.clear();
        ht.clear();
        int64_t L = 1;
        for (int64_t b = 0; b < nm; ++b) {
            if (!active[b]) continue;
            const JMat &J = w->mats[b];
            L = std::max<int64_t>(L, J.nbe - 1);
            for (int t = 0; t < J.nbe / 2; ++t) {
                hm.push_back((int32_t)b);
                ht.push_back(t);
            }
        }
        const int64_t nwg = (int64_t)hm.size();
        // the previous sweep's kernels are complete (the stream was synchronised), so the
        // tables can be rewritten
        TKMI_HIP(hipMemcpy(w->wg_mat, hm.data(), nwg * sizeof(int32_t), hipMemcpyHostToDevice));
        TKMI_HIP(hipMemcpy(w->wg_t, ht.data(), nwg * sizeof(int32_t), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(jac_reset, dim3((unsigned)std::min<int64_t>((nwg + JT - 1) / JT, 1024)), dim3(JT), 0, s,
                           w->slots, nwg);
        for (int64_t r = 0; r < L; ++r, ++g)
            hipLaunchKernelGGL(jac_round, dim3((unsigned)nwg), dim3(RT), 0, s, w->d_mats, w->wg_mat, w->wg_t, g, w->W,
                               w->V, w->slots);
        TKMI_HIP(hipGetLastError());
        hs.assign(nwg, 0.0);
        TKMI_HIP(hipMemcpyAsync(hs.data(), w->slots, nwg * sizeof(double), hipMemcpyDeviceToHost, s));
        TKMI_HIP(hipStreamSynchronize(s));
        // a matrix is done when no pair of its columns had a cosine above its tolerance over
        // a whole sweep (L ≥ nbe − 1 rounds: every pair was visited)
        std::vector<double> worst(nm, 0.0);
        for (int64_t i = 0; i < nwg; ++i) worst[hm[i]] = std::max(worst[hm[i]], hs[i]);
        done = true;
        for (int64_t b = 0; b < nm; ++b) {
            if (!active[b]) continue;
            if (!(worst[b] > w->mats[b].tol)) active[b] = 0;
            else done = false;
        }
    }
    if (sweeps) *sweeps = sw;
    if (!done) TKMI_FAIL(TKMI_ERR_ARGUMENT, "svd: block Jacobi did not converge in 40 sweeps (non-finite input?)");
    hipLaunchKernelGGL(jac_norm, dim3((unsigned)w->ncols), dim3(64), 0, s, w->d_mats, nm, w->W, w->sig);
    hipLaunchKernelGGL(jac_write, dim3((unsigned)w->ncols), dim3(JT), 0, s, w->d_mats, nm, w->W, w->V, w->sig, U, S,
                       Vt);
    hipLaunchKernelGGL(jac_complete, dim3((unsigned)nm), dim3(JT), 0, s, w->d_mats, S, U, Vt, w->W);
    TKMI_HIP(hipGetLastError());
    return TKMI_OK;
}

}  // namespace tkmi
