// dist_topm.hpp -- candidate filter kernels with a fused per-row top-M.
//
// Rows of A (kNN queries, or KMeans samples) stay with one wave while rows
// of B (the kNN database, or the KMeans centres) stream through LDS in tiles
// of 64.  For each (a, b) pair a lower bound L <= fold(a, b) of the exact
// reference fold is formed and each A row keeps its M smallest (L, b) pairs.
// The exact decision is taken later (finalize kernels) on exact f64 folds of
// the kept candidates, and a certificate proves that nothing outside the
// kept set can enter the answer: every pair that was not kept has
// L >= theta, the M-th kept key, and theta is compared with the exact fold
// of the k-th answer.
//
// Two ways to form the per-pair value:
//   MODE_MFMA64 : ||a||^2 + ||b||^2 - 2 a.b with a.b on v_mfma_f64_16x16x4
//                 and L = value - E(a, b), E a rigorous error bound;
//   MODE_EXACT  : the exact reference fold itself (L = fold, or sqrt(fold)
//                 for kNN), so the kept lists are already exact.
//
// Layout of v_mfma_f64_16x16x4f64 on a 64-lane wave (measured on gfx950
// with one-hot inputs): lane l supplies A[l%16][l/16] and B[l/16][l%16] and
// holds D[l/16 + 4*v][l%16] in accumulator register v = 0..3.
#pragma once

#include "common.hpp"

namespace fl {

// Candidate filters: f64 MFMA GEMM form, exact folds, BF16 one product,
// BF16 three products (dist_bf16.hpp).
enum { MODE_MFMA64 = 0, MODE_EXACT = 1, MODE_BF16 = 2, MODE_BF16X3 = 3, MODE_F64 = 4 };

constexpr int TM_WAVES = 4;
constexpr int TM_ROWS_W = 32;   // A rows per wave (two 16-row MFMA tiles)
constexpr int TM_ROWS = TM_WAVES * TM_ROWS_W;
constexpr int TM_BT = 64;       // B rows per tile (four 16-column MFMA tiles)
constexpr int TM_KC = 32;       // feature chunk staged in LDS
// Staging slots per wave; the exact mode with M = 64 keeps a smaller
// staging area so that the whole block fits in 160 KB of LDS.
constexpr int tm_stage(int M, int MODE) { return (M >= 64 && MODE == 1) ? 128 : 256; }

// Per (split, A row): M candidate keys and B indices (ascending (key, idx)),
// how many are filled, and theta (the M-th key, +inf while not full).
struct Cand {
    double *key;
    int32_t *idx;
    int32_t *cnt;
    double *theta;
    int32_t *theta_idx;
};

struct TopmArgs {
    int64_t nA, nB;
    int d;
    int64_t b_per_split;
    const double *an, *as;   // A rows: squared norm, norm
    const double *bn, *bs;   // B rows: squared norm, norm
    double alpha, beta;      // E = alpha*|a||b| + beta*(|a|^2+|b|^2) + tiny
    int sqrt_key;            // MODE_EXACT: key = sqrt(fold) (kNN ordering)
    Cand out;
};

template <int M, int MODE>
struct TopmShared {
    static constexpr int STAGE = tm_stage(M, MODE);
    double bs[TM_BT][TM_KC + 1];
    double as_[MODE == 1 ? TM_ROWS : 1][TM_KC + 1];   // A chunk, exact mode only
    double lkey[TM_WAVES][TM_ROWS_W][M];
    int32_t lidx[TM_WAVES][TM_ROWS_W][M];
    int32_t lcnt[TM_WAVES][TM_ROWS_W];
    double lth[TM_WAVES][TM_ROWS_W];
    int32_t lthi[TM_WAVES][TM_ROWS_W];
    double skey[TM_WAVES][STAGE];
    int32_t sidx[TM_WAVES][STAGE];
    int32_t srow[TM_WAVES][STAGE];
};

// Row (within the wave's 32 A rows) of accumulator register v of 16-row
// tile i held by this lane.
__device__ __forceinline__ int acc_row(int i, int v, int lane) {
    return 16 * i + (lane >> 4) + 4 * v;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Merge the wave's staged (key, idx, row) entries into the per-row sorted
// lists.  Lane r < 32 owns row r of the wave.
template <int M, int MODE>
__device__ void topm_flush(TopmShared<M, MODE> &sh, int w, int lane, int scnt) {
    wave_sync_lds();
    if (lane < TM_ROWS_W) {
        const int r = lane;
        int c = sh.lcnt[w][r];
        // Scan the staged rows four at a time so that the LDS reads overlap.
        for (int e0 = 0; e0 < scnt; e0 += 4) {
            int rows[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) rows[u] = e0 + u < scnt ? sh.srow[w][e0 + u] : -1;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (rows[u] != r) continue;
                const int e = e0 + u;
                const double k = sh.skey[w][e];
                const int32_t ix = sh.sidx[w][e];
                if (c == M && !pair_less(k, ix, sh.lkey[w][r][M - 1], sh.lidx[w][r][M - 1]))
                    continue;
                int p = c < M ? c++ : M - 1;
                while (p > 0 && pair_less(k, ix, sh.lkey[w][r][p - 1], sh.lidx[w][r][p - 1])) {
                    sh.lkey[w][r][p] = sh.lkey[w][r][p - 1];
                    sh.lidx[w][r][p] = sh.lidx[w][r][p - 1];
                    --p;
                }
                sh.lkey[w][r][p] = k;
                sh.lidx[w][r][p] = ix;
            }
        }
        sh.lcnt[w][r] = c;
        if (c == M) {
            sh.lth[w][r] = sh.lkey[w][r][M - 1];
            sh.lthi[w][r] = sh.lidx[w][r][M - 1];
        }
    }
    wave_sync_lds();
}

template <typename TA, typename TB, int M, int MODE>
__global__ __launch_bounds__(256) void dist_topm_kernel(const TA *__restrict__ A,
                                                        const TB *__restrict__ B,
                                                        TopmArgs args) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef TopmShared<M, MODE> Sh;
    Sh &sh = *reinterpret_cast<Sh *>(smem_raw);

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int d = args.d;
    const int64_t nA = args.nA;
    const int64_t blk_row0 = (int64_t)blockIdx.x * TM_ROWS;
    const int64_t row0 = blk_row0 + w * TM_ROWS_W;
    const int split = blockIdx.y;
    const int64_t bbeg = (int64_t)split * args.b_per_split;
    int64_t bend = bbeg + args.b_per_split;
    if (bend > args.nB) bend = args.nB;

    if (lane < TM_ROWS_W) {
        sh.lcnt[w][lane] = 0;
        sh.lth[w][lane] = __builtin_huge_val();
        sh.lthi[w][lane] = 0x7fffffff;
    }

    // The 8 A rows this lane's accumulators belong to (acc_row).
    double a_n[2][4], a_s[2][4];
    bool a_ok[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row0 + acc_row(i, v, lane);
            a_ok[i][v] = r < nA;
            a_n[i][v] = a_ok[i][v] ? args.an[r] : 0.0;
            a_s[i][v] = a_ok[i][v] ? args.as[r] : 0.0;
        }

    // The streamed rows are staged one (64-row tile, 32-feature chunk) step
    // at a time; each thread fetches its 8 values of the next step into
    // registers while the current step is multiplied.
    const int nkc = (d + TM_KC - 1) / TM_KC;
    const int64_t nsteps = (bend > bbeg ? ceil_div(bend - bbeg, TM_BT) : 0) * nkc;
    constexpr int PF = TM_BT * TM_KC / 256;
    constexpr int KQ = TM_KC / 4;   // MFMA k-steps per chunk
    TB pf[PF];
    double anx[KQ][2], acur[KQ][2];   // this lane's A fragments, next / current chunk
    auto fetch = [&](int64_t step) {
        // Branch-free: indices past the end read a valid element; they are
        // zeroed where the values are consumed, so that nothing here waits
        // on a load.
        const int64_t t0 = bbeg + (step / nkc) * TM_BT;
        const int k0 = (int)(step % nkc) * TM_KC;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int t = threadIdx.x + 256 * u;
            const int r = t / TM_KC, c = t % TM_KC;
            const int64_t gb = t0 + r;
            const int gk = k0 + c;
            const bool ok = gb < bend && gk < d;
            pf[u] = B[(ok ? gb : bbeg) * d + (ok ? gk : 0)];
        }
        if (MODE == MODE_MFMA64) {
#pragma unroll
            for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int64_t r = row0 + 16 * i + (lane & 15);
                    const int gk = k0 + 4 * kq + (lane >> 4);
                    const bool ok = r < nA && gk < d;
                    anx[kq][i] = (double)A[(ok ? r : 0) * d + (ok ? gk : 0)];
                }
        }
    };
    wait_vm_loads();   // prologue loads are in registers before the loop
    if (nsteps > 0) fetch(0);
    double4_t acc[2][4];
    // Staged candidates persist across tiles (see dist_bf16.hpp).
    int scnt = 0;

    for (int64_t step = 0; step < nsteps; ++step) {
        const int64_t b0 = bbeg + (step / nkc) * TM_BT;
        const int k0 = (int)(step % nkc) * TM_KC;
        if (k0 == 0) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
        }
        {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int t = threadIdx.x + 256 * u;
                const bool ok = b0 + t / TM_KC < bend && k0 + t % TM_KC < d;
                sh.bs[t / TM_KC][t % TM_KC] = ok ? (double)pf[u] : 0.0;
            }
            if (MODE == MODE_EXACT) {
                for (int t = threadIdx.x; t < TM_ROWS * TM_KC; t += 256) {
                    const int r = t / TM_KC, c = t % TM_KC;
                    const int64_t ga = blk_row0 + r;
                    const int gk = k0 + c;
                    sh.as_[r][c] = (ga < nA && gk < d) ? (double)A[ga * d + gk] : 0.0;
                }
            }
            __syncthreads();
            if (MODE == MODE_MFMA64) {
#pragma unroll
                for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const bool ok = row0 + 16 * i + (lane & 15) < nA &&
                                        k0 + 4 * kq + (lane >> 4) < d;
                        acur[kq][i] = ok ? anx[kq][i] : 0.0;
                    }
            }
            if (step + 1 < nsteps) fetch(step + 1);
            const int kc = d - k0 < TM_KC ? d - k0 : TM_KC;
            if (MODE == MODE_MFMA64) {
                // Features past d are zero in both operands, so all KQ steps
                // can run.
#pragma unroll
                for (int kq = 0; kq < KQ; ++kq) {
                    const int kl = 4 * kq + (lane >> 4);
                    double bf[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) bf[j] = sh.bs[16 * j + (lane & 15)][kl];
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                acur[kq][i], bf[j], acc[i][j], 0, 0, 0);
                }
            } else {
                // Exact fold, feature by feature in ascending order.
                for (int kk = 0; kk < kc; ++kk) {
                    double bv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) bv[j] = sh.bs[16 * j + (lane & 15)][kk];
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const double av = sh.as_[w * TM_ROWS_W + acc_row(i, v, lane)][kk];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const double t = av - bv[j];
                                acc[i][j][v] = acc[i][j][v] + t * t;
                            }
                        }
                }
            }
        }
        if (k0 + TM_KC < d) continue;

        // ---- epilogue: keys, filter against theta, stage, merge ----
        double th[2][4];
        int32_t thi[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                th[i][v] = sh.lth[w][acc_row(i, v, lane)];
                thi[i][v] = sh.lthi[w][acc_row(i, v, lane)];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t col = b0 + 16 * j + (lane & 15);
            const bool cok = col < bend;
            double nb = 0.0, sb = 0.0;
            if (MODE == MODE_MFMA64 && cok) {
                nb = args.bn[col];
                sb = args.bs[col];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rl = acc_row(i, v, lane);
                    const bool ok = cok && a_ok[i][v];
                    double key;
                    if (MODE == MODE_MFMA64) {
                        const double s = a_n[i][v] + nb;
                        const double val = s - 2.0 * acc[i][j][v];
                        const double e =
                            args.alpha * (a_s[i][v] * sb) + args.beta * s + 1e-300;
                        key = val - e;
                    } else {
                        key = args.sqrt_key ? sqrt(acc[i][j][v]) : acc[i][j][v];
                    }
                    bool pass = ok && pair_less(key, (int32_t)col, th[i][v], thi[i][v]);
                    uint64_t mask = __ballot(pass);
                    if (mask == 0) continue;
                    int np = __popcll(mask);
                    if (scnt + np > Sh::STAGE) {
                        topm_flush<M, MODE>(sh, w, lane, scnt);
                        scnt = 0;
#pragma unroll
                        for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
                            for (int v2 = 0; v2 < 4; ++v2) {
                                th[i2][v2] = sh.lth[w][acc_row(i2, v2, lane)];
                                thi[i2][v2] = sh.lthi[w][acc_row(i2, v2, lane)];
                            }
                        pass = ok && pair_less(key, (int32_t)col, th[i][v], thi[i][v]);
                        mask = __ballot(pass);
                        np = __popcll(mask);
                    }
                    if (pass) {
                        const int pos = scnt + __popcll(mask & ((1ull << lane) - 1ull));
                        sh.skey[w][pos] = key;
                        sh.sidx[w][pos] = (int32_t)col;
                        sh.srow[w][pos] = rl;
                    }
                    scnt += np;
                }
        }
    }
    if (scnt > 0) topm_flush<M, MODE>(sh, w, lane, scnt);

    wave_sync_lds();
    if (lane < TM_ROWS_W) {
        const int64_t r = row0 + lane;
        if (r < nA) {
            const int64_t o = (int64_t)split * nA + r;
            const int c = sh.lcnt[w][lane];
            args.out.cnt[o] = c;
            // theta: the M-th key; +inf when the list kept every row of the
            // split; -inf when rows are missing (a key that never compared
            // below +inf, i.e. overflow or NaN), which fails the certificate.
            args.out.theta[o] = c == M ? sh.lth[w][lane]
                                       : (c >= bend - bbeg ? __builtin_huge_val()
                                                           : -__builtin_huge_val());
            args.out.theta_idx[o] = c == M ? sh.lthi[w][lane] : 0x7fffffff;
            for (int m = 0; m < M; ++m) {
                args.out.key[o * M + m] = m < c ? sh.lkey[w][lane][m] : __builtin_huge_val();
                args.out.idx[o * M + m] = m < c ? sh.lidx[w][lane][m] : -1;
            }
        }
    }
}

// Squared norm and norm of each row (f64, any summation order: the error
// bound in the filter allows for it).
template <typename T>
__global__ __launch_bounds__(256) void row_norms_kernel(const T *__restrict__ X, int64_t n,
                                                        int d, double *__restrict__ n2,
                                                        double *__restrict__ n1) {
    // 256 rows per block (launched with 256 threads): the rows are staged
    // through LDS CW columns at a time with the block's loads running
    // along rows, and each thread folds its own row in column order
    // (s = s + v*v, one IEEE op each), the same sum as a loop over the row.
    constexpr int CW = sizeof(T) == 4 ? 32 : 16;
    __shared__ T tile[256 * (CW + 1)];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t i = r0 + t;
    double s = 0.0;
    for (int c0 = 0; c0 < d; c0 += CW) {
        const int cw = d - c0 < CW ? d - c0 : CW;
        __syncthreads();
        for (int e = t; e < 256 * CW; e += 256) {
            const int r = e / CW, c = e % CW;
            if (c < cw && r0 + r < n) tile[r * (CW + 1) + c] = X[(r0 + r) * d + c0 + c];
        }
        __syncthreads();
        if (i < n)
            for (int c = 0; c < cw; ++c) {
                const double v = (double)tile[t * (CW + 1) + c];
                s = s + v * v;
            }
    }
    if (i < n) {
        n2[i] = s;
        n1[i] = sqrt(s);
    }
}

template <typename T>
__global__ void to_f64_kernel(const T *__restrict__ X, int64_t count, double *__restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) Y[i] = (double)X[i];
}

// ------------------------------------------------------------ kNN finalize --
// One wave per query: gather the S*M candidates, fold them exactly in
// increasing key order until the smallest key left exceeds the k-th fold
// (keys are lower bounds of the folds, so no candidate left can enter), order
// by (sqrt(fold), index) -- the reference distance (kdtree.rs:342-348,
// :386-388) with index as the tie rule -- take the first k, and certify.
//
// Certificate: every pair not kept has key >= theta_s for its split s, and
// key <= fold.  With T the exact fold of the k-th answer, a pair outside
// could only enter if its sqrt(fold) <= sqrt(T); correctly rounded sqrt
// makes that impossible once fold > T*(1 + 2^-50), so theta_min > T*(1+2^-48)
// certifies.  A split whose lists may miss rows reports theta = -inf.  A NaN
// fold also leaves the query uncertified.  Uncertified queries are listed for
// the next, tighter filter.
constexpr int FIN_MAXC = 512;        // candidates per query, usual launches
constexpr int FIN_MAXC_BIG = 4096;   // a few queries over many splits (fallbacks)

template <typename TB, int CAP>
__global__ __launch_bounds__(64) void knn_finalize_kernel(
    const double *__restrict__ Q, const TB *__restrict__ B, int d, int64_t nq, int64_t nB,
    int S, int M, int64_t b_per_split, Cand cand, int k, int certify,
    double *__restrict__ dist_out, int64_t *__restrict__ idx_out, int32_t *__restrict__ flag_list,
    int32_t *__restrict__ flag_count, unsigned long long *__restrict__ n_cand) {
    __shared__ double ck[CAP];
    __shared__ double cf[CAP];
    __shared__ double cd[CAP];
    __shared__ int32_t ci[CAP];
    __shared__ int32_t rk[CAP];
    __shared__ double kth_fold;
    __shared__ double next_key;
    __shared__ int kth_found;
    __shared__ int has_nan;
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    if (q >= nq) return;

    // theta per split: -inf flags a split whose lists miss rows.
    bool complete = true;
    double thmin = __builtin_huge_val();
    for (int s = 0; s < S; ++s) {
        const double th = cand.theta[(int64_t)s * nq + q];
        if (th == -__builtin_huge_val()) complete = false;
        thmin = fmin(thmin, th);
    }

    // compact the valid candidates into LDS
    int nc = 0;
    const int total = S * M;
    for (int base = 0; base < total; base += 64) {
        const int e = base + lane;
        int32_t ix = -1;
        double kv = 0.0;
        if (e < total) {
            const int s = e / M, m = e % M;
            const int64_t o = ((int64_t)s * nq + q) * M + m;
            ix = cand.idx[o];
            kv = cand.key[o];
        }
        const bool v = ix >= 0;
        const uint64_t mask = __ballot(v);
        if (v) {
            const int at = nc + __popcll(mask & ((1ull << lane) - 1ull));
            ci[at] = ix;
            ck[at] = kv;
        }
        nc += __popcll(mask);
    }
    if (lane == 0) {
        kth_found = 0;
        has_nan = 0;
        next_key = __builtin_huge_val();
    }
    __syncthreads();

    // The exact form's keys are the reference distances sqrt(fold) and its
    // candidate indices are distinct: the answer is the k smallest
    // (distance, index) pairs, taken by k wave-wide minimum passes, each
    // above the previous pair.  A NaN key flags the query (the host raises).
    if (!certify) {
        bool nan = false;
        for (int e = lane; e < nc; e += 64) nan |= ck[e] != ck[e];
        if (__ballot(nan) != 0 || !complete || nc < k) {
            if (lane == 0) flag_list[atomicAdd(flag_count, 1)] = (int32_t)q;
            return;
        }
        double lk = -__builtin_huge_val();
        int64_t li = -1;
        for (int j = 0; j < k; ++j) {
            double bk = __builtin_huge_val();
            int64_t bi = INT64_MAX;
            for (int e = lane; e < nc; e += 64)
                if (pair_less(lk, li, ck[e], ci[e]) && pair_less(ck[e], ci[e], bk, bi)) {
                    bk = ck[e];
                    bi = ci[e];
                }
            for (int o = 32; o > 0; o >>= 1) {
                const double ok2 = __shfl_xor(bk, o);
                const int64_t oi2 = __shfl_xor(bi, o);
                if (pair_less(ok2, oi2, bk, bi)) {
                    bk = ok2;
                    bi = oi2;
                }
            }
            if (lane == 0) {
                dist_out[q * k + j] = bk;
                idx_out[q * k + j] = bi;
            }
            lk = bk;
            li = bi;
        }
        if (lane == 0) atomicAdd(n_cand, (unsigned long long)nc);
        return;
    }

    // Filter keys are lower bounds of the exact folds, so the candidates are
    // re-scored in increasing key order: first the k+6 smallest; the rest
    // only when the smallest key left could still reach the k-th fold.
    // Ranks by (key, index) of the k+7 smallest candidates, by wave-wide
    // minimum passes (each above the previous pair); the others get
    // INT32_MAX and are ranked in full only if the lazy step below needs them.
    for (int e = lane; e < nc; e += 64) rk[e] = INT32_MAX;
    {
        const int want = nc < k + 7 ? nc : k + 7;
        double lk = -__builtin_huge_val();
        int64_t li = -1;
        for (int j = 0; j < want; ++j) {
            double bk = __builtin_huge_val();
            int64_t bi = INT64_MAX;
            int be = -1;
            for (int e = lane; e < nc; e += 64)
                if (pair_less(lk, li, ck[e], ci[e]) && (be < 0 || pair_less(ck[e], ci[e], bk, bi))) {
                    bk = ck[e];
                    bi = ci[e];
                    be = e;
                }
            for (int o = 32; o > 0; o >>= 1) {
                const double ok2 = __shfl_xor(bk, o);
                const int64_t oi2 = __shfl_xor(bi, o);
                const int oe2 = __shfl_xor(be, o);
                if (oe2 >= 0 && (be < 0 || pair_less(ok2, oi2, bk, bi))) {
                    bk = ok2;
                    bi = oi2;
                    be = oe2;
                }
            }
            if (be < 0) break;
            if (lane == 0) rk[be] = j;
            lk = bk;
            li = bi;
        }
    }
    __syncthreads();
    const double *qr = Q + q * d;
    auto rescore = [&](int lo, int hi) {
        for (int e = lane; e < nc; e += 64) {
            if (rk[e] < lo || rk[e] >= hi) continue;
            const double f = fold_exact(qr, B + (int64_t)ci[e] * d, d);
            cf[e] = f;
            cd[e] = sqrt(f);
            if (f != f) has_nan = 1;
        }
        __syncthreads();
    };
    // k-th of the re-scored set (rank < R) by (distance, index).
    auto kth = [&](int R) {
        for (int e = lane; e < nc; e += 64) {
            if (rk[e] >= R) continue;
            int rank = 0;
            for (int e2 = 0; e2 < nc; ++e2)
                rank += (rk[e2] < R && pair_less(cd[e2], ci[e2], cd[e], ci[e])) ? 1 : 0;
            if (rank == k - 1) {
                kth_fold = cf[e];
                kth_found = 1;
            }
        }
        for (int e = lane; e < nc; e += 64)
            if (rk[e] == R) next_key = ck[e];
        __syncthreads();
    };
    int R = certify ? (nc < k + 6 ? nc : k + 6) : nc;
    rescore(0, R);
    kth(R);
    if (R < nc && !(kth_found && !has_nan && next_key > kth_fold + kth_fold * 0x1.0p-48)) {
        // full ranks (rare): every candidate is re-scored
        for (int e = lane; e < nc; e += 64) {
            int r = 0;
            for (int e2 = 0; e2 < nc; ++e2) r += pair_less(ck[e2], ci[e2], ck[e], ci[e]) ? 1 : 0;
            rk[e] = r;
        }
        __syncthreads();
        rescore(R, nc);
        if (lane == 0) kth_found = 0;
        __syncthreads();
        R = nc;
        kth(R);
    }
    bool ok = complete && (nc >= k) && kth_found && !has_nan;
    if (ok && certify && thmin != __builtin_huge_val()) ok = thmin > kth_fold + kth_fold * 0x1.0p-48;
    if (!ok) {
        if (lane == 0) flag_list[atomicAdd(flag_count, 1)] = (int32_t)q;
        return;
    }
    for (int e = lane; e < nc; e += 64) {
        if (rk[e] >= R) continue;
        int rank = 0;
        for (int e2 = 0; e2 < nc; ++e2)
            rank += (rk[e2] < R && pair_less(cd[e2], ci[e2], cd[e], ci[e])) ? 1 : 0;
        if (rank < k) {
            dist_out[q * k + rank] = cd[e];
            idx_out[q * k + rank] = ci[e];
        }
    }
    if (lane == 0) atomicAdd(n_cand, (unsigned long long)R);
}

// -------------------------------------------------------- kNN epilogues --
// knn.rs:388-440 (class_score_vec + weighted_vote): neighbour order, strict
// `>` over sorted classes so the smallest class position wins a tie.
static __global__ __attribute__((unused)) void knn_vote_kernel(const double *__restrict__ dist, const int64_t *__restrict__ idx,
                                int64_t nq, int k, const int32_t *__restrict__ ypos,
                                const int64_t *__restrict__ classes, int weights,
                                int64_t *__restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double *dq = dist + q * k;
    const int64_t *iq = idx + q * k;
    bool zero = false;
    if (weights) {
        for (int m = 0; m < k; ++m) zero = zero || dq[m] < 1e-15;
    }
    int best_c = 0x7fffffff;
    double best_s = -1.0;
    for (int m = 0; m < k; ++m) {
        const int c = ypos[iq[m]];
        bool seen = false;
        for (int m2 = 0; m2 < m; ++m2) seen = seen || ypos[iq[m2]] == c;
        if (seen) continue;
        double s = 0.0;
        for (int m2 = 0; m2 < k; ++m2) {
            if (ypos[iq[m2]] != c) continue;
            if (!weights) s = s + 1.0;
            else if (zero) {
                if (dq[m2] < 1e-15) s = s + 1.0;
            } else s = s + 1.0 / dq[m2];
        }
        if (s > best_s || (s == best_s && c < best_c)) {
            best_s = s;
            best_c = c;
        }
    }
    // Classes with no neighbour score 0, and the reference's scan starts at
    // class position 0: when no score is positive, position 0 wins.
    if (!(best_s > 0.0)) best_c = 0;
    out[q] = classes[best_c];
}

// knn.rs:452-486 (predict_proba): scores in class order, total by a left
// fold over classes, uniform when the total is not positive.
static __global__ __attribute__((unused)) void knn_proba_kernel(const double *__restrict__ dist, const int64_t *__restrict__ idx,
                                 int64_t nq, int k, const int32_t *__restrict__ ypos,
                                 int n_classes, int weights, double *__restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double *dq = dist + q * k;
    const int64_t *iq = idx + q * k;
    bool zero = false;
    if (weights) {
        for (int m = 0; m < k; ++m) zero = zero || dq[m] < 1e-15;
    }
    double *o = out + q * n_classes;
    double total = 0.0;
    for (int c = 0; c < n_classes; ++c) {
        double s = 0.0;
        for (int m = 0; m < k; ++m) {
            if (ypos[iq[m]] != c) continue;
            if (!weights) s = s + 1.0;
            else if (zero) {
                if (dq[m] < 1e-15) s = s + 1.0;
            } else s = s + 1.0 / dq[m];
        }
        o[c] = s;
        total = total + s;
    }
    if (total > 0.0) {
        for (int c = 0; c < n_classes; ++c) o[c] = o[c] / total;
    } else {
        const double u = 1.0 / (double)n_classes;
        for (int c = 0; c < n_classes; ++c) o[c] = u;
    }
}

// knn.rs:952-988 (weighted_mean), neighbour order.
static __global__ __attribute__((unused)) void knn_mean_kernel(const double *__restrict__ dist, const int64_t *__restrict__ idx,
                                int64_t nq, int k, const double *__restrict__ y, int weights,
                                double *__restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double *dq = dist + q * k;
    const int64_t *iq = idx + q * k;
    if (!weights) {
        double s = 0.0;
        for (int m = 0; m < k; ++m) s = s + y[iq[m]];
        out[q] = s / (double)k;
        return;
    }
    int nz = 0;
    for (int m = 0; m < k; ++m) nz += dq[m] < 1e-15 ? 1 : 0;
    if (nz > 0) {
        double s = 0.0;
        for (int m = 0; m < k; ++m)
            if (dq[m] < 1e-15) s = s + y[iq[m]];
        out[q] = s / (double)nz;
        return;
    }
    double ws = 0.0, wt = 0.0;
    for (int m = 0; m < k; ++m) {
        const double wv = 1.0 / dq[m];
        ws = ws + wv * y[iq[m]];
        wt = wt + wv;
    }
    out[q] = ws / wt;
}

}  // namespace fl
