// batched.hip -- many KMeans problems and all their restarts on the device
// at once.
//
// Reference: each problem is one KMeans::fit (ferrolearn-cluster/src/
// kmeans.rs:403-513): n_init runs of k-means++ (:169-253) then Lloyd
// iterations assign (:267-352) -> recompute centroids (:358-401) -> swap ->
// n_iter = iter + 1 -> stop when max_shift < tol, keeping the run with the
// strictly smallest inertia.  Every (problem, restart) run here gives the
// same labels, n_iter, centres and inertia, bit for bit, as
// fl_kmeans_fit_device on that problem alone.
//
// Layout of the work:
//  * k-means++ runs in lockstep over all runs, one centre step at a time
//    (draw -> search -> trials -> select), with each run's generator state
//    on the device, so no host round trip happens per centre.
//  * One Lloyd iteration is two kernels, captured once in a HIP graph and
//    replayed: a persistent assignment kernel that takes (problem, row tile)
//    items from a work queue and scores each row exactly against the
//    centres of every active restart of its problem (the row is loaded once
//    for all restarts), then an update kernel with one block per run
//    (inertia, reference-order centroid sums, shift, convergence).  Runs
//    that converged are skipped by both; the host looks at the convergence
//    flags every few replays.
//  * The assignment tile shape (rows per lane R, centres per step U) is
//    chosen per (d, k) bucket from a table that the autotuner below fills.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host.hpp"
#include "kmeans_kernels.hpp"

using namespace fl;

namespace fl {
fl_status validate_kmeans_fit(int64_t n, int64_t d, const fl_kmeans_params *p, int32_t dtype);
}

namespace {

// ------------------------------------------------------------ tables ----
constexpr int BK_MAXD = 64;       // features handled by the batched path
constexpr int BK_MAXK = 1024;     // clusters handled by the batched path
constexpr int BK_UPD_WAVES = 8;   // update kernel: waves per run block
constexpr int BK_MAXT = 8;        // k-means++ trials: 2 + floor(ln 1024)

struct BkRun {
    const void *X;      // the problem's rows, n x d (f32 or f64)
    double *cA, *cB;    // centres, two buffers of k x d (current one: BkState::cur_b)
    int32_t *labels;    // n
    double *best;       // n: fold to the assigned centre
    int32_t *order;     // n: rows grouped by label, ascending index within a label
    double *part;       // ceil(n / SUM_BLK) inertia block sums
    int32_t *cnt;       // k: rows per cluster
    int64_t *off;       // k: first position of each cluster in order
    double *pool;       // k-means++ distances: nt + 1 slots of n
    double *bsum;       // k-means++ block sums: nt + 1 slots of nblk
    int64_t n, nblk;    // rows; k-means++ blocks of KPP_BLK rows
    int32_t d, k, f32, nt;
    int32_t max_iter, kpp;   // kpp: 1 = k-means++ init, 0 = given centres in cA
    double tol;
    uint64_t seed;
};

struct BkState {
    uint64_t rng[4];
    double total;       // k-means++: sum of the current minimum distances
    double inertia, shift;
    int32_t it, cur_b, mdslot, skip;
    int64_t cand[BK_MAXT];
    double thr[BK_MAXT];
};

struct BkProb {
    const void *X;
    int64_t n;
    int32_t d, k, f32, nruns, run0, variant;
};

struct BkItem {
    int32_t prob, pad;
    int64_t row0;
};

// Assignment tile shapes: D = feature registers (d <= D), U = centres
// folded side by side, R = rows per lane.  A problem's variant is
// g_tuning[D bucket][k bucket].
struct BkVariant {
    int D, U, R;
};
constexpr BkVariant kVariants[] = {
    {8, 4, 1},  {8, 8, 1},  {8, 4, 2},  {8, 8, 2},  {16, 4, 1}, {16, 8, 1}, {16, 4, 2},
    {16, 8, 2}, {32, 4, 1}, {32, 8, 1}, {32, 4, 2}, {64, 4, 1}, {64, 8, 1}, {64, 2, 1},
};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
constexpr int kDBuckets = 4;   // D = 8, 16, 32, 64
constexpr int kKBuckets = 4;   // k <= 16, <= 64, <= 256, <= 1024

__host__ __device__ inline int d_bucket(int d) { return d <= 8 ? 0 : d <= 16 ? 1 : d <= 32 ? 2 : 3; }
int k_bucket(int k) { return k <= 16 ? 0 : k <= 64 ? 1 : k <= 256 ? 2 : 3; }

std::mutex g_tune_mu;
int g_tuning[kDBuckets][kKBuckets] = {
    {1, 1, 3, 3}, {5, 5, 7, 7}, {9, 9, 9, 9}, {12, 12, 12, 12},
};

int variant_of(int d, int k) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    return g_tuning[d_bucket(d)][k_bucket(k)];
}

// ---------------------------------------------------------- generator ----
// The library's generator (host.hpp Rng) on the device: SplitMix64 seeding
// of xoshiro256**; range(n) = high word of next() * n; f64() = (next() >>
// 11) * 2^-53.  Draw for draw the same values as the host.
__device__ __forceinline__ void rng_seed(uint64_t (&s)[4], uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}
__device__ __forceinline__ uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
__device__ __forceinline__ uint64_t rng_next(uint64_t (&s)[4]) {
    const uint64_t out = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}
__device__ __forceinline__ int64_t rng_range(uint64_t (&s)[4], int64_t n) {
    return (int64_t)__umul64hi(rng_next(s), (uint64_t)n);
}
__device__ __forceinline__ double rng_f64(uint64_t (&s)[4]) {
    return (double)(rng_next(s) >> 11) * 0x1.0p-53;
}

// Pointers kept in the run tables are generic, so loads through them are
// vector loads even at a wave-uniform address.  Data that no thread writes
// during the kernel (centres, candidate rows) is read through the constant
// address space instead, which turns uniform loads into scalar loads.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

__device__ __forceinline__ double ldx(const void *X, int f32, int64_t i) {
    return f32 ? (double)((const float *)X)[i] : ((const double *)X)[i];
}

__device__ __forceinline__ void copy_row(const BkRun &R, int64_t row, double *dst) {
    for (int j = 0; j < R.d; ++j) dst[j] = ldx(R.X, R.f32, row * R.d + j);
}

// ----------------------------------------------------------- k-means++ ----
// The steps of Lloyd::kmeanspp (kmeans.hip) for every run at once, with the
// same draws, two-level sums (sequential within KPP_BLK-row blocks, then
// over the blocks), searches and strict-< choice of the trial.

// One thread per run: generator, first centre, convergence state.
__global__ void bk_init_kernel(const BkRun *__restrict__ runs, BkState *__restrict__ st,
                               int32_t *__restrict__ done, int nruns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    const BkRun R = runs[r];
    BkState &S = st[r];
    uint64_t s[4];
    rng_seed(s, R.seed);
    S.it = 0;
    S.cur_b = 0;
    S.mdslot = 0;
    S.skip = 1;
    S.total = 0.0;
    S.inertia = 1.7976931348623157e308;
    S.shift = 0.0;
    done[r] = R.max_iter <= 0 ? 1 : 0;
    if (R.kpp) {
        const int64_t first = rng_range(s, R.n);
        copy_row(R, first, R.cA);
    }
    for (int i = 0; i < 4; ++i) S.rng[i] = s[i];
}

// Sum of nblk block sums in block order, as one wave (every lane returns it).
__device__ __forceinline__ double wave_seq_sum(const double *__restrict__ v, int64_t m, int lane) {
    double acc = 0.0;
    for (int64_t b0 = 0; b0 < m; b0 += 64) {
        const double x = b0 + lane < m ? v[b0 + lane] : 0.0;
        const int c = m - b0 < 64 ? (int)(m - b0) : 64;
        for (int u = 0; u < c; ++u) acc = acc + wave_bcast(x, u);
    }
    return acc;
}

// One wave per run: the first total.
__global__ __launch_bounds__(64) void bk_kpp_total_kernel(const BkRun *__restrict__ runs,
                                                          BkState *__restrict__ st, int nruns) {
    const int r = blockIdx.x;
    if (r >= nruns) return;
    const BkRun R = runs[r];
    if (!R.kpp || R.k == 1) return;
    const double t = wave_seq_sum(R.bsum, R.nblk, threadIdx.x);
    if (threadIdx.x == 0) st[r].total = t;
}

__device__ __forceinline__ int trial_slot(int t, int mdslot) { return t < mdslot ? t : t + 1; }

// One thread per run, centre c: thresholds, or a uniformly drawn row when the
// potential is zero (kmeans.rs:199-204).
__global__ void bk_kpp_draw_kernel(const BkRun *__restrict__ runs, BkState *__restrict__ st,
                                   const int32_t *__restrict__ rlist, int nruns, int c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nruns) return;
    const int r = rlist[i];
    const BkRun R = runs[r];
    BkState &S = st[r];
    if (!R.kpp || c >= R.k) {
        S.skip = 1;
        return;
    }
    uint64_t s[4] = {S.rng[0], S.rng[1], S.rng[2], S.rng[3]};
    if (S.total <= 0.0) {
        const int64_t idx = rng_range(s, R.n);
        copy_row(R, idx, R.cA + (int64_t)c * R.d);
        S.skip = 1;
    } else {
        for (int t = 0; t < R.nt; ++t) S.thr[t] = rng_f64(s) * S.total;
        S.skip = 0;
    }
    for (int i = 0; i < 4; ++i) S.rng[i] = s[i];
}

// One wave per (run, trial): the first row whose two-level running sum of
// the minimum distances reaches the threshold (n - 1 if none).
__global__ __launch_bounds__(64) void bk_kpp_search_kernel(const BkRun *__restrict__ runs,
                                                           BkState *__restrict__ st,
                                                           const int32_t *__restrict__ rlist,
                                                           int nruns) {
    __shared__ double blk[KPP_BLK];
    const int i = blockIdx.x / BK_MAXT, t = blockIdx.x % BK_MAXT;
    const int lane = threadIdx.x;
    if (i >= nruns) return;
    const int r = rlist[i];
    BkState &S = st[r];
    const BkRun R = runs[r];
    if (S.skip || t >= R.nt) return;
    const double th = S.thr[t];
    const double *md = R.pool + (int64_t)S.mdslot * R.n;
    const double *bs = R.bsum + (int64_t)S.mdslot * R.nblk;
    double run = 0.0;
    int64_t res = R.n - 1;
    bool found = false;
    for (int64_t b0 = 0; b0 < R.nblk && !found; b0 += 64) {
        const double sv = b0 + lane < R.nblk ? bs[b0 + lane] : 0.0;
        const int m = R.nblk - b0 < 64 ? (int)(R.nblk - b0) : 64;
        for (int u = 0; u < m && !found; ++u) {
            const double s = wave_bcast(sv, u);
            const int64_t b = b0 + u;
            if (run + s >= th) {
                // the block's values in LDS (every load in flight at once),
                // then lane 0 runs the sequential sum
                const int64_t lo = b * KPP_BLK;
                const int len = (int)((lo + KPP_BLK < R.n ? lo + KPP_BLK : R.n) - lo);
                for (int e = lane; e < len; e += 64) blk[e] = md[lo + e];
                wave_sync_lds();
                int hit = -1;
                if (lane == 0) {
                    double c = run;
                    int e = 0;
                    for (; e + 8 <= len && hit < 0; e += 8) {
                        double v[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) v[q] = blk[e + q];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            c = c + v[q];
                            if (hit < 0 && c >= th) hit = e + q;
                        }
                    }
                    for (; e < len && hit < 0; ++e) {
                        c = c + blk[e];
                        if (c >= th) hit = e;
                    }
                }
                hit = __shfl(hit, 0);
                wave_sync_lds();
                if (hit >= 0) {
                    res = lo + hit;
                    found = true;
                }
            }
            run = run + s;
        }
    }
    if (lane == 0) S.cand[t] = res;
}

// One wave's share of a k-means++ distance pass over rows [lo, hi) of one
// KPP_BLK block.  FIRST: md = fold(x, c0) (c0 = the first centre) into slot
// 0.  Otherwise, for every trial t: nd = min(md, fold(x, cand_t)) (`if d <
// nd`, kmeans.rs:226-228) into the trial's slot.  Each lane folds its row
// against NT candidate rows at once (features outermost, one accumulator
// per candidate, so each feature value is loaded once); the candidates sit
// in this wave's LDS slice, padded to D features and to NT rows with +0
// (padded features add +0 to every partial sum; padded rows are computed
// and dropped).  The block sums are the sequential folds of the values in
// row order: each chunk's 64 values pass through LDS and lane t adds trial
// t's in order.
template <typename TX, int D, int NT, bool FIRST>
__device__ __forceinline__ void bk_kpp_pass(const BkRun &R, const BkState &S, int b,
                                            double (*lds)[64], double *__restrict__ cs) {
    const int lane = threadIdx.x & 63;
    const int d = R.d;
    const int nt = FIRST ? 1 : R.nt;
    const TX *X = (const TX *)R.X;
    const int64_t lo = (int64_t)b * KPP_BLK;
    const int64_t hi = lo + KPP_BLK < R.n ? lo + KPP_BLK : R.n;
    const int mdslot = S.mdslot;
    const double *md = R.pool + (int64_t)mdslot * R.n;
    for (int e = lane; e < NT * D; e += 64) {
        const int t = e / D, j = e % D;
        double v = 0.0;
        if (t < nt && j < d) v = FIRST ? R.cA[j] : (double)X[S.cand[t] * d + j];
        cs[e] = v;
    }
    wave_sync_lds();
    double s = 0.0;
    for (int64_t i0 = lo; i0 < hi; i0 += 64) {
        const int64_t i = i0 + lane;
        const bool ok = i < hi;
        const TX *xr = X + (ok ? i : lo) * d;   // clamped: loads stay in the row
        double acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = 0.0;
#pragma unroll 8
        for (int j = 0; j < D; ++j) {
            const double xv = (double)xr[j < d ? j : 0];
            const double xj = j < d ? xv : 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double tt = xj - cs[t * D + j];
                acc[t] = acc[t] + tt * tt;
            }
        }
        const double m = (!FIRST && ok) ? md[i] : 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t < nt) {
                const double v = FIRST ? acc[t] : (acc[t] < m ? acc[t] : m);
                if (ok) R.pool[(int64_t)(FIRST ? 0 : trial_slot(t, mdslot)) * R.n + i] = v;
                lds[t][lane] = ok ? v : 0.0;
            }
        }
        wave_sync_lds();
        const int cnt = hi - i0 < 64 ? (int)(hi - i0) : 64;
        if (lane < nt) {
            int u = 0;
            for (; u + 8 <= cnt; u += 8) {
                double w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = lds[lane][u + q];
#pragma unroll
                for (int q = 0; q < 8; ++q) s = s + w[q];
            }
            for (; u < cnt; ++u) s = s + lds[lane][u];
        }
        wave_sync_lds();
    }
    if (lane < nt) R.bsum[(int64_t)(FIRST ? 0 : trial_slot(lane, mdslot)) * R.nblk + b] = s;
}

template <typename TX, int NT, bool FIRST>
__device__ __forceinline__ void bk_kpp_pass_d(int db, const BkRun &R, const BkState &S, int b,
                                              double (*lds)[64], double *cs) {
    switch (db) {
        case 0: return bk_kpp_pass<TX, 8, NT, FIRST>(R, S, b, lds, cs);
        case 1: return bk_kpp_pass<TX, 16, NT, FIRST>(R, S, b, lds, cs);
        case 2: return bk_kpp_pass<TX, 32, NT, FIRST>(R, S, b, lds, cs);
        default: return bk_kpp_pass<TX, 64, NT, FIRST>(R, S, b, lds, cs);
    }
}

// One wave per (run, block) pair; four waves per block.
template <bool FIRST>
__global__ __launch_bounds__(256) void bk_kpp_pass_kernel(const BkRun *__restrict__ runs,
                                                          const BkState *__restrict__ st,
                                                          const int2 *__restrict__ pairs,
                                                          int64_t npairs) {
    __shared__ double lds[4][BK_MAXT][64];
    __shared__ double cs[4][BK_MAXT * BK_MAXD];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wv = (int64_t)blockIdx.x * 4 + w;
    if (wv >= npairs) return;
    const int r = __builtin_amdgcn_readfirstlane(pairs[wv].x);
    const int b = __builtin_amdgcn_readfirstlane(pairs[wv].y);
    const BkRun &R = runs[r];
    const BkState &S = st[r];
    if (FIRST ? (!R.kpp || R.k == 1) : S.skip != 0) return;
    const int db = d_bucket(R.d);
    if (FIRST) {
        if (R.f32) return bk_kpp_pass_d<float, 1, true>(db, R, S, b, lds[w], cs[w]);
        return bk_kpp_pass_d<double, 1, true>(db, R, S, b, lds[w], cs[w]);
    }
    if (R.nt <= 4) {
        if (R.f32) return bk_kpp_pass_d<float, 4, false>(db, R, S, b, lds[w], cs[w]);
        return bk_kpp_pass_d<double, 4, false>(db, R, S, b, lds[w], cs[w]);
    }
    if (R.f32) return bk_kpp_pass_d<float, BK_MAXT, false>(db, R, S, b, lds[w], cs[w]);
    return bk_kpp_pass_d<double, BK_MAXT, false>(db, R, S, b, lds[w], cs[w]);
}

// One wave per run: each trial's potential, the first trial with the
// strictly smallest one, and the chosen centre.
__global__ __launch_bounds__(64) void bk_kpp_select_kernel(const BkRun *__restrict__ runs,
                                                           BkState *__restrict__ st,
                                                           const int32_t *__restrict__ rlist,
                                                           int nruns, int c) {
    const int i = blockIdx.x;
    if (i >= nruns) return;
    const int r = rlist[i];
    BkState &S = st[r];
    if (S.skip) return;
    const BkRun R = runs[r];
    const int lane = threadIdx.x;
    int bt = 0;
    double bp = 0.0;
    for (int t = 0; t < R.nt; ++t) {
        const double p =
            wave_seq_sum(R.bsum + (int64_t)trial_slot(t, S.mdslot) * R.nblk, R.nblk, lane);
        if (t == 0 || p < bp) {
            bt = t;
            bp = p;
        }
    }
    double *dst = R.cA + (int64_t)c * R.d;
    const int64_t row = S.cand[bt];
    for (int j = lane; j < R.d; j += 64) dst[j] = ldx(R.X, R.f32, row * R.d + j);
    if (lane == 0) {
        S.mdslot = trial_slot(bt, S.mdslot);
        S.total = bp;
    }
}

// ---------------------------------------------------------- assignment ----
constexpr int BK_CK = 64;   // centres staged in LDS per chunk

// Rows row0 .. row0 + 256 R of one problem against the centres of each of
// its active restarts.  Lane l of wave w holds rows row0 + (w R + s) 64 + l
// (s < R) in registers; the block stages BK_CK centres at a time in LDS
// (rows padded to D features with +0, read by every lane at the same
// address); each lane folds U centres side by side (kmeans.rs:159-163) and
// keeps the first centre with the strictly smallest fold
// (kmeans.rs:295-309).  Features d..D-1 are +0 on both sides, so their
// terms add +0 and leave every partial sum of the fold unchanged.
template <typename TX, int D, int U, int R>
__device__ __forceinline__ void bk_assign_tile(const BkProb &P, const BkRun *__restrict__ runs,
                                               const BkState *__restrict__ st,
                                               const int32_t *__restrict__ done, int64_t row0,
                                               double *__restrict__ cl) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int d = P.d, k = P.k;
    const TX *X = (const TX *)P.X;
    double x[R][D];
    int64_t row[R];
    bool ok[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        row[s] = row0 + (int64_t)(w * R + s) * 64 + lane;
        ok[s] = row[s] < P.n;
        const TX *xr = X + (ok[s] ? row[s] : 0) * d;   // clamped: loads stay in the row
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const double v = (double)xr[j < d ? j : 0];
            x[s][j] = j < d ? v : 0.0;
        }
    }
    for (int q = 0; q < P.nruns; ++q) {
        const int ri = P.run0 + q;
        if (done[ri]) continue;   // the same for the whole block
        const BkRun &Rn = runs[ri];
        const double *C = st[ri].cur_b ? Rn.cB : Rn.cA;
        double best[R];
        int32_t lab[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            best[s] = 1.7976931348623157e308;
            lab[s] = 0;
        }
        for (int c0 = 0; c0 < k; c0 += BK_CK) {
            const int ck = k - c0 < BK_CK ? k - c0 : BK_CK;
            __syncthreads();   // every wave is done with the previous chunk
            for (int e = threadIdx.x; e < ck * D; e += blockDim.x) {
                const int cc = e / D, j = e % D;
                cl[e] = j < d ? C[(int64_t)(c0 + cc) * d + j] : 0.0;
            }
            __syncthreads();
            int c = 0;
            for (; c + U <= ck; c += U) {
                double acc[R][U];
#pragma unroll
                for (int s = 0; s < R; ++s)
#pragma unroll
                    for (int u = 0; u < U; ++u) acc[s][u] = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const double cj = cl[(c + u) * D + j];
#pragma unroll
                        for (int s = 0; s < R; ++s) {
                            const double t = x[s][j] - cj;
                            acc[s][u] = acc[s][u] + t * t;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int s = 0; s < R; ++s)
                        if (acc[s][u] < best[s]) {
                            best[s] = acc[s][u];
                            lab[s] = c0 + c + u;
                        }
            }
            for (; c < ck; ++c) {
                double acc[R];
#pragma unroll
                for (int s = 0; s < R; ++s) acc[s] = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    const double cj = cl[c * D + j];
#pragma unroll
                    for (int s = 0; s < R; ++s) {
                        const double t = x[s][j] - cj;
                        acc[s] = acc[s] + t * t;
                    }
                }
#pragma unroll
                for (int s = 0; s < R; ++s)
                    if (acc[s] < best[s]) {
                        best[s] = acc[s];
                        lab[s] = c0 + c;
                    }
            }
        }
#pragma unroll
        for (int s = 0; s < R; ++s)
            if (ok[s]) {
                Rn.labels[row[s]] = lab[s];
                Rn.best[row[s]] = best[s];
            }
    }
}

template <typename TX>
__device__ __forceinline__ void bk_assign_dispatch(int v, const BkProb &P,
                                                   const BkRun *__restrict__ runs,
                                                   const BkState *__restrict__ st,
                                                   const int32_t *__restrict__ done, int64_t row0,
                                                   double *cl) {
    switch (v) {
        case 0: return bk_assign_tile<TX, 8, 4, 1>(P, runs, st, done, row0, cl);
        case 1: return bk_assign_tile<TX, 8, 8, 1>(P, runs, st, done, row0, cl);
        case 2: return bk_assign_tile<TX, 8, 4, 2>(P, runs, st, done, row0, cl);
        case 3: return bk_assign_tile<TX, 8, 8, 2>(P, runs, st, done, row0, cl);
        case 4: return bk_assign_tile<TX, 16, 4, 1>(P, runs, st, done, row0, cl);
        case 5: return bk_assign_tile<TX, 16, 8, 1>(P, runs, st, done, row0, cl);
        case 6: return bk_assign_tile<TX, 16, 4, 2>(P, runs, st, done, row0, cl);
        case 7: return bk_assign_tile<TX, 16, 8, 2>(P, runs, st, done, row0, cl);
        case 8: return bk_assign_tile<TX, 32, 4, 1>(P, runs, st, done, row0, cl);
        case 9: return bk_assign_tile<TX, 32, 8, 1>(P, runs, st, done, row0, cl);
        case 10: return bk_assign_tile<TX, 32, 4, 2>(P, runs, st, done, row0, cl);
        case 11: return bk_assign_tile<TX, 64, 4, 1>(P, runs, st, done, row0, cl);
        case 12: return bk_assign_tile<TX, 64, 8, 1>(P, runs, st, done, row0, cl);
        case 13: return bk_assign_tile<TX, 64, 2, 1>(P, runs, st, done, row0, cl);
    }
}

// Persistent: every block takes (problem, row tile) items from the queue
// until it is empty.  Items are ordered by decreasing cost.
__global__ __launch_bounds__(256) void bk_assign_kernel(const BkProb *__restrict__ probs,
                                                        const BkItem *__restrict__ items,
                                                        int64_t nitems,
                                                        const BkRun *__restrict__ runs,
                                                        const BkState *__restrict__ st,
                                                        const int32_t *__restrict__ done,
                                                        unsigned long long *__restrict__ queue) {
    __shared__ long long sh_item;
    __shared__ double cl[BK_CK * BK_MAXD];
    for (;;) {
        if (threadIdx.x == 0) sh_item = (long long)atomicAdd(queue, 1ull);
        __syncthreads();
        // items and queue values stay below 2^31 (nitems + grid size)
        const int64_t it = __builtin_amdgcn_readfirstlane((int)sh_item);
        __syncthreads();
        if (it >= nitems) break;
        const int p = __builtin_amdgcn_readfirstlane(items[it].prob);
        const int64_t row0 = items[it].row0;
        const BkProb P = probs[p];
        if (P.f32)
            bk_assign_dispatch<float>(P.variant, P, runs, st, done, row0, cl);
        else
            bk_assign_dispatch<double>(P.variant, P, runs, st, done, row0, cl);
    }
}

// Runs on the reference's serial path (n * d below KM_SERIAL_WORK) that did
// at least one iteration: the inertia of the last assignment becomes the
// left fold of its folds in row order (kmeans.rs:310-322), as in the
// single fit.  One block of EF_NT threads per run.
__global__ __launch_bounds__(EF_NT) void bk_serial_inertia_kernel(const BkRun *__restrict__ runs,
                                                                  BkState *__restrict__ st) {
    const BkRun &R = runs[blockIdx.x];
    BkState &S = st[blockIdx.x];
    if (R.n * (int64_t)R.d >= KM_SERIAL_WORK || S.it <= 0) return;   // the whole block
    const double s = exact_fold_block(R.best, R.n, nullptr);
    if (threadIdx.x == 0) S.inertia = s;
}

// -------------------------------------------------------------- update ----
// The update of kmeans.rs:358-401 and the loop control of :468-488 for
// every active run, in three kernels:
//  1. bk_sort_kernel, one block per run: inertia of the assignment in the
//     order of block_sum_kernel + runs_sum_kernel (1024-row blocks reduced by
//     a fixed tree, the block sums by 64 contiguous runs, then the runs in
//     order), and a stable counting sort of the rows by label (each wave
//     counts and then places a contiguous range of rows, in row order);
//  2. bk_sums_kernel, one lane per (run, cluster, feature): the cluster's
//     rows folded in ascending index order from +0 (kmeans.rs:370-386),
//     divided by the count, or the old centre for an empty cluster;
//  3. bk_control_kernel, one wave per run: max_shift = sqrt(max_c fold(new_c,
//     old_c)), n_iter = iter + 1, stop when shift < tol or after max_iter
//     iterations; block 0 also resets the assignment queue.
__global__ __launch_bounds__(BK_UPD_WAVES * 64) void bk_sort_kernel(
    const BkRun *__restrict__ runs, BkState *__restrict__ st, const int32_t *__restrict__ done) {
    __shared__ int32_t hist[BK_UPD_WAVES * BK_MAXK];
    __shared__ double red[2][256];
    __shared__ int32_t cnt_sh[BK_MAXK];
    __shared__ int64_t off_sh[BK_MAXK];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r = blockIdx.x;
    if (done[r]) return;
    const BkRun R = runs[r];
    BkState &S = st[r];
    const int64_t n = R.n;
    const int k = R.k;

    // inertia
    {
        const int64_t m = (n + SUM_BLK - 1) / SUM_BLK;
        const int g = tid >> 8, t = tid & 255;
        for (int64_t b0 = 0; b0 < m; b0 += 2) {
            const int64_t b = b0 + g;
            double acc = 0.0;
            if (b < m)
                for (int rr = 0; rr < SUM_BLK / 256; ++rr) {
                    const int64_t i = b * SUM_BLK + rr * 256 + t;
                    if (i < n) acc = acc + R.best[i];
                }
            red[g][t] = acc;
            __syncthreads();
            for (int h = 128; h > 0; h >>= 1) {
                if (t < h) red[g][t] = red[g][t] + red[g][t + h];
                __syncthreads();
            }
            if (t == 0 && b < m) R.part[b] = red[g][0];
            __syncthreads();
        }
        if (w == 0) {
            const int64_t per = (m + 63) / 64;
            const int64_t lo = lane * per, hi = lo + per < m ? lo + per : m;
            double acc = 0.0;
            for (int64_t i = lo; i < hi; ++i) acc = acc + R.part[i];
            double tot = 0.0;
            for (int l = 0; l < 64; ++l) tot = tot + wave_bcast(acc, l);
            if (lane == 0) S.inertia = tot;
        }
    }

    // stable counting sort by label
    for (int i = tid; i < BK_UPD_WAVES * k; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t nch = (n + 63) / 64;
    const int64_t cpw = (nch + BK_UPD_WAVES - 1) / BK_UPD_WAVES;
    const int64_t ch0 = w * cpw, ch1 = ch0 + cpw < nch ? ch0 + cpw : nch;
    for (int64_t ch = ch0; ch < ch1; ++ch) {
        const int64_t i = ch * 64 + lane;
        if (i < n) atomicAdd(&hist[w * k + R.labels[i]], 1);
    }
    __syncthreads();
    for (int l = tid; l < k; l += blockDim.x) {
        int32_t c = 0;
        for (int ww = 0; ww < BK_UPD_WAVES; ++ww) c += hist[ww * k + l];
        cnt_sh[l] = c;
        R.cnt[l] = c;
    }
    __syncthreads();
    if (w == 0) {
        // exclusive scan of the counts: lane l takes a contiguous run of clusters
        const int per = (k + 63) / 64;
        const int lo = lane * per, hi = lo + per < k ? lo + per : k;
        int64_t sum = 0;
        for (int l = lo; l < hi; ++l) sum += cnt_sh[l];
        int64_t incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        int64_t run = incl - sum;
        for (int l = lo; l < hi; ++l) {
            off_sh[l] = run;
            R.off[l] = run;
            run += cnt_sh[l];
        }
    }
    __syncthreads();
    for (int l = tid; l < k; l += blockDim.x) {
        int64_t run = off_sh[l];
        for (int ww = 0; ww < BK_UPD_WAVES; ++ww) {
            const int32_t c = hist[ww * k + l];
            hist[ww * k + l] = (int32_t)run;
            run += c;
        }
    }
    __syncthreads();
    int bits = 0;
    while ((1 << bits) < k) ++bits;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int64_t ch = ch0; ch < ch1; ++ch) {
        const int64_t i = ch * 64 + lane;
        const bool valid = i < n;
        const int32_t lab = valid ? R.labels[i] : 0;
        uint64_t msk = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const bool on = (lab >> bit) & 1;
            const uint64_t bb = __ballot(valid && on);
            msk &= on ? bb : ~bb;
        }
        const int rank = __popcll(msk & below);
        const int leader = __ffsll((unsigned long long)msk) - 1;
        int32_t base = 0;
        if (valid && leader == lane) base = atomicAdd(&hist[w * k + lab], __popcll(msk));
        const int32_t gb = __shfl(base, valid ? leader : lane);
        if (valid) R.order[gb + rank] = (int32_t)i;
    }
}

// One thread per (cluster, feature) pair of item (run, first pair).
__global__ __launch_bounds__(256) void bk_sums_kernel(const BkRun *__restrict__ runs,
                                                      const BkState *__restrict__ st,
                                                      const int32_t *__restrict__ done,
                                                      const int2 *__restrict__ items,
                                                      int64_t nitems) {
    const int64_t it = blockIdx.x;
    if (it >= nitems) return;
    const int r = items[it].x;
    if (done[r]) return;
    const BkRun &R = runs[r];
    const int d = R.d;
    const int64_t p = (int64_t)items[it].y + threadIdx.x;
    if (p >= (int64_t)R.k * d) return;
    const int l = (int)(p / d), j = (int)(p % d);
    const int32_t cnt = R.cnt[l];
    const int32_t *ord = R.order + R.off[l];
    const double *oldc = st[r].cur_b ? R.cB : R.cA;
    double *newc = st[r].cur_b ? R.cA : R.cB;
    double acc = 0.0;
    int32_t m = 0;
    constexpr int B = 16;
    if (R.f32) {
        const float *X = (const float *)R.X;
        for (; m + B <= cnt; m += B) {
            double v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) v[u] = (double)X[(int64_t)ord[m + u] * d + j];
#pragma unroll
            for (int u = 0; u < B; ++u) acc = acc + v[u];
        }
        for (; m < cnt; ++m) acc = acc + (double)X[(int64_t)ord[m] * d + j];
    } else {
        const double *X = (const double *)R.X;
        for (; m + B <= cnt; m += B) {
            double v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) v[u] = X[(int64_t)ord[m + u] * d + j];
#pragma unroll
            for (int u = 0; u < B; ++u) acc = acc + v[u];
        }
        for (; m < cnt; ++m) acc = acc + X[(int64_t)ord[m] * d + j];
    }
    newc[p] = cnt > 0 ? acc / (double)cnt : oldc[p];
}

__global__ __launch_bounds__(64) void bk_control_kernel(const BkRun *__restrict__ runs,
                                                        BkState *__restrict__ st,
                                                        int32_t *__restrict__ done,
                                                        unsigned long long *__restrict__ queue) {
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) *queue = 0ull;
    const int r = blockIdx.x;
    if (done[r]) return;
    const BkRun &R = runs[r];
    BkState &S = st[r];
    const double *oldc = S.cur_b ? R.cB : R.cA;
    const double *newc = S.cur_b ? R.cA : R.cB;
    // the largest positive fold (a NaN fold is skipped, like `if shift >
    // max_shift`); non-negative doubles order like their bit patterns
    unsigned long long mx = 0ull;
    for (int l = lane; l < R.k; l += 64) {
        const double f = fold_exact(newc + (int64_t)l * R.d, oldc + (int64_t)l * R.d, R.d);
        const unsigned long long b = f > 0.0 ? (unsigned long long)__double_as_longlong(f) : 0ull;
        mx = b > mx ? b : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(mx, o);
        mx = v > mx ? v : mx;
    }
    if (lane == 0) {
        const double shift = sqrt(__longlong_as_double((long long)mx));
        const int it = S.it + 1;
        S.shift = shift;
        S.it = it;
        S.cur_b ^= 1;
        if (shift < R.tol || it >= R.max_iter) done[r] = 1;
    }
}

__global__ void bk_labels_out_kernel(const int32_t *__restrict__ a, int64_t n,
                                     int64_t *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ------------------------------------------------------------- driver ----
struct StreamOwner {
    hipStream_t s = nullptr;
    StreamOwner() { FL_HIP(hipStreamCreate(&s)); }
    ~StreamOwner() {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct GraphOwner {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~GraphOwner() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};

// Problems the batched kernels take: validated, d <= 64, k <= 1024, the
// library's generator (the reference's stream runs on the per-problem path).
bool batchable(const fl_kmeans_problem &p) {
    return p.params.rng == FL_RNG_LIBRARY && p.d >= 1 && p.d <= BK_MAXD && p.n <= 0x7fffffffLL && p.params.n_clusters >= 1 &&
           p.params.n_clusters <= BK_MAXK && p.n >= p.params.n_clusters && p.params.n_init >= 1 &&
           p.params.n_init <= 4096 && p.X && p.centers_out && p.labels_out;
}

struct BatchTimes {
    double ms_kpp = 0, ms_lloyd = 0;
    int64_t rows = 0, replays = 0;
};

// Device bytes a problem needs in the batched path.
size_t problem_bytes(const fl_kmeans_problem &p) {
    const int64_t n = p.n, d = p.d, k = p.params.n_clusters, ni = p.params.n_init;
    const int64_t nblk = ceil_div(n, KPP_BLK);
    const int64_t nt = 2 + (int64_t)std::floor(std::log((double)k));
    const bool kpp = p.init_centers == nullptr;
    size_t b = (size_t)ni * ((size_t)(2 * k * d) * 8 + (size_t)n * (4 + 8 + 4) +
                             (size_t)ceil_div(n, SUM_BLK) * 8 + (size_t)k * 12);
    if (kpp) b += (size_t)ni * ((size_t)(nt + 1) * (size_t)(n + nblk) * 8);
    return b + 1024 * 8;
}

// Runs the problems listed in idx (all batchable, validated) to completion.
void run_batch(fl_kmeans_problem *problems, const std::vector<int64_t> &idx, hipStream_t st,
               BatchTimes &tm) {
    const int np = (int)idx.size();
    // ---- tables and buffers
    std::vector<BkRun> runs;
    std::vector<BkProb> probs((size_t)np);
    std::vector<int64_t> run_prob;
    // One allocation for every run's buffers: the sizes of problem_bytes,
    // plus 256-byte alignment of each of a run's eight slices.
    size_t bytes = 0;
    for (int q = 0; q < np; ++q) {
        const fl_kmeans_problem &p = problems[idx[(size_t)q]];
        bytes += problem_bytes(p) + (size_t)p.params.n_init * 10 * 256;
    }
    DevBuf arena;
    unsigned char *base = (unsigned char *)arena.get(bytes);
    size_t used = 0;
    auto take = [&](size_t b) {
        unsigned char *p = base + used;
        used += (b + 255) & ~(size_t)255;
        if (used > bytes) throw DeviceError("internal: batched KMeans arena overflow");
        return p;
    };
    int kmax = 1, max_iter = 0;
    for (int q = 0; q < np; ++q) {
        fl_kmeans_problem &p = problems[idx[(size_t)q]];
        const int64_t n = p.n, k = p.params.n_clusters;
        const int d = (int)p.d;
        const int ni = (int)p.params.n_init;
        const bool kpp = p.init_centers == nullptr;
        const int nt = (int)(2 + (int64_t)std::floor(std::log((double)k)));
        BkProb &P = probs[(size_t)q];
        P.X = p.X;
        P.n = n;
        P.d = d;
        P.k = (int)k;
        P.f32 = p.dtype == FL_F32;
        P.nruns = ni;
        P.run0 = (int)runs.size();
        P.variant = variant_of(d, (int)k);
        kmax = std::max<int>(kmax, (int)k);
        max_iter = std::max<int>(max_iter, (int)std::min<int64_t>(p.params.max_iter, 1 << 30));
        const uint64_t seed0 = p.params.has_random_state ? p.params.random_state : 0;
        for (int r = 0; r < ni; ++r) {
            BkRun R;
            std::memset(&R, 0, sizeof(R));
            R.X = p.X;
            R.n = n;
            R.nblk = ceil_div(n, KPP_BLK);
            R.d = d;
            R.k = (int)k;
            R.f32 = P.f32;
            R.nt = nt;
            R.max_iter = (int)std::min<int64_t>(p.params.max_iter, 1 << 30);
            R.tol = p.params.tol;
            R.kpp = kpp ? 1 : 0;
            R.seed = seed0 + (uint64_t)r;
            R.cA = (double *)take((size_t)(k * d) * 8);
            R.cB = (double *)take((size_t)(k * d) * 8);
            R.labels = (int32_t *)take((size_t)n * 4);
            R.best = (double *)take((size_t)n * 8);
            R.order = (int32_t *)take((size_t)n * 4);
            R.part = (double *)take((size_t)ceil_div(n, SUM_BLK) * 8);
            R.cnt = (int32_t *)take((size_t)k * 4);
            R.off = (int64_t *)take((size_t)k * 8);
            if (kpp) {
                R.pool = (double *)take((size_t)(nt + 1) * (size_t)n * 8);
                R.bsum = (double *)take((size_t)(nt + 1) * (size_t)R.nblk * 8);
            }
            if (!kpp)
                FL_HIP(hipMemcpyAsync(R.cA, p.init_centers + (int64_t)r * k * d,
                                      (size_t)(k * d) * 8, hipMemcpyDeviceToDevice, st));
            FL_HIP(hipMemsetAsync(R.labels, 0, (size_t)n * 4, st));
            runs.push_back(R);
            run_prob.push_back(q);
        }
    }
    const int nruns = (int)runs.size();
    // k-means++ (run, block) pairs; assignment items, most costly first
    // k-means++ runs by decreasing k: at centre step c the runs still
    // drawing (k > c) are a prefix of klist, and their (run, block) pairs a
    // prefix of pairs.
    std::vector<int32_t> klist;
    for (int r = 0; r < nruns; ++r)
        if (runs[(size_t)r].kpp && runs[(size_t)r].k > 1) klist.push_back(r);
    std::stable_sort(klist.begin(), klist.end(), [&](int32_t a, int32_t b) {
        return runs[(size_t)a].k > runs[(size_t)b].k;
    });
    std::vector<int2> pairs;
    std::vector<int64_t> pairs_upto(klist.size() + 1, 0);   // pairs of the first i runs
    for (size_t q = 0; q < klist.size(); ++q) {
        const int r = klist[q];
        for (int64_t b = 0; b < runs[(size_t)r].nblk; ++b) pairs.push_back(make_int2(r, (int)b));
        pairs_upto[q + 1] = (int64_t)pairs.size();
    }
    std::vector<BkItem> items;
    std::vector<double> cost;
    for (int q = 0; q < np; ++q) {
        const BkProb &P = probs[(size_t)q];
        const int64_t tile = 256 * (int64_t)kVariants[P.variant].R;
        for (int64_t r0 = 0; r0 < P.n; r0 += tile) {
            BkItem it;
            it.prob = q;
            it.pad = 0;
            it.row0 = r0;
            items.push_back(it);
            cost.push_back((double)std::min<int64_t>(tile, P.n - r0) * P.k * P.d * P.nruns);
        }
    }
    std::vector<size_t> perm(items.size());
    for (size_t i = 0; i < perm.size(); ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<BkItem> sorted(items.size());
    for (size_t i = 0; i < perm.size(); ++i) sorted[i] = items[perm[i]];

    Workspace ws;
    BkRun *d_runs = ws.get<BkRun>("runs", nruns);
    BkState *d_st = ws.get<BkState>("st", nruns);
    int32_t *d_done = ws.get<int32_t>("done", nruns);
    BkProb *d_probs = ws.get<BkProb>("probs", np);
    BkItem *d_items = ws.get<BkItem>("items", (int64_t)sorted.size());
    int2 *d_pairs = ws.get<int2>("pairs", (int64_t)pairs.size());
    int32_t *d_klist = ws.get<int32_t>("klist", (int64_t)klist.size());
    unsigned long long *d_queue = ws.get<unsigned long long>("queue", 1);
    FL_HIP(hipMemcpyAsync(d_runs, runs.data(), sizeof(BkRun) * runs.size(), hipMemcpyHostToDevice, st));
    FL_HIP(hipMemcpyAsync(d_probs, probs.data(), sizeof(BkProb) * probs.size(), hipMemcpyHostToDevice, st));
    if (!sorted.empty())
        FL_HIP(hipMemcpyAsync(d_items, sorted.data(), sizeof(BkItem) * sorted.size(),
                              hipMemcpyHostToDevice, st));
    if (!pairs.empty()) {
        FL_HIP(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(int2) * pairs.size(),
                              hipMemcpyHostToDevice, st));
        FL_HIP(hipMemcpyAsync(d_klist, klist.data(), sizeof(int32_t) * klist.size(),
                              hipMemcpyHostToDevice, st));
    }
    FL_HIP(hipMemsetAsync(d_queue, 0, 8, st));

    // ---- k-means++ for every run, centre by centre
    PhaseTimer ptm(st);
    const int t_kpp = ptm.begin();
    hipLaunchKernelGGL(bk_init_kernel, dim3(grid1(nruns, 64)), dim3(64), 0, st, d_runs, d_st,
                       d_done, nruns);
    check_launch();
    const int64_t npairs = (int64_t)pairs.size();
    if (npairs > 0) {
        hipLaunchKernelGGL(bk_kpp_pass_kernel<true>, dim3((unsigned)ceil_div(npairs, 4)),
                           dim3(256), 0, st, d_runs, d_st, d_pairs, npairs);
        hipLaunchKernelGGL(bk_kpp_total_kernel, dim3((unsigned)nruns), dim3(64), 0, st, d_runs,
                           d_st, nruns);
        check_launch();
        size_t m = klist.size();   // runs with k > c
        for (int c = 1; c < kmax; ++c) {
            while (m > 0 && runs[(size_t)klist[m - 1]].k <= c) --m;
            if (m == 0) break;
            const int mi = (int)m;
            const int64_t np_c = pairs_upto[m];
            hipLaunchKernelGGL(bk_kpp_draw_kernel, dim3(grid1(mi, 64)), dim3(64), 0, st, d_runs,
                               d_st, d_klist, mi, c);
            hipLaunchKernelGGL(bk_kpp_search_kernel, dim3((unsigned)(mi * BK_MAXT)), dim3(64), 0,
                               st, d_runs, d_st, d_klist, mi);
            hipLaunchKernelGGL(bk_kpp_pass_kernel<false>, dim3((unsigned)ceil_div(np_c, 4)),
                               dim3(256), 0, st, d_runs, d_st, d_pairs, np_c);
            hipLaunchKernelGGL(bk_kpp_select_kernel, dim3((unsigned)mi), dim3(64), 0, st, d_runs,
                               d_st, d_klist, mi, c);
            check_launch();
        }
    }
    ptm.end(t_kpp);

    // ---- Lloyd iterations: one captured graph of [assign, sort, sums, control]
    // centroid-sum items: 256 (cluster, feature) pairs of one run, the runs
    // with the most rows per cluster first
    std::vector<int2> sitems;
    for (int r = 0; r < nruns; ++r)
        for (int64_t p0 = 0; p0 < (int64_t)runs[(size_t)r].k * runs[(size_t)r].d; p0 += 256)
            sitems.push_back(make_int2(r, (int)p0));
    std::stable_sort(sitems.begin(), sitems.end(), [&](const int2 &a, const int2 &b) {
        return runs[(size_t)a.x].n / runs[(size_t)a.x].k > runs[(size_t)b.x].n / runs[(size_t)b.x].k;
    });
    int2 *d_sitems = ws.get<int2>("sitems", (int64_t)sitems.size());
    FL_HIP(hipMemcpyAsync(d_sitems, sitems.data(), sizeof(int2) * sitems.size(),
                          hipMemcpyHostToDevice, st));
    const int t_lloyd = ptm.begin();
    if (max_iter > 0 && !sorted.empty()) {
        const unsigned agrid = (unsigned)std::min<int64_t>((int64_t)sorted.size(),
                                                           (int64_t)device_cu_count() * 3);
        GraphOwner go;
        FL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(bk_assign_kernel, dim3(agrid), dim3(256), 0, st, d_probs, d_items,
                           (int64_t)sorted.size(), d_runs, d_st, d_done, d_queue);
        hipLaunchKernelGGL(bk_sort_kernel, dim3((unsigned)nruns), dim3(BK_UPD_WAVES * 64), 0, st,
                           d_runs, d_st, d_done);
        hipLaunchKernelGGL(bk_sums_kernel, dim3((unsigned)sitems.size()), dim3(256), 0, st,
                           d_runs, d_st, d_done, d_sitems, (int64_t)sitems.size());
        hipLaunchKernelGGL(bk_control_kernel, dim3((unsigned)nruns), dim3(64), 0, st, d_runs,
                           d_st, d_done, d_queue);
        FL_HIP(hipStreamEndCapture(st, &go.g));
        FL_HIP(hipGraphInstantiate(&go.e, go.g, nullptr, nullptr, 0));
        std::vector<int32_t> hd((size_t)nruns);
        const int check_every = 4;
        for (int it = 0; it < max_iter; ++it) {
            FL_HIP(hipGraphLaunch(go.e, st));
            ++tm.replays;
            if ((it + 1) % check_every == 0 && it + 1 < max_iter) {
                FL_HIP(hipMemcpyAsync(hd.data(), d_done, (size_t)nruns * 4, hipMemcpyDeviceToHost,
                                      st));
                FL_HIP(hipStreamSynchronize(st));
                bool all = true;
                for (int r = 0; r < nruns; ++r) all = all && hd[(size_t)r] != 0;
                if (all) break;
            }
        }
    }
    if (max_iter > 0 && nruns > 0) {
        hipLaunchKernelGGL(bk_serial_inertia_kernel, dim3((unsigned)nruns), dim3(EF_NT), 0, st,
                           d_runs, d_st);
        check_launch();
    }
    ptm.end(t_lloyd);

    // ---- per problem: the run with the strictly smallest inertia
    std::vector<BkState> hs((size_t)nruns);
    FL_HIP(hipMemcpyAsync(hs.data(), d_st, sizeof(BkState) * (size_t)nruns, hipMemcpyDeviceToHost,
                          st));
    FL_HIP(hipStreamSynchronize(st));
    tm.ms_kpp += ptm.ms(t_kpp);
    tm.ms_lloyd += ptm.ms(t_lloyd);
    for (int q = 0; q < np; ++q) {
        fl_kmeans_problem &p = problems[idx[(size_t)q]];
        const BkProb &P = probs[(size_t)q];
        int bestr = P.run0;
        for (int r = 0; r < P.nruns; ++r) {
            const BkState &S = hs[(size_t)(P.run0 + r)];
            tm.rows += P.n * (int64_t)S.it;
            if (r == 0 || S.inertia < hs[(size_t)bestr].inertia) bestr = P.run0 + r;
        }
        const BkState &S = hs[(size_t)bestr];
        const BkRun &R = runs[(size_t)bestr];
        FL_HIP(hipMemcpyAsync(p.centers_out, S.cur_b ? R.cB : R.cA, (size_t)(P.k * P.d) * 8,
                              hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(bk_labels_out_kernel, dim3(grid1(P.n, 256)), dim3(256), 0, st, R.labels,
                           P.n, p.labels_out);
        check_launch();
        p.inertia = S.inertia;
        p.n_iter = S.it;
        p.status = FL_OK;
    }
    FL_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

fl_status fl_kmeans_fit_batched(fl_kmeans_problem *problems, int64_t n_problems,
                                int32_t n_workers, fl_report *report) {
    if (!problems && n_problems > 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: problems: must not be NULL");
    if (n_problems < 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_problems: negative");
    WallClock wc;
    std::vector<std::string> errs((size_t)n_problems);
    std::vector<int64_t> batch, single;
    for (int64_t i = 0; i < n_problems; ++i) {
        fl_kmeans_problem &p = problems[i];
        p.status = (int32_t)fl::validate_kmeans_fit(p.n, p.d, &p.params, p.dtype);
        if (p.status != FL_OK) {
            errs[(size_t)i] = fl_last_error();
            continue;
        }
        if (batchable(p))
            batch.push_back(i);
        else
            single.push_back(i);
    }
    BatchTimes tm;
    if (!batch.empty()) {
        fl_status s = guarded([&]() -> fl_status {
            StreamOwner so;
            // groups of problems within a device-memory budget
            const size_t budget = (size_t)24 << 30;
            std::vector<int64_t> group;
            size_t gb = 0;
            for (size_t q = 0; q < batch.size(); ++q) {
                const size_t b = problem_bytes(problems[batch[q]]);
                if (!group.empty() && gb + b > budget) {
                    run_batch(problems, group, so.s, tm);
                    group.clear();
                    gb = 0;
                }
                group.push_back(batch[q]);
                gb += b;
            }
            if (!group.empty()) run_batch(problems, group, so.s, tm);
            return FL_OK;
        });
        if (s != FL_OK) {
            const std::string e = fl_last_error();
            for (int64_t i : batch) {
                problems[i].status = (int32_t)s;
                errs[(size_t)i] = e;
            }
        }
    }
    // Problems outside the batched kernels' limits: whole fits on worker
    // threads, one stream each.
    std::vector<fl_report> reps((size_t)n_problems);
    if (!single.empty()) {
        const int workers = (int)std::max<int64_t>(
            1, std::min<int64_t>(n_workers > 0 ? n_workers : 8, (int64_t)single.size()));
        std::atomic<size_t> next(0);
        auto work = [&]() {
            for (;;) {
                const size_t q = next.fetch_add(1);
                if (q >= single.size()) break;
                const int64_t i = single[q];
                fl_kmeans_problem &p = problems[i];
                p.status = fl_kmeans_fit_device(p.X, p.dtype, p.n, p.d, &p.params, p.init_centers,
                                                p.centers_out, p.labels_out, &p.inertia,
                                                &p.n_iter, &reps[(size_t)i]);
                if (p.status != FL_OK) errs[(size_t)i] = fl_last_error();
            }
        };
        std::vector<std::thread> pool;
        try {
            for (int t = 1; t < workers; ++t) pool.emplace_back(work);
        } catch (...) {
        }
        work();
        for (auto &th : pool) th.join();
    }
    if (report) {
        std::memset(report, 0, sizeof(*report));
        report->ms_total = wc.ms();
        report->n_rows = tm.rows;
        report->ms_distance = tm.ms_lloyd;
        report->ms_update = tm.ms_kpp;
        report->n_dist_launches = tm.replays;
        for (int64_t i : single)
            if (problems[i].status == FL_OK) {
                report->n_rows += reps[(size_t)i].n_rows;
                report->n_fallback += reps[(size_t)i].n_fallback;
                report->ms_distance += reps[(size_t)i].ms_distance;
            }
        for (int64_t i = 0; i < n_problems; ++i)
            if (problems[i].status == FL_OK) report->n_iter += problems[i].n_iter;
    }
    for (int64_t i = 0; i < n_problems; ++i)
        if (problems[i].status != FL_OK)
            return fail((fl_status)problems[i].status,
                        "problem " + std::to_string(i) + ": " + errs[(size_t)i]);
    clear_error();
    return FL_OK;
}

// Tuning table of the batched assignment: table[D bucket][k bucket] =
// variant index (D = 8, 16, 32, 64; k <= 16, 64, 256, 1024).
fl_status fl_kmeans_batched_set_tuning(const int32_t *table) {
    if (!table) return fail(FL_INVALID_PARAMETER, "InvalidParameter: table: must not be NULL");
    for (int i = 0; i < kDBuckets * kKBuckets; ++i) {
        const int v = table[i];
        const int D = 8 << (i / kKBuckets);
        if (v < 0 || v >= kNumVariants || kVariants[v].D != D)
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: table: entry " + std::to_string(i) +
                            " is not a variant for D = " + std::to_string(D));
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (int i = 0; i < kDBuckets * kKBuckets; ++i) g_tuning[i / kKBuckets][i % kKBuckets] = table[i];
    return FL_OK;
}

fl_status fl_kmeans_batched_get_tuning(int32_t *table) {
    if (!table) return fail(FL_INVALID_PARAMETER, "InvalidParameter: table: must not be NULL");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (int i = 0; i < kDBuckets * kKBuckets; ++i) table[i] = g_tuning[i / kKBuckets][i % kKBuckets];
    return FL_OK;
}

// Variant descriptions: (D, U, R) of each index.
int32_t fl_kmeans_batched_n_variants(void) { return kNumVariants; }
fl_status fl_kmeans_batched_variant(int32_t v, int32_t *D, int32_t *U, int32_t *R) {
    if (v < 0 || v >= kNumVariants || !D || !U || !R)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: variant: out of range");
    *D = kVariants[v].D;
    *U = kVariants[v].U;
    *R = kVariants[v].R;
    return FL_OK;
}

}  // extern "C"
