// knn.hip -- brute-force KNeighborsClassifier / KNeighborsRegressor on gfx950.
//
// Reference: ferrolearn-neighbors/src/knn.rs (fit :256-320 / :660-718,
// predict :322-383 / :720-781, predict_proba :452-486, kneighbors
// :823-834 / :875-917) with Algorithm::BruteForce (knn.rs:148) whose search
// is kdtree.rs:364-391.  The fitted handle owns device copies of X and y, as
// the fitted Rust model owns clones of them (knn.rs:311-318).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "host.hpp"

using namespace fl;

struct fl_knn {
    int is_classifier = 0;
    int unsupervised = 0;          // NearestNeighbors (no targets)
    int32_t dtype = FL_F64;
    int64_t n = 0, d = 0, n_neighbors = 5;
    int32_t weights = FL_UNIFORM;
    int32_t precision = FL_PREC_FP64;
    DevBuf X;                      // n*d in dtype
    DevBuf xn, xs;                 // row squared norms and norms (f64)
    std::vector<int64_t> classes;  // sorted unique labels (classifier)
    DevBuf ypos, classes_dev;      // label position per row, classes
    DevBuf yreg;                   // regression targets
    DevBuf bh, bl;                 // BF16 hi / lo split of X, padded rows
    DevBuf bxnp;                   // BF16 per-row key term (1-b)|x|^2
    DevBuf bsmax;                  // largest row norm (f64 bits)
    int bf_split_ready = 0;        // hi/lo arrays built
    int db_range = -1;             // range_flags_kernel bits of X (-1: unknown)
    int bf_terms_passes = 0;       // passes the key terms were built for
    DevBuf xpad;                   // f64 filter: X with rows padded to 4*D4 (if d differs)
    DevBuf fxnp;                   // f64 filter: per-row key term (1-b)|x|^2
    int f64_ready = 0;             // xpad / fxnp / bsmax built
    int smax_ready = 0;            // bsmax built
    int f64_ok = 0;                // this call's data fits the f64 filter's range
    hipStream_t stream = nullptr;
    Workspace ws;
    ~fl_knn() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int64_t kMaxFilterK = 64;   // largest k of the filter kernels' lists

std::string shape_msg(int64_t expected, int64_t actual, const char *ctx) {
    return "ShapeMismatch: expected [" + std::to_string(expected) + "], actual [" +
           std::to_string(actual) + "]: " + ctx;
}

int pick_m_filter(int64_t k) {
    for (int m : {16, 32, 64})
        if (k + 6 <= m) return m;
    return 64;
}

int pick_m_exact(int64_t k) {
    for (int m : {16, 32, 64})
        if (k <= m) return m;
    return 64;
}

struct Acc {
    double ms_dist = 0, ms_fin = 0, ms_fb = 0, ms_upd = 0;
    int64_t n_cand = 0, n_fb = 0, n_launch = 0;
};

__global__ void gather_rows_kernel(const double *__restrict__ Q, int d,
                                   const int32_t *__restrict__ list, int64_t cnt,
                                   double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * d) return;
    const int64_t r = t / d, j = t % d;
    out[t] = Q[(int64_t)list[r] * d + j];
}

__global__ void scatter_results_kernel(const double *__restrict__ dsrc,
                                       const int64_t *__restrict__ isrc,
                                       const int32_t *__restrict__ list, int64_t cnt, int k,
                                       double *__restrict__ dist, int64_t *__restrict__ idx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * k) return;
    const int64_t r = t / k, m = t % k;
    dist[(int64_t)list[r] * k + m] = dsrc[t];
    idx[(int64_t)list[r] * k + m] = isrc[t];
}

// ---------------------------------------------------------------------------
// Exact top-k by sorting every distance (any k <= n; also the path for the
// few queries a filter leaves, which it spreads over the whole device).
//
// Keys are the bits of sqrt(fold) -- non-negative doubles order as their bit
// patterns, NaN after +inf -- and values the row positions; a stable
// segmented radix sort per query leaves equal distances in index order, so
// the first k of a segment are the reference's k smallest in (distance,
// index) order (kdtree.rs:364-391, ties as the oracle defines them).
// ---------------------------------------------------------------------------

template <typename TB>
__global__ void knn_all_dist_kernel(const double *__restrict__ Q, const TB *__restrict__ X,
                                    int64_t N, int d, uint64_t *__restrict__ keys,
                                    int *__restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t qq = blockIdx.y;
    if (i >= N) return;
    const double f = fold_exact(X + i * d, Q + qq * d, d);
    keys[qq * N + i] = (uint64_t)__double_as_longlong(sqrt(f));
    pos[qq * N + i] = (int)i;
}

__global__ void knn_seg_offsets_kernel(int64_t qc, int64_t N, int *__restrict__ offs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= qc) offs[t] = (int)(t * N);
}

// First k of each sorted segment into rows q0.. of the outputs; a NaN
// distance (sorted last) sets *nan.
__global__ void knn_take_sorted_kernel(const uint64_t *__restrict__ keys,
                                       const int *__restrict__ pos, int64_t qc, int64_t N, int k,
                                       double *__restrict__ dist, int64_t *__restrict__ idx,
                                       int *__restrict__ nan) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= qc * k) return;
    const int64_t qq = t / k, m = t % k;
    const int64_t p = qq * N + m;
    dist[t] = __longlong_as_double((long long)keys[p]);
    idx[t] = pos[p];
    if (m == 0) {
        const double last = __longlong_as_double((long long)keys[qq * N + N - 1]);
        if (last != last) atomicOr(nan, 1);
    }
}

// Queries per sort: about 64M keys (about 40 bytes each with the sort's buffers).
// At most 65535 (the grid's y extent).
inline int64_t sorted_chunk(int64_t N) {
    return std::min<int64_t>(65535,
                             std::max<int64_t>(1, ((int64_t)1 << 26) / std::max<int64_t>(N, 1)));
}

// Queries sorted one at a time by the device-wide radix sort (every CU on
// one query's N keys); more use the segmented sort, one block per query.
constexpr int64_t kSortEachMax = 16;

template <typename TB>
void knn_sorted(fl_knn *h, const double *Q, int64_t nq, int k, double *dist, int64_t *idx,
                Acc &acc, bool fallback) {
    if (nq == 0) return;
    hipStream_t st = h->stream;
    Workspace &ws = h->ws;
    const int64_t N = h->n;
    const int d = (int)h->d;
    if (nq <= kSortEachMax && N < ((int64_t)1 << 31)) {
        uint64_t *kin = ws.get<uint64_t>("srt_kin", nq * N);
        uint64_t *kout = ws.get<uint64_t>("srt_kout", N);
        int *pin = ws.get<int>("srt_pin", nq * N);
        int *pout = ws.get<int>("srt_pout", N);
        int *nan = ws.get<int>("srt_nan", 1);
        size_t tmp_bytes = 0;
        FL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, pin, pout,
                                                  (int)N, 0, 64, st));
        void *tmp = ws.get<unsigned char>("srt_tmp1", (int64_t)tmp_bytes + 256);
        FL_HIP(hipMemsetAsync(nan, 0, 4, st));
        PhaseTimer tm(st);
        const int t0 = tm.begin();
        hipLaunchKernelGGL(knn_all_dist_kernel<TB>, dim3((unsigned)grid1(N, 256), (unsigned)nq),
                           dim3(256), 0, st, Q, (const TB *)h->X.ptr(), N, d, kin, pin);
        check_launch();
        for (int64_t q = 0; q < nq; ++q) {
            size_t tb = tmp_bytes;
            // LSD radix sort: stable, so equal distances stay in row order
            FL_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin + q * N, kout, pin + q * N,
                                                      pout, (int)N, 0, 64, st));
            hipLaunchKernelGGL(knn_take_sorted_kernel, dim3(grid1(k, 256)), dim3(256), 0, st,
                               kout, pout, (int64_t)1, N, k, dist + q * k, idx + q * k, nan);
            check_launch();
        }
        tm.end(t0);
        int hn = 0;
        FL_HIP(hipMemcpyAsync(&hn, nan, 4, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        (fallback ? acc.ms_fb : acc.ms_dist) += tm.ms(t0);
        acc.n_cand += nq * N;
        acc.n_launch += 1;
        if (hn)
            throw StatusError{FL_NUMERICAL_INSTABILITY,
                              "NumericalInstability: a distance is NaN (the reference panics "
                              "here, kdtree.rs:386)"};
        return;
    }
    const int64_t chunk = std::min<int64_t>(nq, sorted_chunk(N));
    const int64_t cap = chunk * N;
    uint64_t *kin = ws.get<uint64_t>("srt_kin", cap);
    uint64_t *kout = ws.get<uint64_t>("srt_kout", cap);
    int *pin = ws.get<int>("srt_pin", cap);
    int *pout = ws.get<int>("srt_pout", cap);
    int *offs = ws.get<int>("srt_offs", chunk + 1);
    int *nan = ws.get<int>("srt_nan", 1);
    size_t tmp_bytes = 0;
    FL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(
        nullptr, tmp_bytes, kin, kout, pin, pout, (int)cap,
        (int)chunk, offs, offs + 1, 0, 64, st));
    void *tmp = ws.get<unsigned char>("srt_tmp", (int64_t)tmp_bytes + 256);
    FL_HIP(hipMemsetAsync(nan, 0, 4, st));
    PhaseTimer tm(st);
    const int t0 = tm.begin();
    for (int64_t q0 = 0; q0 < nq; q0 += chunk) {
        const int64_t qc = std::min<int64_t>(chunk, nq - q0);
        hipLaunchKernelGGL(knn_all_dist_kernel<TB>, dim3((unsigned)grid1(N, 256), (unsigned)qc),
                           dim3(256), 0, st, Q + q0 * d, (const TB *)h->X.ptr(), N, d, kin, pin);
        hipLaunchKernelGGL(knn_seg_offsets_kernel, dim3(grid1(qc + 1, 256)), dim3(256), 0, st, qc,
                           N, offs);
        check_launch();
        size_t tb = tmp_bytes;
        FL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(
            tmp, tb, kin, kout, pin, pout, (int)(qc * N),
            (int)qc, offs, offs + 1, 0, 64, st));
        hipLaunchKernelGGL(knn_take_sorted_kernel, dim3(grid1(qc * k, 256)), dim3(256), 0, st,
                           kout, pout, qc, N, k, dist + q0 * k, idx + q0 * k, nan);
        check_launch();
    }
    tm.end(t0);
    int hn = 0;
    FL_HIP(hipMemcpyAsync(&hn, nan, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    (fallback ? acc.ms_fb : acc.ms_dist) += tm.ms(t0);
    acc.n_cand += nq * N;
    acc.n_launch += 1;
    if (hn)
        throw StatusError{FL_NUMERICAL_INSTABILITY,
                          "NumericalInstability: a distance is NaN (the reference panics "
                          "here, kdtree.rs:386)"};
}

// Exact top-k of nq f64 queries against the fitted rows.  mode selects the
// candidate filter; queries the certificate cannot settle are re-run in the
// exact mode, whose kept lists are exact by construction.
template <typename TB>
void knn_search(fl_knn *h, const double *Q, int64_t nq, int k, int mode, double *dist,
                int64_t *idx, Acc &acc, int depth) {
    if (nq == 0) return;
    hipStream_t st = h->stream;
    Workspace &ws = h->ws;
    const std::string p = depth ? "d" + std::to_string(depth) + "_" : "";
    const TB *B = (const TB *)h->X.ptr();
    const int64_t nB = h->n;
    const int d = (int)h->d;

    double *qn = ws.get<double>(p + "qn", nq);
    double *qs = ws.get<double>(p + "qs", nq);
    hipLaunchKernelGGL(row_norms_kernel<double>, dim3(grid1(nq, 256)), dim3(256), 0, st, Q, nq,
                       d, qn, qs);
    check_launch();

    const bool bf = mode == MODE_BF16 || mode == MODE_BF16X3;
    const bool f64 = mode == MODE_F64;
    // BF16 and f64-filter kernels output MP candidates per split (merged
    // from four register lists per resident row); the other kernels keep one
    // list of M in LDS.
    const int MP = k <= 6 ? 8 : 16;
    const int M = (bf || f64) ? MP : (mode == MODE_EXACT ? pick_m_exact(k) : pick_m_filter(k));
    const int rows_per_block = bf ? BF_Q : (f64 ? F64_Q : TM_ROWS);
    const int64_t nblk = ceil_div(nq, rows_per_block);
    // The exact form needs no certificate, so a few queries may use many
    // splits (parallelism) up to the large finalize capacity.
    const int64_t cap = mode == MODE_EXACT ? FIN_MAXC_BIG : FIN_MAXC;
    const int64_t smax = std::max<int64_t>(
        1, std::min<int64_t>(cap / M, std::max<int64_t>(1, nB / 256)));
    int64_t S;
    if (f64) {
        const int64_t row_bytes = (int64_t)f64_d4(d) * 4 * (int64_t)sizeof(TB);
        const int64_t smin = ceil_div(nB * row_bytes, (int64_t)1 << 31);
        S = std::max<int64_t>(smin, bf16_splits(nblk, nB, smax));
    } else if (bf) {
        // Each split's hi/lo rows must stay within 2^31 bytes (32-bit buffer
        // offsets of the staged copies).
        const int64_t row_bytes = (int64_t)bf16_nkc(d) * 64;
        const int64_t smin = ceil_div(nB * row_bytes, (int64_t)1 << 31);
        S = std::max<int64_t>(smin, bf16_splits(nblk, nB, smax));
    } else {
        S = ceil_div(2048, nblk);
        S = std::min<int64_t>(S, smax);
        S = std::max<int64_t>(S, 1);
    }
    const int64_t bps = ceil_div(nB, S);
    S = ceil_div(nB, bps);

    Cand cand;
    cand.key = ws.get<double>(p + "ckey", S * nq * M);
    cand.idx = ws.get<int32_t>(p + "cidx", S * nq * M);
    cand.cnt = ws.get<int32_t>(p + "ccnt", S * nq);
    cand.theta = ws.get<double>(p + "cth", S * nq);
    cand.theta_idx = ws.get<int32_t>(p + "cthi", S * nq);

    PhaseTimer tm(st);
    int t0;
    if (!h->smax_ready && (bf || f64)) {
        unsigned long long *smaxp = (unsigned long long *)h->bsmax.get(8);
        FL_HIP(hipMemsetAsync(smaxp, 0, 8, st));
        hipLaunchKernelGGL(max_nonneg_kernel, dim3(256), dim3(256), 0, st,
                           (const double *)h->xs.ptr(), nB, smaxp);
        check_launch();
        h->smax_ready = 1;
    }
    if (f64) {
        const int d4 = f64_d4(d);
        const int dp = 4 * d4;
        const Bound bd = bound_fp64(d);
        if (!h->f64_ready) {
            if (dp != d) {
                TB *xp = (TB *)h->xpad.get((size_t)(nB * dp) * sizeof(TB));
                hipLaunchKernelGGL(pad_rows_kernel<TB>, dim3(grid1(nB * dp, 256)), dim3(256), 0,
                                   st, B, nB, d, dp, xp);
            }
            double *xnp = (double *)h->fxnp.get((size_t)nB * 8);
            hipLaunchKernelGGL(f64_stream_terms_kernel, dim3(grid1(nB, 256)), dim3(256), 0, st,
                               (const double *)h->xn.ptr(), nB, bd.beta, xnp);
            check_launch();
            h->f64_ready = 1;
        }
        double *qoff = ws.get<double>(p + "qoff", nq);
        hipLaunchKernelGGL(f64_resident_terms_kernel, dim3(grid1(nq, 256)), dim3(256), 0, st, qn,
                           qs, nq, d, bd.alpha, bd.beta, (const double *)h->bsmax.ptr(), qoff);
        check_launch();
        F64Args a;
        a.nA = nq;
        a.nB = nB;
        a.d = d;
        a.b_per_split = bps;
        a.A = Q;
        a.B = dp != d ? h->xpad.ptr() : (const void *)B;
        a.qoff = qoff;
        a.xnp = (const double *)h->fxnp.ptr();
        a.out = cand;
        t0 = tm.begin();
        launch_f64_knn<TB>(MP, d4, a, (int)S, st);
        tm.end(t0);
    } else if (mode == MODE_BF16 || mode == MODE_BF16X3) {
        const int passes = mode == MODE_BF16X3 ? 3 : 1;
        const int nkc = bf16_nkc(d);
        const int dpad = nkc * 32;
        const Bound bd = bound_bf16(d, passes);
        if (!h->bf_split_ready) {
            unsigned short *bh = (unsigned short *)h->bh.get((size_t)(nB * dpad) * 2);
            unsigned short *bl = (unsigned short *)h->bl.get((size_t)(nB * dpad) * 2);
            hipLaunchKernelGGL(bf16_split_kernel<TB>, dim3(grid1(nB * dpad, 256)), dim3(256), 0,
                               st, B, nB, d, dpad, bh, bl);
            check_launch();
            h->bf_split_ready = 1;
        }
        if (h->bf_terms_passes != passes) {
            float *xnp = (float *)h->bxnp.get((size_t)nB * 4);
            hipLaunchKernelGGL(bf16_stream_terms_kernel, dim3(grid1(nB, 256)), dim3(256), 0, st,
                               (const double *)h->xn.ptr(), nB, bd.beta, xnp);
            check_launch();
            h->bf_terms_passes = passes;
        }
        unsigned short *qh = ws.get<unsigned short>(p + "qh", nq * dpad);
        unsigned short *ql = ws.get<unsigned short>(p + "ql", nq * dpad);
        double *qoff = ws.get<double>(p + "qoff", nq);
        hipLaunchKernelGGL(bf16_split_kernel<double>, dim3(grid1(nq * dpad, 256)), dim3(256), 0,
                           st, Q, nq, d, dpad, qh, ql);
        hipLaunchKernelGGL(bf16_resident_terms_kernel, dim3(grid1(nq, 256)), dim3(256), 0, st, qn,
                           qs, nq, d, bd.alpha, bd.beta, (const double *)h->bsmax.ptr(), qoff);
        check_launch();
        Bf16Args a;
        a.nA = nq;
        a.nB = nB;
        a.b_per_split = bps;
        a.ah = qh;
        a.al = ql;
        a.bh = (const unsigned short *)h->bh.ptr();
        a.bl = (const unsigned short *)h->bl.ptr();
        a.qoff = qoff;
        a.xnp = (const float *)h->bxnp.ptr();
        a.out = cand;
        t0 = tm.begin();
        launch_bf16(passes, MP, nkc, a, (int)S, st);
        tm.end(t0);
    } else {
        TopmArgs a;
        a.nA = nq;
        a.nB = nB;
        a.d = d;
        a.b_per_split = bps;
        a.an = qn;
        a.as = qs;
        a.bn = (const double *)h->xn.ptr();
        a.bs = (const double *)h->xs.ptr();
        const Bound bd = bound_fp64(d);
        a.alpha = bd.alpha;
        a.beta = bd.beta;
        a.sqrt_key = 1;
        a.out = cand;
        t0 = tm.begin();
        launch_topm<double, TB>(M, mode, Q, B, a, (int)S, st);
        tm.end(t0);
    }
    acc.n_launch += 1;

    int32_t *flags = ws.get<int32_t>(p + "flags", nq);
    int32_t *fcnt = ws.get<int32_t>(p + "fcnt", 1);
    unsigned long long *ncand = ws.get<unsigned long long>(p + "ncand", 1);
    FL_HIP(hipMemsetAsync(fcnt, 0, sizeof(int32_t), st));
    FL_HIP(hipMemsetAsync(ncand, 0, sizeof(unsigned long long), st));
    const int t1 = tm.begin();
    if (S * M > FIN_MAXC)
        hipLaunchKernelGGL((knn_finalize_kernel<TB, FIN_MAXC_BIG>), dim3((unsigned)nq), dim3(64), 0,
                           st, Q, B, d, nq, nB, (int)S, M, bps, cand, k,
                           mode == MODE_EXACT ? 0 : 1, dist, idx, flags, fcnt, ncand);
    else
        hipLaunchKernelGGL((knn_finalize_kernel<TB, FIN_MAXC>), dim3((unsigned)nq), dim3(64), 0, st,
                           Q, B, d, nq, nB, (int)S, M, bps, cand, k, mode == MODE_EXACT ? 0 : 1,
                           dist, idx, flags, fcnt, ncand);
    check_launch();
    tm.end(t1);
    int32_t nflag = 0;
    unsigned long long nc = 0;
    FL_HIP(hipMemcpyAsync(&nflag, fcnt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FL_HIP(hipMemcpyAsync(&nc, ncand, sizeof(nc), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (depth == 0) {
        acc.ms_dist += tm.ms(t0);
        acc.ms_fin += tm.ms(t1);
    } else {
        acc.ms_fb += tm.ms(t0) + tm.ms(t1);
    }
    acc.n_cand += (int64_t)nc;
    if (nflag == 0) return;
    if (mode == MODE_EXACT)
        throw StatusError{FL_NUMERICAL_INSTABILITY,
                          "NumericalInstability: a distance is NaN (the reference panics "
                          "here, kdtree.rs:386)"};
    // Uncertified queries go to the next, tighter filter: BF16 -> f64 GEMM
    // form -> exact folds.
    const int next_mode = bf ? (h->f64_ok && k <= 24 && f64_d4(d) > 0 ? MODE_F64 : MODE_MFMA64)
                             : MODE_EXACT;
    if (depth == 0) acc.n_fb += nflag;
    const std::string fp = "fb" + std::to_string(depth) + "_";
    double *Qf = ws.get<double>(fp + "q", (int64_t)nflag * d);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid1((int64_t)nflag * d, 256)), dim3(256), 0, st,
                       Q, d, flags, (int64_t)nflag, Qf);
    check_launch();
    double *df = ws.get<double>(fp + "d", (int64_t)nflag * k);
    int64_t *xf = ws.get<int64_t>(fp + "i", (int64_t)nflag * k);
    // A few queries (up to 2^24 distances) are sorted exactly over the whole
    // device; more go through the next filter, whose grid follows the queries.
    if ((int64_t)nflag * h->n <= ((int64_t)1 << 24))
        knn_sorted<TB>(h, Qf, nflag, k, df, xf, acc, true);
    else
        knn_search<TB>(h, Qf, nflag, k, next_mode, df, xf, acc, depth + 1);
    hipLaunchKernelGGL(scatter_results_kernel, dim3(grid1((int64_t)nflag * k, 256)), dim3(256), 0,
                       st, df, xf, flags, (int64_t)nflag, k, dist, idx);
    check_launch();
}

// The BF16 filters keep register lists sized for k <= 24 and d <= 256; other
// shapes use the f64 GEMM-form filter (same results).
int filter_mode(int32_t precision, int64_t d, int64_t k, int64_t n) {
    if (precision == FL_PREC_EXACT) return MODE_EXACT;
    if (precision == FL_PREC_FP8) precision = FL_PREC_BF16X3;   // FP8 is a KMeans form
    // BF16 kernels index streamed rows with int32.
    if ((precision == FL_PREC_BF16 || precision == FL_PREC_BF16X3) && bf16_nkc(d) > 0 && k <= 24 &&
        n < ((int64_t)1 << 31) - 64)
        return precision == FL_PREC_BF16 ? MODE_BF16 : MODE_BF16X3;
    return MODE_MFMA64;
}

// Top-k search for device queries of either dtype into device outputs.
void search_device(fl_knn *h, const void *dQ, int32_t dtype, int64_t nq, int k, double *dist,
                   int64_t *idx, Acc &acc) {
    const double *Q = dtype == FL_F32
                          ? as_f64<float>((const float *)dQ, nq * h->d, h->ws, "q64", h->stream)
                          : (const double *)dQ;
    if (k > kMaxFilterK) {   // beyond the filters' candidate lists
        if (h->dtype == FL_F32)
            knn_sorted<float>(h, Q, nq, k, dist, idx, acc, false);
        else
            knn_sorted<double>(h, Q, nq, k, dist, idx, acc, false);
        return;
    }
    int mode = filter_mode(h->precision, h->d, k, h->n);
    if (mode != MODE_EXACT) {
        // Range of the data: the BF16 bound needs finite values below 2^56 in
        // magnitude, the f64 filter below 2^250; other data takes the next
        // form (f64 GEMM form with its own incompleteness flag, or exact).
        unsigned int *flag = h->ws.get<unsigned int>("range", 1);
        if (h->db_range < 0) {
            FL_HIP(hipMemsetAsync(flag, 0, 4, h->stream));
            if (h->dtype == FL_F32)
                hipLaunchKernelGGL(range_flags_kernel<float>, dim3(1024), dim3(256), 0, h->stream,
                                   (const float *)h->X.ptr(), h->n * h->d, flag);
            else
                hipLaunchKernelGGL(range_flags_kernel<double>, dim3(1024), dim3(256), 0,
                                   h->stream, (const double *)h->X.ptr(), h->n * h->d, flag);
            check_launch();
            unsigned int hf = 0;
            FL_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, h->stream));
            FL_HIP(hipStreamSynchronize(h->stream));
            h->db_range = (int)hf;
        }
        unsigned int hq = 0;
        if (nq > 0) {
            FL_HIP(hipMemsetAsync(flag, 0, 4, h->stream));
            hipLaunchKernelGGL(range_flags_kernel<double>, dim3(256), dim3(256), 0, h->stream, Q,
                               nq * h->d, flag);
            check_launch();
            FL_HIP(hipMemcpyAsync(&hq, flag, 4, hipMemcpyDeviceToHost, h->stream));
            FL_HIP(hipStreamSynchronize(h->stream));
        }
        const unsigned int fl = (unsigned int)h->db_range | hq;
        h->f64_ok = (fl & 2u) == 0 && h->n < ((int64_t)1 << 31) - 64;
        if ((mode == MODE_BF16 || mode == MODE_BF16X3) && (fl & 1u)) mode = MODE_MFMA64;
        if (mode == MODE_MFMA64 && h->f64_ok && k <= 24 && f64_d4(h->d) > 0) mode = MODE_F64;
    }
    if (h->dtype == FL_F32)
        knn_search<float>(h, Q, nq, k, mode, dist, idx, acc, 0);
    else
        knn_search<double>(h, Q, nq, k, mode, dist, idx, acc, 0);
}

void fill_report(fl_report *r, const Acc &a, int64_t nq, double wall) {
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->ms_total = wall;
    r->ms_distance = a.ms_dist;
    r->ms_finalize = a.ms_fin;
    r->ms_fallback = a.ms_fb;
    r->ms_update = a.ms_upd;
    r->n_rows = nq;
    r->n_candidates = a.n_cand;
    r->n_fallback = a.n_fb;
    r->n_dist_launches = a.n_launch;
}

fl_status check_dtype(int32_t dtype) {
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    return FL_OK;
}

fl_status check_precision(int32_t p) {
    if (p != FL_PREC_FP64 && p != FL_PREC_EXACT && p != FL_PREC_BF16 && p != FL_PREC_BF16X3 &&
        p != FL_PREC_FP8)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: precision: unknown mode");
    return FL_OK;
}

// knn.rs:277-301 validation order, then the device copy.
fl_status knn_fit_common(const void *X, int32_t dtype, int64_t n, int64_t d, int64_t y_len,
                         int64_t n_neighbors, int32_t weights, int32_t precision,
                         const char *what, fl_knn **out) {
    if (!out) return fail(FL_INVALID_PARAMETER, "InvalidParameter: out: must not be NULL");
    *out = nullptr;
    if (fl_status s = check_dtype(dtype)) return s;
    if (n != y_len)
        return fail(FL_SHAPE_MISMATCH,
                    shape_msg(n, y_len, "y length must match number of samples in X"));
    if (n_neighbors == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: must be at least 1");
    if (n < n_neighbors)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(n_neighbors) + ", actual " +
                        std::to_string(n) + ": " + what + " requires at least n_neighbors=" +
                        std::to_string(n_neighbors) + " samples");
    if (n_neighbors < 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: must be at least 1");
    if (d <= 0 || d > (1 << 20))
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_features: must be in 1..2^20");
    if (n > 0x7fffffffLL)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_samples: at most 2^31-1 rows");
    if (weights != FL_UNIFORM && weights != FL_DISTANCE)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: weights: unknown scheme");
    if (fl_status s = check_precision(precision)) return s;
    if (!X) return fail(FL_INVALID_PARAMETER, "InvalidParameter: X: must not be NULL");

    fl_knn *h = new fl_knn();
    try {
        h->dtype = dtype;
        h->n = n;
        h->d = d;
        h->n_neighbors = n_neighbors;
        h->weights = weights;
        h->precision = precision;
        FL_HIP(hipStreamCreate(&h->stream));
        const size_t bytes = (size_t)(n * d) * dtype_size(dtype);
        void *dX = h->X.get(bytes);
        FL_HIP(hipMemcpyAsync(dX, X, bytes, hipMemcpyHostToDevice, h->stream));
        double *xn = (double *)h->xn.get((size_t)n * 8);
        double *xs = (double *)h->xs.get((size_t)n * 8);
        if (dtype == FL_F32)
            hipLaunchKernelGGL(row_norms_kernel<float>, dim3(grid1(n, 256)), dim3(256), 0,
                               h->stream, (const float *)dX, n, (int)d, xn, xs);
        else
            hipLaunchKernelGGL(row_norms_kernel<double>, dim3(grid1(n, 256)), dim3(256), 0,
                               h->stream, (const double *)dX, n, (int)d, xn, xs);
        check_launch();
        FL_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    return FL_OK;
}

fl_status check_query(fl_knn *h, const void *Q, int32_t dtype, int64_t nq, int64_t d) {
    if (!h) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
    if (fl_status s = check_dtype(dtype)) return s;
    if (d != h->d)
        return fail(FL_SHAPE_MISMATCH,
                    shape_msg(h->d, d, "number of features must match training data"));
    if (nq < 0) return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_queries: negative");
    if (nq > 0 && !Q) return fail(FL_INVALID_PARAMETER, "InvalidParameter: Q: must not be NULL");
    return FL_OK;
}

// k of a kneighbors call after check_query: 0 means the fitted n_neighbors,
// a negative k (a caller's explicit zero) and k == 0 after defaulting are
// InvalidParameter; k > n is InsufficientSamples for the supervised models
// (knn.rs:826-834) and InvalidParameter for NearestNeighbors
// (nearest_neighbors.rs:274-287).
fl_status check_k(fl_knn *h, int64_t &k) {
    if (k == 0) k = h->n_neighbors;
    if (k <= 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: must be at least 1");
    if (k > h->n) {
        if (h->unsupervised)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: n_neighbors=" +
                                                  std::to_string(k) +
                                                  " exceeds number of training samples=" +
                                                  std::to_string(h->n));
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(k) + ", actual " +
                        std::to_string(h->n) + ": n_neighbors exceeds number of training samples");
    }
    return FL_OK;
}

// Host-pointer wrapper: copy Q in, run on device buffers, copy out.
template <typename Run>
fl_status host_call(fl_knn *h, const void *Q, int32_t dtype, int64_t nq, Run &&run) {
    void *dQ = nullptr;
    if (nq > 0) {
        const size_t bytes = (size_t)(nq * h->d) * dtype_size(dtype);
        dQ = h->ws.get<unsigned char>("qhost", (int64_t)bytes);
        FL_HIP(hipMemcpyAsync(dQ, Q, bytes, hipMemcpyHostToDevice, h->stream));
    }
    return run(dQ);
}

fl_status kneighbors_dev(fl_knn *h, const void *dQ, int32_t dtype, int64_t nq, int64_t k,
                         double *d_dist, int64_t *d_idx, fl_report *rep) {
    WallClock wc;
    Acc acc;
    search_device(h, dQ, dtype, nq, (int)k, d_dist, d_idx, acc);
    FL_HIP(hipStreamSynchronize(h->stream));
    fill_report(rep, acc, nq, wc.ms());
    return FL_OK;
}

fl_status classify_dev(fl_knn *h, const void *dQ, int32_t dtype, int64_t nq, int64_t *d_labels,
                       double *d_proba, fl_report *rep) {
    WallClock wc;
    Acc acc;
    const int k = (int)h->n_neighbors;
    double *dist = h->ws.get<double>("kdist", nq * k);
    int64_t *idx = h->ws.get<int64_t>("kidx", nq * k);
    search_device(h, dQ, dtype, nq, k, dist, idx, acc);
    if (nq > 0) {
        PhaseTimer tm(h->stream);
        const int t = tm.begin();
        if (d_labels)
            hipLaunchKernelGGL(knn_vote_kernel, dim3(grid1(nq, 256)), dim3(256), 0, h->stream,
                               dist, idx, nq, k, (const int32_t *)h->ypos.ptr(),
                               (const int64_t *)h->classes_dev.ptr(), h->weights, d_labels);
        else
            hipLaunchKernelGGL(knn_proba_kernel, dim3(grid1(nq, 256)), dim3(256), 0, h->stream,
                               dist, idx, nq, k, (const int32_t *)h->ypos.ptr(),
                               (int)h->classes.size(), h->weights, d_proba);
        check_launch();
        tm.end(t);
        acc.ms_upd += tm.ms(t);
    }
    FL_HIP(hipStreamSynchronize(h->stream));
    fill_report(rep, acc, nq, wc.ms());
    return FL_OK;
}

fl_status regress_dev(fl_knn *h, const void *dQ, int32_t dtype, int64_t nq, double *d_y,
                      fl_report *rep) {
    WallClock wc;
    Acc acc;
    const int k = (int)h->n_neighbors;
    double *dist = h->ws.get<double>("kdist", nq * k);
    int64_t *idx = h->ws.get<int64_t>("kidx", nq * k);
    search_device(h, dQ, dtype, nq, k, dist, idx, acc);
    if (nq > 0) {
        PhaseTimer tm(h->stream);
        const int t = tm.begin();
        hipLaunchKernelGGL(knn_mean_kernel, dim3(grid1(nq, 256)), dim3(256), 0, h->stream, dist,
                           idx, nq, k, (const double *)h->yreg.ptr(), h->weights, d_y);
        check_launch();
        tm.end(t);
        acc.ms_upd += tm.ms(t);
    }
    FL_HIP(hipStreamSynchronize(h->stream));
    fill_report(rep, acc, nq, wc.ms());
    return FL_OK;
}

// ---------------------------------------------------------------------------
// kneighbors_graph (graph.rs:91-123): row i of the CSR holds the k neighbours
// of query i with their columns ascending, valued 1 (connectivity) or the
// distance.  Columns of one row are distinct, so the order is unique.
// ---------------------------------------------------------------------------

__global__ void graph_pairs_kernel(const int64_t *__restrict__ idx,
                                   const double *__restrict__ dist, int64_t total, int mode,
                                   int *__restrict__ cols, double *__restrict__ vals) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    cols[t] = (int)idx[t];
    vals[t] = mode == FL_GRAPH_DISTANCE ? dist[t] : 1.0;
}

__global__ void graph_offsets_kernel(int64_t rows, int k, int *__restrict__ offs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= rows) offs[t] = (int)(t * k);
}

__global__ void graph_out_kernel(const int *__restrict__ cols, int64_t q0, int64_t rows,
                                 int k, int64_t *__restrict__ indptr,
                                 int64_t *__restrict__ indices) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < rows * k) indices[q0 * k + t] = cols[t];
    if (t <= rows) indptr[q0 + t] = (q0 + t) * k;   // chunk ends written twice, same value
}

fl_status kneighbors_graph_dev(fl_knn *h, const void *dQ, int32_t dtype, int64_t nq, int64_t k,
                               int32_t mode, int64_t *d_indptr, int64_t *d_indices,
                               double *d_data, fl_report *rep) {
    WallClock wc;
    Acc acc;
    hipStream_t st = h->stream;
    Workspace &ws = h->ws;
    double *dist = ws.get<double>("gdist", nq * k);
    int64_t *idx = ws.get<int64_t>("gidx", nq * k);
    search_device(h, dQ, dtype, nq, (int)k, dist, idx, acc);
    if (nq == 0) {
        hipLaunchKernelGGL(graph_out_kernel, dim3(1), dim3(64), 0, st, (const int *)nullptr,
                           (int64_t)0, (int64_t)0, (int)k, d_indptr, d_indices);
        check_launch();
    }
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) <= h->n) ++bits;
    const int64_t rows_per = std::max<int64_t>(1, ((int64_t)1 << 30) / k);
    for (int64_t q0 = 0; q0 < nq; q0 += rows_per) {
        const int64_t rows = std::min<int64_t>(rows_per, nq - q0);
        const int64_t tot = rows * k;
        int *cols = ws.get<int>("gcols", tot);
        int *cols_s = ws.get<int>("gcols_s", tot);
        double *vals = ws.get<double>("gvals", tot);
        int *offs = ws.get<int>("goffs", rows + 1);
        hipLaunchKernelGGL(graph_pairs_kernel, dim3(grid1(tot, 256)), dim3(256), 0, st,
                           idx + q0 * k, dist + q0 * k, tot, mode, cols, vals);
        hipLaunchKernelGGL(graph_offsets_kernel, dim3(grid1(rows + 1, 256)), dim3(256), 0, st,
                           rows, (int)k, offs);
        check_launch();
        size_t tb = 0;
        FL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb, cols, cols_s, vals,
                                                           d_data + q0 * k, (int)tot, (int)rows,
                                                           offs, offs + 1, 0, bits, st));
        void *tmp = ws.get<unsigned char>("gtmp", (int64_t)tb + 256);
        FL_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, tb, cols, cols_s, vals,
                                                           d_data + q0 * k, (int)tot, (int)rows,
                                                           offs, offs + 1, 0, bits, st));
        hipLaunchKernelGGL(graph_out_kernel, dim3(grid1(std::max(tot, rows + 1), 256)), dim3(256),
                           0, st, cols_s, q0, rows, (int)k, d_indptr, d_indices);
        check_launch();
    }
    FL_HIP(hipStreamSynchronize(st));
    fill_report(rep, acc, nq, wc.ms());
    return FL_OK;
}

fl_status check_graph_mode(int32_t mode) {
    if (mode != FL_GRAPH_CONNECTIVITY && mode != FL_GRAPH_DISTANCE)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: mode: unknown graph mode");
    return FL_OK;
}

}  // namespace

extern "C" {

fl_status fl_knn_fit_classifier(const void *X, int32_t dtype, int64_t n, int64_t d,
                                const int64_t *y, int64_t y_len, int64_t n_neighbors,
                                int32_t weights, int32_t precision, fl_knn_handle *out) {
    return guarded([&]() -> fl_status {
        fl_knn *h = nullptr;
        fl_status s = knn_fit_common(X, dtype, n, d, y_len, n_neighbors, weights, precision,
                                     "KNeighborsClassifier", &h);
        if (s) return s;
        try {
            h->is_classifier = 1;
            // knn.rs:304-306: classes = sort + dedup of y.
            h->classes.assign(y, y + n);
            std::sort(h->classes.begin(), h->classes.end());
            h->classes.erase(std::unique(h->classes.begin(), h->classes.end()),
                             h->classes.end());
            std::vector<int32_t> pos((size_t)n);
            for (int64_t i = 0; i < n; ++i)
                pos[(size_t)i] = (int32_t)(std::lower_bound(h->classes.begin(),
                                                            h->classes.end(), y[i]) -
                                           h->classes.begin());
            void *dp = h->ypos.get((size_t)n * 4);
            FL_HIP(hipMemcpy(dp, pos.data(), (size_t)n * 4, hipMemcpyHostToDevice));
            void *dc = h->classes_dev.get(h->classes.size() * 8);
            FL_HIP(hipMemcpy(dc, h->classes.data(), h->classes.size() * 8,
                             hipMemcpyHostToDevice));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return FL_OK;
    });
}

fl_status fl_knn_fit_regressor(const void *X, int32_t dtype, int64_t n, int64_t d,
                               const double *y, int64_t y_len, int64_t n_neighbors,
                               int32_t weights, int32_t precision, fl_knn_handle *out) {
    return guarded([&]() -> fl_status {
        fl_knn *h = nullptr;
        fl_status s = knn_fit_common(X, dtype, n, d, y_len, n_neighbors, weights, precision,
                                     "KNeighborsRegressor", &h);
        if (s) return s;
        try {
            void *dy = h->yreg.get((size_t)n * 8);
            FL_HIP(hipMemcpy(dy, y, (size_t)n * 8, hipMemcpyHostToDevice));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return FL_OK;
    });
}

fl_status fl_knn_kneighbors_device(fl_knn_handle h, const void *dQ, int32_t dtype, int64_t nq,
                                   int64_t d, int64_t k, double *d_dist_out, int64_t *d_idx_out,
                                   fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, dQ, dtype, nq, d)) return s;
        if (fl_status s = check_k(h, k)) return s;
        return kneighbors_dev(h, dQ, dtype, nq, k, d_dist_out, d_idx_out, report);
    });
}

fl_status fl_knn_kneighbors(fl_knn_handle h, const void *Q, int32_t dtype, int64_t nq, int64_t d,
                            int64_t k, double *dist_out, int64_t *idx_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, Q, dtype, nq, d)) return s;
        if (fl_status s = check_k(h, k)) return s;
        return host_call(h, Q, dtype, nq, [&](const void *dQ) -> fl_status {
            double *dd = h->ws.get<double>("odist", nq * k);
            int64_t *di = h->ws.get<int64_t>("oidx", nq * k);
            fl_status s = kneighbors_dev(h, dQ, dtype, nq, k, dd, di, report);
            if (s) return s;
            if (nq > 0) {
                FL_HIP(hipMemcpy(dist_out, dd, (size_t)(nq * k) * 8, hipMemcpyDeviceToHost));
                FL_HIP(hipMemcpy(idx_out, di, (size_t)(nq * k) * 8, hipMemcpyDeviceToHost));
            }
            return FL_OK;
        });
    });
}

fl_status fl_knn_predict_classify_device(fl_knn_handle h, const void *dQ, int32_t dtype,
                                         int64_t nq, int64_t d, int64_t *d_labels_out,
                                         fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, dQ, dtype, nq, d)) return s;
        if (!h->is_classifier)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: not a classifier");
        return classify_dev(h, dQ, dtype, nq, d_labels_out, nullptr, report);
    });
}

fl_status fl_knn_predict_classify(fl_knn_handle h, const void *Q, int32_t dtype, int64_t nq,
                                  int64_t d, int64_t *labels_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, Q, dtype, nq, d)) return s;
        if (!h->is_classifier)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: not a classifier");
        return host_call(h, Q, dtype, nq, [&](const void *dQ) -> fl_status {
            int64_t *dl = h->ws.get<int64_t>("olab", nq);
            fl_status s = classify_dev(h, dQ, dtype, nq, dl, nullptr, report);
            if (s) return s;
            if (nq > 0)
                FL_HIP(hipMemcpy(labels_out, dl, (size_t)nq * 8, hipMemcpyDeviceToHost));
            return FL_OK;
        });
    });
}

fl_status fl_knn_predict_proba(fl_knn_handle h, const void *Q, int32_t dtype, int64_t nq,
                               int64_t d, double *proba_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, Q, dtype, nq, d)) return s;
        if (!h->is_classifier)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: not a classifier");
        const int64_t C = (int64_t)h->classes.size();
        return host_call(h, Q, dtype, nq, [&](const void *dQ) -> fl_status {
            double *dp = h->ws.get<double>("oproba", nq * C);
            fl_status s = classify_dev(h, dQ, dtype, nq, nullptr, dp, report);
            if (s) return s;
            if (nq > 0)
                FL_HIP(hipMemcpy(proba_out, dp, (size_t)(nq * C) * 8, hipMemcpyDeviceToHost));
            return FL_OK;
        });
    });
}

fl_status fl_knn_predict_regress_device(fl_knn_handle h, const void *dQ, int32_t dtype,
                                        int64_t nq, int64_t d, double *d_y_out,
                                        fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, dQ, dtype, nq, d)) return s;
        if (h->is_classifier)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: not a regressor");
        return regress_dev(h, dQ, dtype, nq, d_y_out, report);
    });
}

fl_status fl_knn_predict_regress(fl_knn_handle h, const void *Q, int32_t dtype, int64_t nq,
                                 int64_t d, double *y_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, Q, dtype, nq, d)) return s;
        if (h->is_classifier)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: not a regressor");
        return host_call(h, Q, dtype, nq, [&](const void *dQ) -> fl_status {
            double *dy = h->ws.get<double>("oy", nq);
            fl_status s = regress_dev(h, dQ, dtype, nq, dy, report);
            if (s) return s;
            if (nq > 0) FL_HIP(hipMemcpy(y_out, dy, (size_t)nq * 8, hipMemcpyDeviceToHost));
            return FL_OK;
        });
    });
}

fl_status fl_knn_set_precision(fl_knn_handle h, int32_t precision) {
    return guarded([&]() -> fl_status {
        if (!h) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
        if (fl_status s = check_precision(precision)) return s;
        h->precision = precision;
        return FL_OK;
    });
}

int64_t fl_knn_n_classes(fl_knn_handle h) { return h ? (int64_t)h->classes.size() : 0; }

fl_status fl_knn_classes(fl_knn_handle h, int64_t *classes_out) {
    if (!h) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
    std::copy(h->classes.begin(), h->classes.end(), classes_out);
    return FL_OK;
}

fl_status fl_nn_fit(const void *X, int32_t dtype, int64_t n, int64_t d, int64_t n_neighbors,
                    int32_t precision, fl_knn_handle *out) {
    return guarded([&]() -> fl_status {
        fl_knn *h = nullptr;
        fl_status s = knn_fit_common(X, dtype, n, d, n, n_neighbors, FL_UNIFORM, precision,
                                     "NearestNeighbors", &h);
        if (s) return s;
        h->unsupervised = 1;
        *out = h;
        return FL_OK;
    });
}

fl_status fl_knn_kneighbors_graph_device(fl_knn_handle h, const void *dQ, int32_t dtype,
                                         int64_t nq, int64_t d, int64_t k, int32_t mode,
                                         int64_t *d_indptr_out, int64_t *d_indices_out,
                                         double *d_data_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, dQ, dtype, nq, d)) return s;
        if (fl_status s = check_k(h, k)) return s;
        if (fl_status s = check_graph_mode(mode)) return s;
        if (!d_indptr_out || (nq > 0 && (!d_indices_out || !d_data_out)))
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: outputs: must not be NULL");
        return kneighbors_graph_dev(h, dQ, dtype, nq, k, mode, d_indptr_out, d_indices_out,
                                    d_data_out, report);
    });
}

fl_status fl_knn_kneighbors_graph(fl_knn_handle h, const void *Q, int32_t dtype, int64_t nq,
                                  int64_t d, int64_t k, int32_t mode, int64_t *indptr_out,
                                  int64_t *indices_out, double *data_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_query(h, Q, dtype, nq, d)) return s;
        if (fl_status s = check_k(h, k)) return s;
        if (fl_status s = check_graph_mode(mode)) return s;
        if (!indptr_out || (nq > 0 && (!indices_out || !data_out)))
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: outputs: must not be NULL");
        return host_call(h, Q, dtype, nq, [&](const void *dQ) -> fl_status {
            int64_t *dp = h->ws.get<int64_t>("g_indptr", nq + 1);
            int64_t *di = h->ws.get<int64_t>("g_indices", nq * k);
            double *dd = h->ws.get<double>("g_data", nq * k);
            fl_status s = kneighbors_graph_dev(h, dQ, dtype, nq, k, mode, dp, di, dd, report);
            if (s) return s;
            FL_HIP(hipMemcpy(indptr_out, dp, (size_t)(nq + 1) * 8, hipMemcpyDeviceToHost));
            if (nq > 0) {
                FL_HIP(hipMemcpy(indices_out, di, (size_t)(nq * k) * 8, hipMemcpyDeviceToHost));
                FL_HIP(hipMemcpy(data_out, dd, (size_t)(nq * k) * 8, hipMemcpyDeviceToHost));
            }
            return FL_OK;
        });
    });
}

int64_t fl_knn_n_samples_fit(fl_knn_handle h) { return h ? h->n : 0; }

void fl_knn_free(fl_knn_handle h) { delete h; }

}  // extern "C"

namespace fl {
// The fitted handle's device rows and stream, for the cell router built on it.
void knn_describe(const fl_knn *h, const void **X, int32_t *dtype, int64_t *n, int64_t *d) {
    *X = h->X.ptr();
    *dtype = h->dtype;
    *n = h->n;
    *d = h->d;
}
int64_t knn_n_neighbors(const fl_knn *h) { return h->n_neighbors; }
}  // namespace fl
