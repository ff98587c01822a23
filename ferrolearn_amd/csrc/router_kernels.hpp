// router_kernels.hpp -- kernels of the routed exact kNN search.
//
// The database is partitioned into cells (a coarse KMeans); each cell c has
// a centre mu_c and a radius r_c >= max over its rows of |x - mu_c|.  A query
// scans the rows of n_probe cells, chosen by the router network or by the
// cells' triangle-inequality lower bounds, with the exact reference
// distance sqrt(fold) (kdtree.rs:342-348) and the (distance, index) order of
// brute_force_knn (kdtree.rs:364-391).  The answer is certified when every
// unscanned cell's lower bound on the reference distance of its rows exceeds
// the k-th distance found; other queries go to the brute-force search.
#pragma once

#include "common.hpp"
#include "dist_bf16.hpp"

namespace fl {

constexpr int RT_Q = 64;   // lists (queries or query-cell pairs) per block
constexpr int RT_R = 64;   // rows per tile
constexpr int RT_F = 32;   // features per staged chunk
constexpr int RT_FPAD = RT_R + 1;   // fold buffer row stride (doubles)
constexpr int32_t RT_NOID = 0x7fffffff;

// Insert (key, id) into an ascending (key, id) list of K entries held in
// registers, dropping the last entry.  Caller has checked that the pair is
// below the last entry.
template <int K>
__device__ __forceinline__ void rt_insert(double (&lk)[K], int32_t (&li)[K], double key,
                                          int32_t id) {
    bool placed = false;
#pragma unroll
    for (int s = K - 1; s > 0; --s) {
        if (!placed) {
            if (pair_less(key, id, lk[s - 1], li[s - 1])) {
                lk[s] = lk[s - 1];
                li[s] = li[s - 1];
            } else {
                lk[s] = key;
                li[s] = id;
                placed = true;
            }
        }
    }
    if (!placed) {
        lk[0] = key;
        li[0] = id;
    }
}

// Lower bound on the reference distance sqrt(fold(q, x)) of every row x of
// a cell, from rho = sqrt(fold(q, mu)) and r = computed radius:
// |q - mu| >= rho (1 - e), |x - mu| <= r (1 + e) and sqrt(fold(q, x)) >=
// |q - x| (1 - e) with e covering the fold's d+2 roundings and the sqrt's;
// ep = e plus the roundings of this expression.  NaN gives -inf (no bound).
__device__ __forceinline__ double rt_lower_bound(double rho, double r, double ep) {
    const double a = rho * (1.0 - ep);
    const double b = r * (1.0 + ep);
    const double lb = (a - b) * (1.0 - ep);
    return lb == lb ? lb : -__builtin_huge_val();
}

// Lower bound from the bisector of the query's nearest centre a and cell c
// (cells are exact nearest-centre cells, so every row x of c has
// fold(x, mu_c) <= fold(x, mu_a)).  In real arithmetic that puts x in the
// half-space 2 x.(mu_a - mu_c) <= |mu_a|^2 - |mu_c|^2 + s, with s = 2.02 g
// (|x - mu_c| + |mu_c - mu_a|)^2 covering the folds' roundings, so
// |q - x| >= (|q - mu_c|^2 - |q - mu_a|^2 - s) / (2 |mu_c - mu_a|).  fc, fa:
// computed folds of q with mu_c and mu_a; D: computed |mu_c - mu_a|; r:
// computed radius of c; g bounds the relative error of a fold (d + 2
// roundings) or of a distance, plus the roundings of this expression.
__device__ __forceinline__ double rt_bisector_bound(double fc, double fa, double D, double r,
                                                    double g) {
    const double w = (r + D) * (1.0 + g);
    const double s = 2.02 * g * w * w;
    const double num = fc * (1.0 - g) - fa * (1.0 + g) - s;
    const double lb = num / (2.0 * D * (1.0 + g)) * (1.0 - g);
    return lb == lb ? lb : -__builtin_huge_val();
}

// Shared core of the two list kernels: 64 lists, each fed by the exact folds
// of one query row against the rows [r_begin, r_end) of X.  Thread t folds
// queries (t & 15) + 16i against rows (t >> 4) + 16j of each 64-row tile, in
// 32-feature chunks staged as f64 in LDS (the fold stays a left fold in
// feature order).  The tile's folds then go through LDS to the list owners:
// thread t keeps list (t & 63) for rows 16 (t >> 6) .. +15 of every tile.
// key(fold, row, &key, &id) maps a fold to its list key; thr(list_last_key)
// gives a fold above which the key cannot enter the list.
template <typename TQ, typename TX, int K, typename KeyFn, typename ThrFn>
__device__ __forceinline__ void rt_fold_lists(const TQ *const *qrow, int d, const TX *__restrict__ X,
                                              int64_t r_begin, int64_t r_end, double *lds,
                                              double (&lk)[K], int32_t (&li)[K], KeyFn keyfn,
                                              ThrFn thrfn) {
    const int t = threadIdx.x;
    const int qg = t & 15, rg = t >> 4;
    const int own = t & 63, sub = t >> 6;
    double *Qs = lds;                  // [RT_F][RT_Q]
    double *Xt = lds + RT_F * RT_Q;    // [RT_F][RT_R]
    double *F = lds;                   // [RT_Q][RT_FPAD], after the folds
    double thr = thrfn(lk[K - 1]);
    for (int64_t r0 = r_begin; r0 < r_end; r0 += RT_R) {
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int f0 = 0; f0 < d; f0 += RT_F) {
            {
                const int q = t >> 2, fb = (t & 3) * 8;
                const TQ *qp = qrow[q];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int f = f0 + fb + u;
                    Qs[(fb + u) * RT_Q + q] = (qp && f < d) ? (double)qp[f] : 0.0;
                }
                const int64_t row = r0 + q;
                const TX *xp = X + row * d;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int f = f0 + fb + u;
                    Xt[(fb + u) * RT_R + q] = (row < r_end && f < d) ? (double)xp[f] : 0.0;
                }
            }
            __syncthreads();
            const int fc = d - f0 < RT_F ? d - f0 : RT_F;
            auto step = [&](int f) {
                double qv[4], xv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) qv[i] = Qs[f * RT_Q + qg + 16 * i];
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = Xt[f * RT_R + rg + 16 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double df = qv[i] - xv[j];
                        acc[i][j] = acc[i][j] + df * df;
                    }
            };
            if (fc == RT_F) {
#pragma unroll 8
                for (int f = 0; f < RT_F; ++f) step(f);
            } else {
                for (int f = 0; f < fc; ++f) step(f);
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) F[(qg + 16 * i) * RT_FPAD + rg + 16 * j] = acc[i][j];
        __syncthreads();
        for (int jj = 0; jj < 16; ++jj) {
            const int r = sub * 16 + jj;
            const int64_t row = r0 + r;
            if (row >= r_end) break;
            const double fold = F[own * RT_FPAD + r];
            if (fold > thr) continue;   // NaN goes on to the key function
            double key;
            int32_t id;
            keyfn(fold, row, key, id);
            if (pair_less(key, id, lk[K - 1], li[K - 1])) {
                rt_insert<K>(lk, li, key, id);
                thr = thrfn(lk[K - 1]);
            }
        }
        __syncthreads();
    }
}

// Merge the four sub-lists of each of the 64 lists (one per wave) and hand
// the first kout entries of list `own` to emit(own, j, key, id), run by the
// threads of wave 0.
template <int K, typename Emit>
__device__ __forceinline__ void rt_merge4(double *lds, const double (&lk)[K],
                                          const int32_t (&li)[K], int kout, Emit emit) {
    const int t = threadIdx.x;
    const int own = t & 63, sub = t >> 6;
    double *Lk = lds;                                        // [4][64][K]
    int32_t *Li = reinterpret_cast<int32_t *>(lds + 4 * RT_Q * K);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        Lk[(sub * RT_Q + own) * K + s] = lk[s];
        Li[(sub * RT_Q + own) * K + s] = li[s];
    }
    __syncthreads();
    if (sub == 0) {
        int pos[4] = {0, 0, 0, 0};
        for (int j = 0; j < kout; ++j) {
            int best = 0;
            double bk = Lk[(0 * RT_Q + own) * K + pos[0]];
            int32_t bi = Li[(0 * RT_Q + own) * K + pos[0]];
#pragma unroll
            for (int s = 1; s < 4; ++s) {
                const double k2 = Lk[(s * RT_Q + own) * K + pos[s]];
                const int32_t i2 = Li[(s * RT_Q + own) * K + pos[s]];
                if (pair_less(k2, i2, bk, bi)) {
                    best = s;
                    bk = k2;
                    bi = i2;
                }
            }
            emit(own, j, bk, bi);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s == best) ++pos[s];
        }
    }
}

inline size_t rt_lds_bytes(int K) {
    const size_t folds = (size_t)RT_Q * RT_FPAD * 8;
    const size_t merge = (size_t)4 * RT_Q * K * 12;
    return folds > merge ? folds : merge;
}

// Scan: one block per work item (a cell and up to 64 of the query-cell pairs
// routed to it).  Pair p = query * n_probe + slot; its list holds the k
// smallest (sqrt(fold), global index) over the cell's rows, written to
// out_key / out_id [p][k] (absent entries: +inf, RT_NOID).
struct RtScanArgs {
    const void *Q;                 // queries, nq x d (TQ)
    const void *Xs;                // cell-sorted rows, n x d (TX)
    const int32_t *row_id;         // global index of each sorted row
    const int64_t *cell_first;     // first sorted row of each cell
    const int32_t *cell_count;     // rows per cell
    const int32_t *pair_order;     // pairs sorted by cell (stable)
    const int64_t *pair_first;     // first sorted pair of each cell
    const int32_t *pair_count;     // pairs per cell
    const int32_t *work_cell;      // per block: cell
    const int32_t *work_start;     // per block: offset into the cell's pairs
    int d, n_probe, k;
    double *out_key;
    int32_t *out_id;
};

template <typename TQ, typename TX, int K>
__global__ __launch_bounds__(256) void rt_scan_kernel(RtScanArgs a) {
    extern __shared__ double lds[];
    __shared__ const TQ *qrow[RT_Q];
    __shared__ int32_t pid[RT_Q];
    const int t = threadIdx.x;
    const int c = a.work_cell[blockIdx.x];
    const int s0 = a.work_start[blockIdx.x];
    const int np = min(RT_Q, a.pair_count[c] - s0);
    if (t < RT_Q) {
        if (t < np) {
            const int32_t p = a.pair_order[a.pair_first[c] + s0 + t];
            pid[t] = p;
            qrow[t] = (const TQ *)a.Q + (int64_t)(p / a.n_probe) * a.d;
        } else {
            pid[t] = -1;
            qrow[t] = nullptr;
        }
    }
    __syncthreads();
    double lk[K];
    int32_t li[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        lk[s] = __builtin_huge_val();
        li[s] = RT_NOID;
    }
    const int64_t rb = a.cell_first[c], re = rb + a.cell_count[c];
    const int32_t *rid = a.row_id;
    rt_fold_lists<TQ, TX, K>(
        qrow, a.d, (const TX *)a.Xs, rb, re, lds, lk, li,
        [&](double fold, int64_t row, double &key, int32_t &id) {
            key = sqrt(fold);
            id = rid[row];
        },
        // A fold above last^2 (1 + 2^-48) has a correctly rounded sqrt above
        // `last`, so it cannot enter the list.
        [](double last) { return last * last * (1.0 + 0x1.0p-48); });
    double *ok = a.out_key;
    int32_t *oi = a.out_id;
    const int kk = a.k;
    rt_merge4<K>(lds, lk, li, kk, [&](int own, int j, double key, int32_t id) {
        const int32_t p = pid[own];
        if (p >= 0) {
            ok[(int64_t)p * kk + j] = key;
            oi[(int64_t)p * kk + j] = id;
        }
    });
}

// Bounds: one block per 64 queries; list = the kb smallest (lower bound,
// cell) over all cells (empty cells: +inf), written to out_key / out_id
// [q][kb].
struct RtBoundsArgs {
    const void *Q;
    int64_t nq;
    int d;
    const double *centers;         // C x d
    const double *radius;          // C (negative: empty cell)
    const double *cdist;           // C x C centre distances
    const double *near_fold;       // per query: fold with its nearest centre
    const int32_t *near_cell;      // per query: its nearest centre
    int C;
    double ep;
    int kb;
    int centre_key;                // 1: key = fold with the centre (no bounds)
    double *out_key;
    int32_t *out_id;
};

// The key of a cell is the larger of its ball bound and its bisector bound
// against the query's nearest centre; with centre_key, the query's fold with
// the cell's centre (the inverted-file probe order).  An empty cell's key is
// +inf.
template <typename TQ, int K>
__global__ __launch_bounds__(256) void rt_bounds_kernel(RtBoundsArgs a) {
    extern __shared__ double lds[];
    __shared__ const TQ *qrow[RT_Q];
    const int t = threadIdx.x;
    const int64_t q0 = (int64_t)blockIdx.x * RT_Q;
    if (t < RT_Q) qrow[t] = q0 + t < a.nq ? (const TQ *)a.Q + (q0 + t) * a.d : nullptr;
    __syncthreads();
    double lk[K];
    int32_t li[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        lk[s] = __builtin_huge_val();
        li[s] = RT_NOID;
    }
    const double *rad = a.radius;
    const double ep = a.ep;
    const int64_t qo = q0 + (t & 63);
    const int32_t near = qo < a.nq ? a.near_cell[qo] : 0;
    const double fnear = qo < a.nq ? a.near_fold[qo] : 0.0;
    const double *drow = a.cdist + (int64_t)near * a.C;
    rt_fold_lists<TQ, double, K>(
        qrow, a.d, a.centers, 0, a.C, lds, lk, li,
        [&](double fold, int64_t c, double &key, int32_t &id) {
            id = (int32_t)c;
            const double r = rad[c];
            if (r < 0.0) {
                key = __builtin_huge_val();
                return;
            }
            if (a.centre_key) {
                key = fold;
                return;
            }
            key = rt_lower_bound(sqrt(fold), r, ep);
            const double D = drow[c];
            if (c != near && D > 0.0) {
                const double v = rt_bisector_bound(fold, fnear, D, r, ep);
                key = v > key ? v : key;
            }
        },
        [](double) { return __builtin_huge_val(); });
    double *ok = a.out_key;
    int32_t *oi = a.out_id;
    const int kb = a.kb;
    const int64_t nq = a.nq;
    rt_merge4<K>(lds, lk, li, kb, [&](int own, int j, double key, int32_t id) {
        if (q0 + own < nq) {
            ok[(q0 + own) * kb + j] = key;
            oi[(q0 + own) * kb + j] = id;
        }
    });
}

// The same lists and keys with every fold replaced by a lower bound from a
// BF16x3 MFMA dot product: q.mu as hi*hi + hi*lo + lo*hi (bf16 splits of the
// f64-widened values) in one FP32 accumulator, then in f64
//   flo = |q|^2 + |mu|^2 - 2 q.mu - (alpha |q||mu| + beta (|q|^2 + |mu|^2))(1 + 2^-50)
//         - 2^-50 (|q|^2 + |mu|^2) - eta (1 + |q| + |mu|),
// which the kNN filters' error bound (alpha, beta of bound_bf16 with three
// passes; eta for FP32 flushing) puts at or below the exact fold; the 2^-50
// terms cover this expression's own roundings.  flo is kept as the f32 at or
// below it.  The ball bound grows with sqrt(fold) and the bisector bound with
// fold(q, mu_c), so both stay valid lower bounds with flo in place of the
// fold, and the certificate that uses them is unchanged.  Queries and
// centres must be finite with magnitudes <= 2^56 (the BF16 filters' range);
// other searches use rt_bounds_kernel.
//
// One block of 256 threads per 64 queries, 64-centre tiles: the queries' and
// the tile's hi / lo rows sit in LDS with a row stride of dp + 8 halfwords
// (16-byte MFMA operand loads of 16 consecutive rows hit distinct banks);
// wave w computes the 32 x 32 (query, centre) quarter (w >> 1, w & 1) as
// 2 x 2 tiles of v_mfma_f32_16x16x32_bf16, writes its lower bounds to the
// [query][centre] tile in LDS, and the list owners of rt_fold_lists read
// them.
constexpr int RTM_MAXDP = 128;   // padded features the LDS plan allows
constexpr int RTM_FS = 68;       // lower-bound tile row stride (floats): 16-byte rows

struct RtBoundsMfmaArgs {
    RtBoundsArgs b;                  // queries, cells, nearest centres, lists
    const uint16_t *qh, *ql;         // nq x dp bf16 hi / lo of the queries
    const double *qn2, *qn1;         // |q|^2, |q|
    const uint16_t *ch, *cl;         // C x dp bf16 hi / lo of the centres
    const double *cn2, *cn1;         // |mu|^2, |mu|
    int dp;                          // features padded to a multiple of 32
    double alpha, beta, eta;
};

inline size_t rtm_lds_bytes(int K, int dp) {
    const size_t rows = (size_t)2 * 2 * RT_Q * (dp + 8) * 2;   // q and centre tiles, hi / lo
    const size_t norms = (size_t)6 * RT_Q * 8;
    const size_t tile = (size_t)RT_Q * RTM_FS * 4;             // f32 lower bounds
    const size_t merge = (size_t)4 * RT_Q * K * 12;
    const size_t work = rows + norms + tile;
    return work > merge ? work : merge;
}

template <int K>
__global__ __launch_bounds__(256) void rt_bounds_mfma_kernel(RtBoundsMfmaArgs m) {
    extern __shared__ __align__(16) unsigned char rtm_lds[];
    const RtBoundsArgs &a = m.b;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int dp = m.dp, ds = dp + 8;
    uint16_t *Qh = reinterpret_cast<uint16_t *>(rtm_lds);
    uint16_t *Ql = Qh + RT_Q * ds;
    uint16_t *Ch = Ql + RT_Q * ds;
    uint16_t *Cl = Ch + RT_Q * ds;
    double *sq2 = reinterpret_cast<double *>(Cl + RT_Q * ds);
    double *sq1 = sq2 + RT_Q;
    double *sc2 = sq1 + RT_Q;
    double *sc1 = sc2 + RT_Q;
    double *srr = sc1 + RT_Q;                           // r_c (1 + ep) / (1 - ep)
    double *srad = srr + RT_Q;                          // r_c
    float *F = reinterpret_cast<float *>(srad + RT_Q);  // [RT_Q][RTM_FS]
    const int64_t q0 = (int64_t)blockIdx.x * RT_Q;
    const int chunks = dp / 8;   // 16-byte chunks per row
    // the block's queries, once
    for (int e = t; e < RT_Q * chunks; e += 256) {
        const int r = e / chunks, c = e % chunks;
        u16x8_t h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
        if (q0 + r < a.nq) {
            h = *reinterpret_cast<const u16x8_t *>(m.qh + (q0 + r) * dp + c * 8);
            l = *reinterpret_cast<const u16x8_t *>(m.ql + (q0 + r) * dp + c * 8);
        }
        *reinterpret_cast<u16x8_t *>(Qh + r * ds + c * 8) = h;
        *reinterpret_cast<u16x8_t *>(Ql + r * ds + c * 8) = l;
    }
    if (t < RT_Q) {
        sq2[t] = q0 + t < a.nq ? m.qn2[q0 + t] : 0.0;
        sq1[t] = q0 + t < a.nq ? m.qn1[q0 + t] : 0.0;
    }
    double lk[K];
    int32_t li[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        lk[s] = __builtin_huge_val();
        li[s] = RT_NOID;
    }
    const int own = t & 63, sub = t >> 6;
    const int64_t qo = q0 + own;
    const int32_t near = qo < a.nq ? a.near_cell[qo] : 0;
    const double fnear = qo < a.nq ? a.near_fold[qo] : 0.0;
    const double *drow = a.cdist + (int64_t)near * a.C;
    const double ep = a.ep;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    for (int c0 = 0; c0 < a.C; c0 += RT_R) {
        for (int e = t; e < RT_R * chunks; e += 256) {
            const int r = e / chunks, c = e % chunks;
            u16x8_t h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
            if (c0 + r < a.C) {
                h = *reinterpret_cast<const u16x8_t *>(m.ch + (int64_t)(c0 + r) * dp + c * 8);
                l = *reinterpret_cast<const u16x8_t *>(m.cl + (int64_t)(c0 + r) * dp + c * 8);
            }
            *reinterpret_cast<u16x8_t *>(Ch + r * ds + c * 8) = h;
            *reinterpret_cast<u16x8_t *>(Cl + r * ds + c * 8) = l;
        }
        if (t < RT_R) {
            sc2[t] = c0 + t < a.C ? m.cn2[c0 + t] : 0.0;
            sc1[t] = c0 + t < a.C ? m.cn1[c0 + t] : 0.0;
            const double rc = c0 + t < a.C ? a.radius[c0 + t] : -1.0;
            srr[t] = rc < 0.0 ? __builtin_huge_val() : rc * (1.0 + ep) / (1.0 - ep);
            srad[t] = rc;
        }
        __syncthreads();
        float4_t acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < dp; k0 += 32) {
            const int ko = k0 + (lane >> 4) * 8;
            bf16x8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ra = wm + i * 16 + (lane & 15), rb = wn + i * 16 + (lane & 15);
                ah[i] = __builtin_bit_cast(bf16x8_t,
                                           *reinterpret_cast<const u16x8_t *>(Qh + ra * ds + ko));
                al[i] = __builtin_bit_cast(bf16x8_t,
                                           *reinterpret_cast<const u16x8_t *>(Ql + ra * ds + ko));
                bh[i] = __builtin_bit_cast(bf16x8_t,
                                           *reinterpret_cast<const u16x8_t *>(Ch + rb * ds + ko));
                bl[i] = __builtin_bit_cast(bf16x8_t,
                                           *reinterpret_cast<const u16x8_t *>(Cl + rb * ds + ko));
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j],
                                                                        0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j],
                                                                        0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j],
                                                                        0, 0, 0);
                }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int mq = wm + i * 16 + bf_acc_row(lane, v);
                    const int nc = wn + j * 16 + (lane & 15);
                    const double q2 = sq2[mq], q1 = sq1[mq], c2 = sc2[nc], c1 = sc1[nc];
                    const double s2 = q2 + c2;
                    const double val = s2 - 2.0 * (double)acc[i][j][v];
                    const double err = (m.alpha * q1 * c1 + m.beta * s2) * (1.0 + 0x1p-50) +
                                       0x1p-50 * s2 + m.eta * (1.0 + q1 + c1);
                    const double flo = val - err;
                    float f = (float)flo;
                    if ((double)f > flo) {   // one f32 step down (f is not NaN or -inf here)
                        const uint32_t fb = __builtin_bit_cast(uint32_t, f);
                        f = __builtin_bit_cast(
                            float, f > 0.f ? fb - 1u : (f < 0.f ? fb + 1u : 0x80000001u));
                    }
                    F[mq * RTM_FS + nc] = f;
                }
        __syncthreads();
        // A cell whose ball bound alone exceeds the list's last key T cannot
        // enter: ball > T once sqrt(flo) > S = T / (1-ep)^2 + r (1+ep)/(1-ep).
        // S is widened by 2^-40 (|T'| + r') before the test flo > S^2 (1 +
        // 2^-40), which leaves room for the roundings of S (T' + r' may
        // cancel) and of the key itself, so a skipped pair's computed key is
        // above T.  The owner's 16 bounds are tested together; only the
        // candidates left compute their key.
        const double tpf = 1.0 / ((1.0 - ep) * (1.0 - ep));
        const float *Fr = F + own * RTM_FS + sub * 16;
        float fv[16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4_t v4 = *reinterpret_cast<const float4_t *>(Fr + 4 * u);
            fv[4 * u] = v4[0], fv[4 * u + 1] = v4[1], fv[4 * u + 2] = v4[2], fv[4 * u + 3] = v4[3];
        }
        auto skip = [&](double flo, int r, double tl) {
            if (a.centre_key) return flo > tl;
            const double tp = tl * tpf;
            const double S = tp + srr[r] + 0x1p-40 * (__builtin_fabs(tp) + srr[r]);
            return S > 0.0 && flo > S * S * (1.0 + 0x1p-40);
        };
        unsigned cand = 0;
        {
            const double tl = lk[K - 1];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int r = sub * 16 + jj;
                const bool keep = c0 + r < a.C && !skip((double)fv[jj], r, tl);
                cand |= (unsigned)keep << jj;
            }
        }
#pragma unroll 1
        while (cand) {
            const int jj = __builtin_ctz(cand);
            cand &= cand - 1;
            const int r = sub * 16 + jj;
            const int c = c0 + r;
            const double flo = (double)Fr[jj];
            if (skip(flo, r, lk[K - 1])) continue;   // the list moved on
            double key;
            const double rc = srad[r];
            if (rc < 0.0) {
                key = __builtin_huge_val();
            } else if (a.centre_key) {
                key = flo;
            } else {
                const double fl0 = flo < 0.0 ? 0.0 : flo;   // NaN stays NaN: no bound
                key = rt_lower_bound(sqrt(fl0), rc, ep);
                const double D = drow[c];
                if (c != near && D > 0.0) {
                    const double vb = rt_bisector_bound(fl0, fnear, D, rc, ep);
                    key = vb > key ? vb : key;
                }
            }
            if (pair_less(key, (int32_t)c, lk[K - 1], li[K - 1])) rt_insert<K>(lk, li, key, c);
        }
        __syncthreads();
    }
    double *lds = reinterpret_cast<double *>(rtm_lds);
    double *ok = a.out_key;
    int32_t *oi = a.out_id;
    const int kb = a.kb;
    const int64_t nq = a.nq;
    rt_merge4<K>(lds, lk, li, kb, [&](int o, int j, double key, int32_t id) {
        if (q0 + o < nq) {
            ok[(q0 + o) * kb + j] = key;
            oi[(q0 + o) * kb + j] = id;
        }
    });
}

// Router choice: the n_probe cells with the largest logits (ties: lower
// cell), per row, one wave per row.  Each lane keeps its best M in
// registers; the wave then pops the best head n_probe times.
template <int M>
__global__ __launch_bounds__(256) void rt_top_cells_kernel(const float *__restrict__ Z, int64_t ldz,
                                                           int64_t rows, int C, int n_probe,
                                                           int32_t *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const float *z = Z + row * ldz;
    // keys: -logit ascending, so the list code orders (key, cell) ascending
    double lk[M];
    int32_t li[M];
#pragma unroll
    for (int s = 0; s < M; ++s) {
        lk[s] = __builtin_huge_val();
        li[s] = RT_NOID;
    }
    for (int c = lane; c < C; c += 64) {
        float v = z[c];
        const double key = v == v ? -(double)v : __builtin_huge_val();
        if (pair_less(key, c, lk[M - 1], li[M - 1])) rt_insert<M>(lk, li, key, c);
    }
    int head = 0;
    for (int j = 0; j < n_probe; ++j) {
        double hk = __builtin_huge_val();
        int32_t hi = RT_NOID;
#pragma unroll
        for (int s = 0; s < M; ++s)
            if (s == head) {
                hk = lk[s];
                hi = li[s];
            }
        double bk = hk;
        int32_t bi = hi;
        for (int o = 32; o > 0; o >>= 1) {
            const double k2 = __shfl_xor(bk, o);
            const int32_t i2 = __shfl_xor(bi, o);
            if (pair_less(k2, i2, bk, bi)) {
                bk = k2;
                bi = i2;
            }
        }
        if (hi == bi && hk == bk) ++head;
        if (lane == 0) out[row * n_probe + j] = bi == RT_NOID ? 0 : bi;
    }
}

// Whether (key, id) is already in a list.
template <int K>
__device__ __forceinline__ bool rt_contains(const double (&lk)[K], const int32_t (&li)[K],
                                            double key, int32_t id) {
    bool f = false;
#pragma unroll
    for (int s = 0; s < K; ++s) f |= (lk[s] == key) & (li[s] == id);
    return f;
}

// Merge a query's n_probe pair lists, then certify: the query is finite and
// the smallest lower bound over cells it did not scan exceeds its k-th
// distance.  Thread per query.  Uncertified queries get flag 1.
//
// A rescan passes the earlier search's k best (prev_d, prev_i; index -1 for
// an empty entry) and the routes of the cells the earlier search scanned
// are set to C in scan_routes: those slots were not scanned again and are
// skipped here.  A (distance, index) pair already in the list is not
// inserted again, so a cell scanned twice cannot duplicate a row.
template <typename TQ, int K>
__global__ void rt_merge_certify_kernel(const double *__restrict__ pkey,
                                        const int32_t *__restrict__ pid, int64_t nq, int n_probe,
                                        int k, const int32_t *__restrict__ routes,
                                        const int32_t *__restrict__ scan_routes, int C,
                                        const double *__restrict__ prev_d,
                                        const int64_t *__restrict__ prev_i,
                                        const double *__restrict__ bkey,
                                        const int32_t *__restrict__ bid, int kb,
                                        const TQ *__restrict__ Q, int d,
                                        const int32_t *__restrict__ cell_count,
                                        double *__restrict__ dist, int64_t *__restrict__ idx,
                                        uint8_t *__restrict__ flag,
                                        unsigned long long *__restrict__ rows_scanned) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double lk[K];
    int32_t li[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        lk[s] = __builtin_huge_val();
        li[s] = RT_NOID;
    }
    if (prev_d)
        for (int j = 0; j < k; ++j) {
            const int64_t pi = prev_i[q * k + j];
            if (pi < 0) break;
            const double key = prev_d[q * k + j];
            if (!pair_less(key, (int32_t)pi, lk[K - 1], li[K - 1])) break;
            rt_insert<K>(lk, li, key, (int32_t)pi);
        }
    unsigned long long scanned = 0;
    for (int s = 0; s < n_probe; ++s) {
        const int64_t p = q * n_probe + s;
        const int32_t sc = scan_routes[p];
        if (sc == C) continue;
        scanned += (unsigned long long)cell_count[sc];
        for (int j = 0; j < k; ++j) {
            const double key = pkey[p * k + j];
            const int32_t id = pid[p * k + j];
            if (!pair_less(key, id, lk[K - 1], li[K - 1])) break;
            if (prev_d && rt_contains<K>(lk, li, key, id)) continue;
            rt_insert<K>(lk, li, key, id);
        }
    }
    double dk = __builtin_huge_val();
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (j < k) {
            dist[q * k + j] = lk[j];
            idx[q * k + j] = li[j] == RT_NOID ? (int64_t)-1 : (int64_t)li[j];
            if (j == k - 1) dk = lk[j];
        }
    // smallest lower bound over unscanned cells: the first entry of the
    // (n_probe + 1)-smallest bound list that is not a routed cell
    double lbx = __builtin_huge_val();
    for (int j = 0; j < kb; ++j) {
        const int32_t c = bid[q * kb + j];
        bool routed = false;
        for (int s = 0; s < n_probe; ++s) routed |= routes[q * n_probe + s] == c;
        if (!routed) {
            lbx = c == RT_NOID ? __builtin_huge_val() : bkey[q * kb + j];
            break;
        }
    }
    bool finite = true;
    for (int j = 0; j < d; ++j) {
        const double v = (double)Q[q * d + j];
        finite &= v - v == 0.0;
    }
    flag[q] = (finite && lbx > dk) ? 0 : 1;
    atomicAdd(rows_scanned, scanned);
}

// Cell radii: max over a cell's rows of fold(x, mu) as ordered f64 bits
// (non-negative doubles order like their bit patterns).
template <typename TX>
__global__ void rt_radius_kernel(const TX *__restrict__ Xs, int64_t n, int d,
                                 const int32_t *__restrict__ row_cell,
                                 const double *__restrict__ centers,
                                 unsigned long long *__restrict__ maxfold) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = row_cell[i];
    const double f = fold_exact(Xs + i * d, centers + (int64_t)c * d, d);
    atomicMax(maxfold + c, (unsigned long long)__double_as_longlong(f));
}

// radius[c] = sqrt(max fold), or -1 for an empty cell.
__global__ void rt_radius_finish_kernel(const unsigned long long *__restrict__ maxfold,
                                        const int32_t *__restrict__ count, int C,
                                        double *__restrict__ radius) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    radius[c] = count[c] > 0 ? sqrt(__longlong_as_double((long long)maxfold[c])) : -1.0;
}

// Rows in cell order: Xs[pos] = X[order[pos]], and the cell of each sorted
// row.
template <typename TX>
__global__ void rt_gather_sorted_kernel(const TX *__restrict__ X, int64_t n, int d,
                                        const int32_t *__restrict__ order,
                                        const int32_t *__restrict__ label,
                                        TX *__restrict__ Xs, int32_t *__restrict__ row_cell) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * d) return;
    const int64_t pos = t / d;
    const int j = (int)(t % d);
    const int32_t r = order[pos];
    Xs[t] = X[(int64_t)r * d + j];
    if (j == 0) row_cell[pos] = label[r];
}

// int64 labels (KMeans output) -> int32.
__global__ void rt_labels32_kernel(const int64_t *__restrict__ in, int64_t n,
                                   int32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)in[i];
}

// Training targets: the cell of each of a query's k exact neighbours.
__global__ void rt_target_cells_kernel(const int64_t *__restrict__ nbr, int64_t count,
                                       const int32_t *__restrict__ label,
                                       int32_t *__restrict__ cells) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) cells[i] = nbr[i] >= 0 ? label[nbr[i]] : -1;
}

// Work items of the scan: per cell, ceil(pairs / 64) blocks, the cells taken
// in `order` (largest first, so that their items start first).  One
// 256-thread block scans the per-cell block counts; total[0] = number of
// items.
__global__ __launch_bounds__(256) void rt_work_scan_kernel(const int32_t *__restrict__ pair_count,
                                                           const int32_t *__restrict__ order,
                                                           int C, int32_t *__restrict__ first,
                                                           int32_t *__restrict__ total) {
    __shared__ int32_t part[256];
    const int t = threadIdx.x;
    const int per = (C + 255) / 256;
    const int lo = t * per, hi = lo + per < C ? lo + per : C;
    int32_t s = 0;
    for (int i = lo; i < hi; ++i) s += (pair_count[order[i]] + RT_Q - 1) / RT_Q;
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = t > 0 ? part[t - 1] : 0;
    for (int i = lo; i < hi; ++i) {
        const int c = order[i];
        first[c] = run;
        run += (pair_count[c] + RT_Q - 1) / RT_Q;
    }
    if (t == 255) total[0] = part[255];
}

__global__ void rt_work_fill_kernel(const int32_t *__restrict__ pair_count,
                                    const int32_t *__restrict__ first, int C,
                                    int32_t *__restrict__ work_cell,
                                    int32_t *__restrict__ work_start) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int nb = (pair_count[c] + RT_Q - 1) / RT_Q;
    for (int j = 0; j < nb; ++j) {
        work_cell[first[c] + j] = c;
        work_start[first[c] + j] = j * RT_Q;
    }
}

// Rows of the queries that failed their certificate, for the brute force.
template <typename TQ>
__global__ void rt_gather_queries_kernel(const TQ *__restrict__ Q, int d,
                                         const int32_t *__restrict__ which, int64_t m,
                                         TQ *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * d) return;
    out[t] = Q[(int64_t)which[t / d] * d + t % d];
}

__global__ void rt_scatter_results_kernel(const double *__restrict__ fd,
                                          const int64_t *__restrict__ fi,
                                          const int32_t *__restrict__ which, int64_t m, int k,
                                          double *__restrict__ dist, int64_t *__restrict__ idx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * k) return;
    const int64_t q = which[t / k];
    dist[q * k + t % k] = fd[t];
    idx[q * k + t % k] = fi[t];
}

// A rescan's inputs from the search before it: the k best of each
// uncertified query and the cells that search scanned.
__global__ void rt_gather_prev_kernel(const double *__restrict__ dist,
                                      const int64_t *__restrict__ idx,
                                      const int32_t *__restrict__ routes, int n_probe,
                                      const int32_t *__restrict__ which, int64_t m, int k,
                                      double *__restrict__ pd, int64_t *__restrict__ pi,
                                      int32_t *__restrict__ pr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m * k) {
        const int64_t q = which[t / k];
        pd[t] = dist[q * k + t % k];
        pi[t] = idx[q * k + t % k];
    }
    if (t < m * n_probe) pr[t] = routes[(int64_t)which[t / n_probe] * n_probe + t % n_probe];
}

// scan_routes = routes, with C in place of every cell that need not be
// scanned (the pair goes to bucket C, which gets no work items): a cell the
// earlier search already scanned for that query, and, when the routes are
// the smallest bounds in order (bkey non-null), a cell whose bound exceeds
// the earlier search's k-th distance.  That distance only decreases as rows
// are added, so such a cell's rows cannot enter the answer and the cell
// passes the certificate unscanned.
__global__ void rt_skip_routes_kernel(const int32_t *__restrict__ routes, int64_t nq,
                                      int n_probe, const int32_t *__restrict__ prev_routes,
                                      int prev_np, int C, const double *__restrict__ bkey,
                                      int kb, const double *__restrict__ prev_d, int k,
                                      int32_t *__restrict__ scan_routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * n_probe) return;
    const int64_t q = t / n_probe;
    const int32_t c = routes[t];
    bool skip = false;
    for (int s = 0; s < prev_np; ++s) skip |= prev_routes[q * prev_np + s] == c;
    if (bkey) skip |= bkey[q * kb + t % n_probe] > prev_d[q * k + k - 1];
    scan_routes[t] = skip ? C : c;
}

// The all-cells rescan's scan_routes (routes[q][c] = c): C in place of a
// cell the earlier search scanned, an empty cell, or a cell whose ball bound
// exceeds the earlier search's k-th distance (as in rt_skip_routes_kernel).
template <typename TQ>
__global__ void rt_skip_all_kernel(const TQ *__restrict__ Q, int d, int64_t nq, int C,
                                   const double *__restrict__ centers,
                                   const double *__restrict__ radius, double ep,
                                   const int32_t *__restrict__ prev_routes, int prev_np,
                                   const double *__restrict__ prev_d, int k,
                                   int32_t *__restrict__ scan_routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * C) return;
    const int64_t q = t / C;
    const int32_t c = (int32_t)(t % C);
    bool skip = radius[c] < 0.0;
    for (int s = 0; s < prev_np; ++s) skip |= prev_routes[q * prev_np + s] == c;
    if (!skip) {
        const double f = fold_exact(Q + q * d, centers + (int64_t)c * d, d);
        skip = rt_lower_bound(sqrt(f), radius[c], ep) > prev_d[q * k + k - 1];
    }
    scan_routes[t] = skip ? C : c;
}

// First pass of a search over n_probe >= 2 cells: slot 0 of each query.
__global__ void rt_first_slot_kernel(const int32_t *__restrict__ routes, int64_t P, int n_probe,
                                     int C, int32_t *__restrict__ scan_routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    scan_routes[t] = t % n_probe == 0 ? routes[t] : C;
}

// Second pass: slots 1.. whose cell's key (the larger of the ball and
// bisector bounds, as in rt_bounds_kernel) does not exceed the k-th distance
// of the first pass; C elsewhere.  The skipped cells' rows cannot enter the
// answer (that distance only decreases) and the cells pass the certificate
// unscanned.
template <typename TQ>
__global__ void rt_skip_by_bound_kernel(const TQ *__restrict__ Q, int d, int64_t nq, int n_probe,
                                        const int32_t *__restrict__ routes, int C,
                                        const double *__restrict__ centers,
                                        const double *__restrict__ radius,
                                        const double *__restrict__ cdist,
                                        const double *__restrict__ near_fold,
                                        const int32_t *__restrict__ near_cell, double ep,
                                        const double *__restrict__ prev_d, int k,
                                        int32_t *__restrict__ scan_routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * n_probe) return;
    const int64_t q = t / n_probe;
    const int32_t c = routes[t];
    bool skip = t % n_probe == 0 || radius[c] < 0.0;
    if (!skip) {
        const double fold = fold_exact(Q + q * d, centers + (int64_t)c * d, d);
        double key = rt_lower_bound(sqrt(fold), radius[c], ep);
        const int32_t a = near_cell[q];
        const double D = cdist[(int64_t)a * C + c];
        if (c != a && D > 0.0) {
            const double v = rt_bisector_bound(fold, near_fold[q], D, radius[c], ep);
            key = v > key ? v : key;
        }
        skip = key > prev_d[q * k + k - 1];
    }
    scan_routes[t] = skip ? C : c;
}

// Selected rows widened to f64 (initial cell centres).
template <typename TX>
__global__ void rt_rows_f64_kernel(const TX *__restrict__ X, int d, const int32_t *__restrict__ rows,
                                   int64_t m, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * d) return;
    out[t] = (double)X[(int64_t)rows[t / d] * d + t % d];
}

// Bound-ordered routing: the n_probe cells with the smallest lower bounds.
__global__ void rt_routes_from_bounds_kernel(const int32_t *__restrict__ bid, int kb, int64_t nq,
                                             int n_probe, int32_t *__restrict__ routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * n_probe) return;
    // every cell enters the bound list (NaN bounds become -inf), so with
    // n_cells >= n_probe the first n_probe entries are cells; the clamp
    // only keeps the counting sort's labels in range
    const int32_t c = bid[(t / n_probe) * kb + t % n_probe];
    routes[t] = c == RT_NOID ? 0 : c;
}

// Centre distances D[a][c] = sqrt(fold(mu_a, mu_c)).
__global__ void rt_center_dist_kernel(const double *__restrict__ centers, int C, int d,
                                      double *__restrict__ D) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)C * C) return;
    const int64_t a = t / C, c = t % C;
    D[t] = sqrt(fold_exact(centers + a * d, centers + c * d, d));
}

// Certification targets: the cells whose lower bound does not exceed the
// query's exact k-th distance (from the T smallest bounds; -1 pads).  A
// search that scans exactly these cells finds the exact answer and
// certifies it; the cells of the k neighbours are among them.
__global__ void rt_certify_targets_kernel(const double *__restrict__ kth, int64_t nq,
                                          const double *__restrict__ bkey,
                                          const int32_t *__restrict__ bid, int T,
                                          int32_t *__restrict__ cells) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * T) return;
    const int64_t q = t / T;
    const int32_t c = bid[t];
    cells[t] = (c != RT_NOID && bkey[t] <= kth[q]) ? c : -1;
}

// k-th distance of each query from its sorted kNN distances.
__global__ void rt_kth_kernel(const double *__restrict__ dist, int64_t nq, int k,
                              double *__restrict__ kth) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) kth[q] = dist[q * k + k - 1];
}

// Every cell, in order, for each query.
__global__ void rt_routes_all_kernel(int64_t nq, int C, int32_t *__restrict__ routes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nq * C) routes[t] = (int32_t)(t % C);
}

}  // namespace fl
