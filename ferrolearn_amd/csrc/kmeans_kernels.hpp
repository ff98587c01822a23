// kmeans_kernels.hpp -- KMeans assignment decision, inertia, reference-order
// centroid update and k-means++ passes on gfx950.
//
// Reference: ferrolearn-cluster/src/kmeans.rs.  Every decision uses the
// exact fold (kmeans.rs:159-163) and the strict `<` argmin that keeps the
// lowest centre index on ties (kmeans.rs:295-309).
#pragma once

#include "common.hpp"
#include "exact_fold.hpp"
#include "dist_topm.hpp"

namespace fl {

// -------------------------------------------------------------- decision --
// After the filter kernel: fold the kept candidates exactly, pick the
// smallest (fold, index), and certify it: every centre that was not kept has
// key >= theta and key <= fold, so theta > best fold proves that no other
// centre has a smaller fold or an equal fold with a lower index.
template <typename TX>
__global__ void km_finalize_kernel(const TX *__restrict__ X, const double *__restrict__ C,
                                   int64_t n, int d, int64_t k, int M, Cand cand,
                                   int32_t *__restrict__ labels, double *__restrict__ best,
                                   int32_t *__restrict__ flag_list,
                                   int32_t *__restrict__ flag_count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cand.cnt[i];
    const TX *x = X + i * d;
    // Fold the candidate with the smallest key first, then only candidates
    // whose key (a lower bound of their fold) does not exceed the best fold.
    int first = 0;
    for (int m = 1; m < c; ++m)
        if (pair_less(cand.key[i * M + m], cand.idx[i * M + m], cand.key[i * M + first],
                      cand.idx[i * M + first]))
            first = m;
    double bf = 0.0;
    int32_t bl = -1;
    for (int r = 0; r < c; ++r) {
        const int m = r == 0 ? first : (r <= first ? r - 1 : r);
        const int32_t ix = cand.idx[i * M + m];
        if (bl >= 0 && !(cand.key[i * M + m] <= bf)) continue;
        const double f = fold_exact(x, C + (int64_t)ix * d, d);
        if (bl < 0 || pair_less(f, ix, bf, bl)) {
            bf = f;
            bl = ix;
        }
    }
    bool ok = bl >= 0 && (c == M || (int64_t)c == k);
    if (ok && c == M) ok = cand.theta[i] > bf;
    if (!ok) {
        flag_list[atomicAdd(flag_count, 1)] = (int32_t)i;
        return;
    }
    labels[i] = bl;
    best[i] = bf;
}

// Exact scan over every centre for one row: kmeans.rs:295-309 verbatim
// (init (0, F::MAX), update iff fold < best).
template <typename TX>
__device__ __forceinline__ void km_exact_row(const TX *__restrict__ x,
                                             const double *__restrict__ C, int d, int64_t k,
                                             int32_t &lab, double &dist) {
    int32_t bl = 0;
    double bd = 1.7976931348623157e308;
    for (int64_t c = 0; c < k; ++c) {
        const double f = fold_exact(x, C + c * d, d);
        if (f < bd) {
            bd = f;
            bl = (int32_t)c;
        }
    }
    lab = bl;
    dist = bd;
}

// Direct exact assignment of every row (used when k*d is small).
template <typename TX>
__global__ void km_assign_exact_kernel(const TX *__restrict__ X, const double *__restrict__ C,
                                       int64_t n, int d, int64_t k,
                                       int32_t *__restrict__ labels,
                                       double *__restrict__ best) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t l;
    double v;
    km_exact_row(X + i * d, C, d, k, l, v);
    labels[i] = l;
    best[i] = v;
}

// Exact scan of every centre for one row by one wave: lane l folds centres
// l, l+64, ...; the wave then takes the smallest (fold, index), which is the
// result of kmeans.rs:295-309's sequential scan (init (0, F::MAX), update iff
// fold < best): the lowest index among the smallest folds below F::MAX, else
// (0, F::MAX).  Every lane returns the result.
template <typename TX>
__device__ __forceinline__ void wave_exact_argmin(const TX *__restrict__ x,
                                                  const double *__restrict__ C, int d, int64_t k,
                                                  int lane, int32_t &lab, double &dist) {
    double bd = 1.7976931348623157e308;
    int64_t bl = -1;
    for (int64_t c = lane; c < k; c += 64) {
        const double f = fold_exact(x, C + c * d, d);
        if (f < bd) {
            bd = f;
            bl = c;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double od = __shfl_xor(bd, o);
        const int64_t ol = __shfl_xor(bl, o);
        if (ol >= 0 && (bl < 0 || od < bd || (od == bd && ol < bl))) {
            bd = od;
            bl = ol;
        }
    }
    lab = bl < 0 ? 0 : (int32_t)bl;
    dist = bd;
}

// Exact scan for the rows the certificate could not settle: one wave per
// listed row, four per 256-thread block, the grid striding over the count
// the filter left on the device (so the host need not read it first).
template <typename TX>
__global__ __launch_bounds__(256) void km_fallback_kernel(
    const TX *__restrict__ X, const double *__restrict__ C, int d, int64_t k,
    const int32_t *__restrict__ list, const int32_t *__restrict__ count,
    int32_t *__restrict__ labels, double *__restrict__ best) {
    const int lane = threadIdx.x & 63;
    const int64_t m = *count;
    for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < m;
         t += (int64_t)gridDim.x * 4) {
        const int64_t i = list[t];
        int32_t l;
        double v;
        wave_exact_argmin(X + i * d, C, d, k, lane, l, v);
        if (lane == 0) {
            labels[i] = l;
            best[i] = v;
        }
    }
}

// Feature-major copy of the centres: Ct[j][c] = C[c][j].
static __global__ __attribute__((unused)) void km_transpose_kernel(const double *__restrict__ C, int64_t k, int d,
                                    double *__restrict__ Ct) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * d) return;
    const int64_t j = t / k, c = t % k;
    Ct[t] = C[c * d + j];
}

// Same scan, one 256-thread block per listed row: the row is staged in LDS
// as f64 (dynamic LDS: d doubles), the centres are read from a
// feature-major copy Ct (d x k, so adjacent lanes read adjacent centres),
// and thread t folds centres t, t + 256, ... two at a time (two independent
// left folds).  Each thread visits its centres in increasing order and
// replaces its best only on a strictly smaller fold; the waves' and then
// the block's (fold, index) minima are combined with the lowest index
// winning ties, as in wave_exact_argmin.  The grid strides over the count.
template <typename TX>
__global__ __launch_bounds__(256) void km_fallback_lds_kernel(
    const TX *__restrict__ X, const double *__restrict__ Ct, int d, int64_t k,
    const int32_t *__restrict__ list, const int32_t *__restrict__ count,
    int32_t *__restrict__ labels, double *__restrict__ best) {
    extern __shared__ double xr[];
    __shared__ double wd[4];
    __shared__ int64_t wl[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t m = *count;
    for (int64_t e = blockIdx.x; e < m; e += gridDim.x) {
        const int64_t i = list[e];
        for (int j = t; j < d; j += 256) xr[j] = (double)X[i * d + j];
        __syncthreads();
        double bd = 1.7976931348623157e308;
        int64_t bl = -1;
        int64_t c = t;
        for (; c + 256 < k; c += 512) {
            const double *c0 = Ct + c, *c1 = Ct + c + 256;
            double a0 = 0.0, a1 = 0.0;
#pragma unroll 8
            for (int j = 0; j < d; ++j) {
                const double x = xr[j];
                const double t0 = x - c0[(int64_t)j * k];
                const double t1 = x - c1[(int64_t)j * k];
                a0 = a0 + t0 * t0;
                a1 = a1 + t1 * t1;
            }
            if (a0 < bd) {
                bd = a0;
                bl = c;
            }
            if (a1 < bd) {
                bd = a1;
                bl = c + 256;
            }
        }
        for (; c < k; c += 256) {
            const double *c0 = Ct + c;
            double a0 = 0.0;
            for (int j = 0; j < d; ++j) {
                const double t0 = xr[j] - c0[(int64_t)j * k];
                a0 = a0 + t0 * t0;
            }
            if (a0 < bd) {
                bd = a0;
                bl = c;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double od = __shfl_xor(bd, o);
            const int64_t ol = __shfl_xor(bl, o);
            if (ol >= 0 && (bl < 0 || od < bd || (od == bd && ol < bl))) {
                bd = od;
                bl = ol;
            }
        }
        if (lane == 0) {
            wd[w] = bd;
            wl[w] = bl;
        }
        __syncthreads();
        if (t == 0) {
            for (int v = 1; v < 4; ++v) {
                const double od = wd[v];
                const int64_t ol = wl[v];
                if (ol >= 0 && (bl < 0 || od < bd || (od == bd && ol < bl))) {
                    bd = od;
                    bl = ol;
                }
            }
            labels[i] = bl < 0 ? 0 : (int32_t)bl;
            best[i] = bd;
        }
        __syncthreads();
    }
}

// The same scan for R listed rows per block (R * d doubles staged in LDS):
// each centre feature read from memory serves the R rows, and the R * 2
// folds per thread are independent chains.  Per (row, centre) the fold, the
// visiting order and the tie rule are those of km_fallback_lds_kernel.
// Slots past the count take the first listed row of the group and are not
// written.  The grid strides over the count in groups of R.
template <typename TX, int R>
__global__ __launch_bounds__(256) void km_fallback_rows_kernel(
    const TX *__restrict__ X, const double *__restrict__ Ct, int d, int64_t k,
    const int32_t *__restrict__ list, const int32_t *__restrict__ count,
    int32_t *__restrict__ labels, double *__restrict__ best) {
    extern __shared__ double xr[];   // R rows of d
    __shared__ double wd[R][4];
    __shared__ int64_t wl[R][4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t m = *count;
    for (int64_t e0 = (int64_t)blockIdx.x * R; e0 < m; e0 += (int64_t)gridDim.x * R) {
        int64_t rows[R];
#pragma unroll
        for (int r = 0; r < R; ++r) rows[r] = list[e0 + r < m ? e0 + r : e0];
#pragma unroll
        for (int r = 0; r < R; ++r)
            for (int j = t; j < d; j += 256) xr[r * d + j] = (double)X[rows[r] * d + j];
        __syncthreads();
        double bd[R];
        int64_t bl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bd[r] = 1.7976931348623157e308;
            bl[r] = -1;
        }
        int64_t c = t;
        for (; c + 256 < k; c += 512) {
            const double *c0 = Ct + c, *c1 = Ct + c + 256;
            double a0[R], a1[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                a0[r] = 0.0;
                a1[r] = 0.0;
            }
            for (int j = 0; j < d; ++j) {
                const double y0 = c0[(int64_t)j * k], y1 = c1[(int64_t)j * k];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double x = xr[r * d + j];
                    const double t0 = x - y0;
                    const double t1 = x - y1;
                    a0[r] = a0[r] + t0 * t0;
                    a1[r] = a1[r] + t1 * t1;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (a0[r] < bd[r]) {
                    bd[r] = a0[r];
                    bl[r] = c;
                }
                if (a1[r] < bd[r]) {
                    bd[r] = a1[r];
                    bl[r] = c + 256;
                }
            }
        }
        for (; c < k; c += 256) {
            const double *c0 = Ct + c;
            double a0[R];
#pragma unroll
            for (int r = 0; r < R; ++r) a0[r] = 0.0;
            for (int j = 0; j < d; ++j) {
                const double y0 = c0[(int64_t)j * k];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double t0 = xr[r * d + j] - y0;
                    a0[r] = a0[r] + t0 * t0;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (a0[r] < bd[r]) {
                    bd[r] = a0[r];
                    bl[r] = c;
                }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            for (int o = 32; o > 0; o >>= 1) {
                const double od = __shfl_xor(bd[r], o);
                const int64_t ol = __shfl_xor(bl[r], o);
                if (ol >= 0 && (bl[r] < 0 || od < bd[r] || (od == bd[r] && ol < bl[r]))) {
                    bd[r] = od;
                    bl[r] = ol;
                }
            }
            if (lane == 0) {
                wd[r][w] = bd[r];
                wl[r][w] = bl[r];
            }
        }
        __syncthreads();
        if (t < R && e0 + t < m) {
            const int r = t;
            double qd = wd[r][0];
            int64_t ql = wl[r][0];
            for (int v = 1; v < 4; ++v) {
                const double od = wd[r][v];
                const int64_t ol = wl[r][v];
                if (ol >= 0 && (ql < 0 || od < qd || (od == qd && ol < ql))) {
                    qd = od;
                    ql = ol;
                }
            }
            const int64_t i = list[e0 + r];
            labels[i] = ql < 0 ? 0 : (int32_t)ql;
            best[i] = qd;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- mini-batch --
// mini_batch_kmeans.rs:381-404 (assign_batch): nearest centre of each batch
// row, the same sequential scan; one wave per row.
template <typename TX>
__global__ __launch_bounds__(256) void mb_assign_batch_kernel(
    const TX *__restrict__ X, const double *__restrict__ C, int d, int64_t k,
    const int32_t *__restrict__ bidx, int64_t b, int32_t *__restrict__ blab) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= b) return;
    int32_t l;
    double v;
    wave_exact_argmin(X + (int64_t)bidx[t] * d, C, d, k, lane, l, v);
    if (lane == 0) blab[t] = l;
}

// mini_batch_kmeans.rs:410-426 (update_centers_mini_batch): for the batch
// rows in batch order, counts[l] += 1, lr = 1 / counts[l], and
// c[l][j] = c[l][j] + lr * (x[j] - c[l][j]).  One 64-thread block per centre:
// it scans the batch labels in order and updates its features in place
// (thread t owns features t, t+64, ...); different centres are independent.
template <typename TX>
__global__ __launch_bounds__(64) void mb_update_kernel(const TX *__restrict__ X, int d,
                                                       const int32_t *__restrict__ bidx,
                                                       const int32_t *__restrict__ blab, int64_t b,
                                                       double *__restrict__ C,
                                                       int64_t *__restrict__ counts) {
    __shared__ int32_t labs[1024];
    const int64_t c = blockIdx.x;
    const int t = threadIdx.x;
    double *cr = C + c * d;
    int64_t cnt = counts[c];
    for (int64_t s0 = 0; s0 < b; s0 += 1024) {
        const int nb = b - s0 < 1024 ? (int)(b - s0) : 1024;
        wave_sync_lds();
        for (int u = t; u < nb; u += 64) labs[u] = blab[s0 + u];
        wave_sync_lds();
        for (int s = 0; s < nb; ++s) {
            if (labs[s] != (int32_t)c) continue;
            cnt += 1;
            const double lr = 1.0 / (double)cnt;
            const TX *x = X + (int64_t)bidx[s0 + s] * d;
            for (int j = t; j < d; j += 64) cr[j] = cr[j] + lr * ((double)x[j] - cr[j]);
        }
    }
    if (t == 0) counts[c] = cnt;
}

static __global__ __attribute__((unused)) void fill_i64_kernel(int64_t *__restrict__ v, int64_t n,
                                                                int64_t value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

// mini_batch_kmeans.rs:366-373: inertia as the left fold of the rows' best
// folds in row order (one lane; loads in batches of 32).
static __global__ __attribute__((unused)) void seq_fold_kernel(const double *__restrict__ v,
                                                                int64_t n, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    int64_t i = 0;
    for (; i + 32 <= n; i += 32) {
        double b[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) b[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < 32; ++u) s = s + b[u];
    }
    for (; i < n; ++i) s = s + v[i];
    *out = s;
}

// ---------------------------------------------------------------- inertia --
// Deterministic two-level sum of the per-row best folds: 1024-row blocks are
// reduced by a fixed tree, then the block sums in block order.  The
// reference's own parallel inertia is order dependent (kmeans.rs:326-352);
// both are within a few ulps of the exact sum.
constexpr int SUM_BLK = 1024;

static __global__ __attribute__((unused)) __launch_bounds__(256) void block_sum_kernel(const double *__restrict__ v, int64_t n,
                                                        double *__restrict__ part) {
    __shared__ double s[256];
    const int64_t b0 = (int64_t)blockIdx.x * SUM_BLK;
    double acc = 0.0;
    for (int r = 0; r < SUM_BLK / 256; ++r) {
        const int64_t i = b0 + r * 256 + threadIdx.x;
        if (i < n) acc = acc + v[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Sequential fold of m values (one lane): the second level of k-means++'s
// two-level sums, which the oracle restates in the same order.
static __global__ __attribute__((unused)) void seq_sum_kernel(const double *__restrict__ part, int64_t m,
                               double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = 0.0;
    for (int64_t i = 0; i < m; ++i) acc = acc + part[i];
    *out = acc;
}

// Second level of the inertia: lane l folds the l-th of 64 contiguous runs
// of block sums, then lane 0 folds the 64 run sums in order (a fixed order;
// launched with one wave).
static __global__ __attribute__((unused)) void runs_sum_kernel(const double *__restrict__ part,
                                                                int64_t m, double *__restrict__ out) {
    __shared__ double stage[8192];
    const int lane = threadIdx.x;
    if (blockIdx.x != 0 || lane >= 64) return;
    const int64_t per = (m + 63) / 64;
    const int64_t lo = lane * per, hi = lo + per < m ? lo + per : m;
    double acc = 0.0;
    // the partials pass through LDS 8192 at a time (coalesced loads); each
    // lane folds the part of its run inside the staged window, in order
    for (int64_t s0 = 0; s0 < m; s0 += 8192) {
        const int64_t s1 = s0 + 8192 < m ? s0 + 8192 : m;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (int64_t i = s0 + lane; i < s1; i += 64) stage[i - s0] = part[i];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const int64_t a = lo > s0 ? lo : s0, b = hi < s1 ? hi : s1;
        for (int64_t i = a; i < b; ++i) acc = acc + stage[i - s0];
    }
    double tot = 0.0;
    for (int l = 0; l < 64; ++l) tot = tot + wave_bcast(acc, l);
    if (lane == 0) *out = tot;
}

// ---------------------------------------------------------- centroid update --
// kmeans.rs:358-401 sums x[i][j] into sum[label][j] for i ascending.  The
// GPU reproduces that order exactly: a stable counting sort groups the rows
// of each cluster in ascending index order, then one lane per (cluster,
// feature) folds its cluster's rows left to right from +0.
constexpr int UPD_RB = 4096;   // rows per counting block

// Bits that distinguish labels 0..k-1.
inline int label_bits(int64_t k) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << b) < k) ++b;
    return b;
}

static __global__ __attribute__((unused)) __launch_bounds__(256) void km_block_hist_kernel(const int32_t *__restrict__ labels,
                                                            int64_t n, int64_t k,
                                                            int32_t *__restrict__ blockcnt) {
    extern __shared__ int32_t hist[];
    for (int64_t c = threadIdx.x; c < k; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * UPD_RB;
    for (int r = threadIdx.x; r < UPD_RB; r += blockDim.x) {
        const int64_t i = b0 + r;
        if (i < n) atomicAdd(&hist[labels[i]], 1);
    }
    __syncthreads();
    for (int64_t c = threadIdx.x; c < k; c += blockDim.x)
        blockcnt[(int64_t)blockIdx.x * k + c] = hist[c];
}

// Per cluster: running position of each block's first member, and the
// cluster's total count.  One 256-thread block per cluster: each thread sums
// a contiguous run of counting blocks, an LDS scan orders the runs.
static __global__ __attribute__((unused)) __launch_bounds__(256) void km_col_scan_kernel(
    const int32_t *__restrict__ blockcnt, int64_t nblk, int64_t k, int32_t *__restrict__ blockbase,
    int32_t *__restrict__ count) {
    __shared__ int32_t part[256];
    const int64_t c = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t per = (nblk + 255) / 256;
    const int64_t lo = t * per, hi = lo + per < nblk ? lo + per : nblk;
    int32_t s = 0;
    for (int64_t b = lo; b < hi; ++b) s += blockcnt[b * k + c];
    part[t] = s;
    __syncthreads();
    // exclusive scan of part[] (Hillis-Steele in LDS)
    for (int o = 1; o < 256; o <<= 1) {
        const int32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t b = lo; b < hi; ++b) {
        blockbase[b * k + c] = run;
        run += blockcnt[b * k + c];
    }
    if (t == 255) count[c] = part[255];
}

// Exclusive scan of the cluster counts (one 256-thread block).
// v[0..m) = value.
static __global__ __attribute__((unused)) void km_fill_i32_kernel(int32_t *__restrict__ v,
                                                                   int64_t m, int32_t value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] = value;
}

// Members per cluster by global atomics (counts are exact in any order) and
// the row indices 0..n-1 to be sorted by label.
static __global__ __attribute__((unused)) void km_count_ids_kernel(
    const int32_t *__restrict__ labels, int64_t n, int32_t *__restrict__ count,
    int32_t *__restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&count[labels[i]], 1);
    ids[i] = (int32_t)i;
}

static __global__ __attribute__((unused)) __launch_bounds__(256) void km_offsets_kernel(
    const int32_t *__restrict__ count, int64_t k, int64_t *__restrict__ offset) {
    __shared__ int64_t part[256];
    const int t = threadIdx.x;
    const int64_t per = (k + 255) / 256;
    const int64_t lo = t * per, hi = lo + per < k ? lo + per : k;
    int64_t s = 0;
    for (int64_t c = lo; c < hi; ++c) s += count[c];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t c = lo; c < hi; ++c) {
        offset[c] = run;
        run += count[c];
    }
}

// One wave per counting block: the lanes of a 64-row chunk that share a
// label are found with one ballot per label bit (their mask is the AND of
// the matching bit masks), rank = lanes before with that label, so the
// scatter is stable; the first lane of each group reserves the group's
// slots with one LDS atomic, and the chunks are taken in row order.
static __global__ __attribute__((unused)) __launch_bounds__(64) void km_scatter_kernel(
    const int32_t *__restrict__ labels, int64_t n, int64_t k, int bits,
    const int32_t *__restrict__ blockbase, const int64_t *__restrict__ offset,
    int32_t *__restrict__ order) {
    extern __shared__ int64_t ctr[];
    const int lane = threadIdx.x;
    const int64_t b = blockIdx.x;
    for (int64_t c = lane; c < k; c += 64) ctr[c] = offset[c] + blockbase[b * k + c];
    wave_sync_lds();
    const int64_t b0 = b * UPD_RB;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int r0 = 0; r0 < UPD_RB; r0 += 64) {
        const int64_t i = b0 + r0 + lane;
        const bool valid = i < n;
        const int32_t lab = valid ? labels[i] : 0;
        uint64_t m = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const bool on = (lab >> bit) & 1;
            const uint64_t bb = __ballot(valid && on);
            m &= on ? bb : ~bb;
        }
        const int rank = __popcll(m & below);
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned long long base = 0;
        if (valid && leader == lane)
            base = atomicAdd((unsigned long long *)&ctr[lab], (unsigned long long)__popcll(m));
        const unsigned long long gb = __shfl(base, valid ? leader : lane);
        if (valid) order[(int64_t)gb + rank] = (int32_t)i;
    }
}

// sum[c][j] = fold over the cluster's rows in ascending index order; then
// divide by the count or keep the old centre for an empty cluster.  The
// member indices are staged through LDS 512 at a time and the values loaded
// 64 at a time, so that the loads of a batch overlap.
template <typename TX>
__global__ __launch_bounds__(64) void km_sum_kernel(const TX *__restrict__ X, int d,
                                                    const int32_t *__restrict__ order,
                                                    const int64_t *__restrict__ offset,
                                                    const int32_t *__restrict__ count,
                                                    const double *__restrict__ oldc,
                                                    double *__restrict__ newc) {
    __shared__ int32_t ord[512];
    const int64_t c = blockIdx.x;
    const int j = blockIdx.y * 64 + threadIdx.x;
    const bool col = j < d;
    const int64_t off = offset[c];
    const int32_t cnt = count[c];
    double acc = 0.0;
    for (int32_t m0 = 0; m0 < cnt; m0 += 512) {
        const int nb = cnt - m0 < 512 ? cnt - m0 : 512;
        wave_sync_lds();
        for (int u = threadIdx.x; u < nb; u += 64) ord[u] = order[off + m0 + u];
        wave_sync_lds();
        if (!col) continue;
        int m = 0;
        constexpr int U = 64;
        for (; m + U <= nb; m += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = (double)X[(int64_t)ord[m + u] * d + j];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = acc + v[u];
        }
        for (; m < nb; ++m) acc = acc + (double)X[(int64_t)ord[m] * d + j];
    }
    if (col) newc[c * d + j] = cnt > 0 ? acc / (double)cnt : oldc[c * d + j];
}

// max_shift = sqrt(max_c fold(new_c, old_c)) (kmeans.rs:389-400); the max
// ignores NaN folds the way `if shift > max_shift` does.  One thread per
// centre; the maximum of the non-negative folds is taken on their bit
// patterns (zeroed before the launch), and km_shift_finish takes the sqrt.
static __global__ __attribute__((unused)) void km_shift_kernel(const double *__restrict__ newc,
                                                               const double *__restrict__ oldc,
                                                               int64_t k, int d,
                                                               unsigned long long *__restrict__ mx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    const double f = fold_exact(newc + c * d, oldc + c * d, d);
    if (f > 0.0) atomicMax(mx, (unsigned long long)__double_as_longlong(f));
}

static __global__ __attribute__((unused)) void km_shift_finish(
    const unsigned long long *__restrict__ mx, double *__restrict__ shift) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *shift = sqrt(__longlong_as_double((long long)*mx));
}

// --------------------------------------------------------------- transform --
template <typename TX>
__global__ void km_transform_kernel(const TX *__restrict__ X, const double *__restrict__ C,
                                    int64_t n, int d, int64_t k, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int64_t i = t / k, c = t % k;
    out[t] = sqrt(fold_exact(X + i * d, C + c * d, d));
}

// ---------------------------------------------------------------- k-means++ --
// The library's k-means++ (kmeans.rs:169-253 with two-level sums, see
// KPP_BLK): the distance passes run here, the generator on the host.
constexpr int KPP_BLK = 4096;

template <typename TX>
__global__ void kpp_first_kernel(const TX *__restrict__ X, int64_t n, int d,
                                 const double *__restrict__ c0, double *__restrict__ md) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) md[i] = fold_exact(X + i * d, c0, d);
}

// nd[i] = min(md[i], fold(x_i, cand)) with the reference's `if d < nd`.
template <typename TX>
__global__ void kpp_trial_kernel(const TX *__restrict__ X, int64_t n, int d, int64_t cand,
                                 const double *__restrict__ md, double *__restrict__ nd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = fold_exact(X + i * d, X + cand * d, d);
    const double m = md[i];
    nd[i] = v < m ? v : m;
}

// ------------------------------------------- k-means++ in the reference's order --
// kmeans.rs:169-253 with every sum the reference's plain left fold, each
// centre's steps enqueued without a host round trip.  The generator's draws
// for every centre are made on the host up front (they do not depend on the
// data while the reference's `total > 0` holds); the device keeps the
// running state:
//   total  = left fold of the current minimum distances md (the previous
//            centre's best potential, which is the same fold);
//   pre[i] = running left fold of md up to and including i, so a trial's
//            threshold search (the reference's cumsum loop) is the first i
//            with pre[i] >= threshold;
//   stop   = the first centre whose total is <= 0 (the reference then draws
//            a row uniformly instead; the host replays that tail).
struct KppSeqCtl {
    double total;
    int32_t stop;
    int32_t best;
};

// nd[t*n + i] = min(md[i], fold(x_i, x_cand[t])) with `if d < nd`
// (:219-226), every trial from one read of the row.
template <typename TX>
__global__ __launch_bounds__(256) void kpp_seq_trials_kernel(
    const KppSeqCtl *__restrict__ ctl, const TX *__restrict__ X, int64_t n, int d, int nt,
    const unsigned long long *__restrict__ cand, const double *__restrict__ md,
    double *__restrict__ nd) {
    if (ctl->stop) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = md[i];
    const TX *xi = X + i * d;
    for (int t0 = 0; t0 < nt; t0 += 4) {
        const int tc = nt - t0 < 4 ? nt - t0 : 4;
        const TX *c[4];
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = X + (int64_t)cand[t0 + (u < tc ? u : 0)] * d;
        for (int j = 0; j < d; ++j) {
            const double xv = (double)xi[j];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double df = xv - (double)c[u][j];
                acc[u] = acc[u] + df * df;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < tc) nd[(int64_t)(t0 + u) * n + i] = acc[u] < m ? acc[u] : m;
    }
}

// Centre c (>= 1), one launch: stop when the total is <= 0 (kmeans.rs:201;
// every block returns, block 0 records c), else cand[t] = first i with
// pre[i] >= r_t * total (:207-213).  The running sums of non-negative
// distances never decrease until they turn NaN, after which every comparison
// is false, so at most one i has pre[i] >= thr and not pre[i-1] >= thr: that
// thread writes it.  cand holds n - 1 beforehand (no index reaches it).  Grid:
// ceil(n / 256) x nt.
static __global__ __attribute__((unused)) __launch_bounds__(256) void kpp_seq_threshold_kernel(
    KppSeqCtl *__restrict__ ctl, const double *__restrict__ r, int c,
    const double *__restrict__ pre, int64_t n, unsigned long long *__restrict__ cand) {
    if (ctl->stop) return;
    const double total = ctl->total;
    if (total <= 0.0) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) ctl->stop = c;
        return;
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double th = r[blockIdx.y] * total;   // kmeans.rs:209-210
    if (pre[i] >= th && !(i > 0 && pre[i - 1] >= th)) cand[blockIdx.y] = (unsigned long long)i;
}

// The trial with the strictly smallest potential, first trial first
// (:231-235), found by every block; block 0 makes its candidate centre c,
// keeps its potential as the next total and sets the next centre's
// candidates to n - 1; every block adopts its distances and running sums
// (md = nd[best], pre = pre_t[best], :239-241).  Grid: ceil(n / 256).
template <typename TX>
__global__ __launch_bounds__(256) void kpp_seq_select_adopt_kernel(
    KppSeqCtl *__restrict__ ctl, const double *__restrict__ pot, int nt,
    const unsigned long long *__restrict__ cand, const TX *__restrict__ X, int d,
    double *__restrict__ centre, int64_t n, const double *__restrict__ nd,
    const double *__restrict__ pre_t, double *__restrict__ md, double *__restrict__ pre,
    unsigned long long *__restrict__ cand_next) {
    if (ctl->stop) return;
    __shared__ int bt;
    if (threadIdx.x == 0) {
        int b = 0;
        for (int t = 1; t < nt; ++t)
            if (pot[t] < pot[b]) b = t;
        bt = b;
    }
    __syncthreads();
    const int b = bt;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            ctl->best = b;
            ctl->total = pot[b];
        }
        const int64_t row = (int64_t)cand[b];
        for (int j = threadIdx.x; j < d; j += blockDim.x) centre[j] = (double)X[row * d + j];
        for (int t = threadIdx.x; t < nt; t += blockDim.x)
            cand_next[t] = (unsigned long long)(n - 1);
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = (int64_t)b * n + i;
    md[i] = nd[o];
    pre[i] = pre_t[o];
}

// v[0..m) = value.
static __global__ __attribute__((unused)) void kpp_fill_u64_kernel(unsigned long long *__restrict__ v,
                                                                    int64_t m,
                                                                    unsigned long long value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] = value;
}

// nd[t*n + i] = min(md[i], fold(x_i, x_cand[t])) for every trial t (grid.y).
template <typename TX>
__global__ void kpp_trials_multi_kernel(const TX *__restrict__ X, int64_t n, int d,
                                        const int64_t *__restrict__ cand,
                                        const double *__restrict__ md, double *__restrict__ nd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t t = blockIdx.y;
    const double v = fold_exact(X + i * d, X + cand[t] * d, d);
    const double m = md[i];
    nd[t * n + i] = v < m ? v : m;
}

// Sequential fold of each KPP_BLK block (one lane per block).
static __global__ __attribute__((unused)) void kpp_block_sums_kernel(const double *__restrict__ v, int64_t n,
                                      double *__restrict__ bsum) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lo = b * KPP_BLK;
    if (lo >= n) return;
    const int64_t hi = lo + KPP_BLK < n ? lo + KPP_BLK : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s = s + v[i];
    bsum[b] = s;
}

// out[0] = total (block sums folded in order); for each of nthr thresholds,
// the first index whose two-level running sum reaches it (n-1 if none).
static __global__ __attribute__((unused)) void kpp_search_kernel(const double *__restrict__ v, int64_t n,
                                  const double *__restrict__ bsum, int64_t nblk,
                                  const double *__restrict__ thr, int nthr,
                                  int64_t *__restrict__ cand) {
    const int t = threadIdx.x;
    if (t >= nthr) return;
    const double th = thr[t];
    double run = 0.0;
    int64_t res = n - 1;
    for (int64_t b = 0; b < nblk; ++b) {
        const double s = bsum[b];
        if (run + s >= th) {
            const int64_t lo = b * KPP_BLK;
            const int64_t hi = lo + KPP_BLK < n ? lo + KPP_BLK : n;
            double c = run;
            int64_t found = -1;
            for (int64_t i = lo; i < hi; ++i) {
                c = c + v[i];
                if (c >= th) {
                    found = i;
                    break;
                }
            }
            if (found >= 0) {
                res = found;
                break;
            }
        }
        run = run + s;
    }
    cand[t] = res;
}

// Batched form of the passes above, for d small enough to stage every
// trial's candidate row in LDS.  Distances are stored block-interleaved:
// the value of row b * KPP_BLK + i lives at [i * nblk + b] (rows past n
// hold +0, which changes no sum and never reaches a threshold first), so
// that one lane per block folds its block in row order with coalesced
// loads.  All sums keep the two-level order of the passes above.
constexpr int KPP_MAXT = 11;   // trials per batched pass (n_trials <= 11 for k <= 8192)

template <typename TX>
__global__ void kpp_first_il_kernel(const TX *__restrict__ X, int64_t n, int d,
                                    const double *__restrict__ c0, int64_t nblk,
                                    double *__restrict__ md) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk * KPP_BLK) return;
    const int64_t r = (t % nblk) * KPP_BLK + t / nblk;
    md[t] = r < n ? fold_exact(X + r * d, c0, d) : 0.0;
}

// nd_t = min(md, fold(x, cand_t)) for every trial t in one pass over X
// (the reference's `if d < nd`).  A block (64 * I threads) takes I rows from
// each of 64 consecutive distance blocks: the rows are staged in LDS with
// whole-row coalesced loads (row stride d + 1 against bank conflicts),
// thread (ii, bb) folds row (b0 + bb) * KPP_BLK + i0 + ii against the
// candidate rows (LDS, f64), and its results land at consecutive
// interleaved addresses.
template <typename TX, int NT>
__global__ __launch_bounds__(256) void kpp_trials_il_kernel(
    const TX *__restrict__ X, int64_t n, int d, const int64_t *__restrict__ cand,
    int64_t nblk, int I, const double *__restrict__ md, double *__restrict__ nd) {
    constexpr int nt = NT;
    extern __shared__ double cr[];                            // nt * d, then the row tile
    TX *tile = reinterpret_cast<TX *>(cr + (int64_t)nt * d);  // (64 * I) x (d + 1)
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int u = t; u < nt * d; u += blockDim.x) cr[u] = (double)X[cand[u / d] * d + u % d];
    const int64_t b0 = (int64_t)blockIdx.x * 64;
    const int64_t i0 = (int64_t)blockIdx.y * I;
    // wave w stages rows slot = ii * 64 + bb for ii == w, 32 rows' loads in
    // flight before their LDS stores
    for (int j = lane; j < d; j += 64)
        for (int bb0 = 0; bb0 < 64; bb0 += 32) {
            TX v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                // unconditional load (row 0 stands in for rows past n), then select
                const int64_t r = (b0 + bb0 + u) * KPP_BLK + i0 + w;
                const bool ok = b0 + bb0 + u < nblk && r < n;
                const TX x = X[(ok ? r : 0) * d + j];
                v[u] = ok ? x : (TX)0;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) tile[(int64_t)(w * 64 + bb0 + u) * (d + 1) + j] = v[u];
        }
    __syncthreads();
    const int ii = w, bb = lane;
    const int64_t b = b0 + bb;
    if (b >= nblk) return;
    const int64_t stride = nblk * KPP_BLK;
    const int64_t slot = (i0 + ii) * nblk + b;
    const int64_t r = b * KPP_BLK + i0 + ii;
    if (r >= n) {
        for (int q = 0; q < nt; ++q) nd[q * stride + slot] = 0.0;
        return;
    }
    const TX *x = tile + (int64_t)(ii * 64 + bb) * (d + 1);
    double acc[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[q] = 0.0;
    for (int j = 0; j < d; ++j) {
        const double xj = (double)x[j];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const double df = xj - cr[q * d + j];
            acc[q] = acc[q] + df * df;
        }
    }
    const double m = md[slot];
#pragma unroll
    for (int q = 0; q < NT; ++q) nd[q * stride + slot] = acc[q] < m ? acc[q] : m;
}

// Block sums of nt interleaved arrays: lane (q, b) folds block b of array q
// in row order.
static __global__ __attribute__((unused)) void kpp_bsums_il_kernel(const double *__restrict__ v,
                                                                   int64_t nblk, int nt,
                                                                   double *__restrict__ bsum) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk * nt) return;
    const int64_t q = t / nblk, b = t % nblk;
    const double *p = v + q * nblk * KPP_BLK + b;
    double s = 0.0;
    for (int i0 = 0; i0 < KPP_BLK; i0 += 32) {
        double w[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) w[u] = p[(int64_t)(i0 + u) * nblk];
#pragma unroll
        for (int u = 0; u < 32; ++u) s = s + w[u];
    }
    bsum[t] = s;
}

// tot[q] = the block sums of array q folded in order (one wave per array;
// 64 values are loaded at once and added in order through shuffles).
static __global__ __attribute__((unused)) __launch_bounds__(64) void kpp_totals_kernel(
    const double *__restrict__ bsum, int64_t nblk, double *__restrict__ tot) {
    const int lane = threadIdx.x;
    const double *p = bsum + (int64_t)blockIdx.x * nblk;
    double acc = 0.0;
    for (int64_t b0 = 0; b0 < nblk; b0 += 64) {
        const double v = b0 + lane < nblk ? p[b0 + lane] : 0.0;
        const int m = nblk - b0 < 64 ? (int)(nblk - b0) : 64;
        for (int u = 0; u < m; ++u) acc = acc + wave_bcast(v, u);
    }
    if (lane == 0) tot[blockIdx.x] = acc;
}

// The search of kpp_search_kernel on an interleaved array: one wave per
// threshold, values broadcast through shuffles so that every lane follows
// the same sequential running sums.
static __global__ __attribute__((unused)) __launch_bounds__(64) void kpp_search_il_kernel(
    const double *__restrict__ v, int64_t n, const double *__restrict__ bsum, int64_t nblk,
    const double *__restrict__ thr, int64_t *__restrict__ cand) {
    const int lane = threadIdx.x;
    const double th = thr[blockIdx.x];
    double run = 0.0;
    int64_t res = n - 1;
    bool done = false;
    for (int64_t b0 = 0; b0 < nblk && !done; b0 += 64) {
        const double sv = b0 + lane < nblk ? bsum[b0 + lane] : 0.0;
        const int m = nblk - b0 < 64 ? (int)(nblk - b0) : 64;
        for (int u = 0; u < m && !done; ++u) {
            const double s = wave_bcast(sv, u);
            const int64_t b = b0 + u;
            if (run + s >= th) {
                double c = run;
                for (int i0 = 0; i0 < KPP_BLK && !done; i0 += 64) {
                    const double w = v[(int64_t)(i0 + lane) * nblk + b];
                    for (int uu = 0; uu < 64; ++uu) {
                        c = c + wave_bcast(w, uu);
                        if (c >= th) {
                            res = b * KPP_BLK + i0 + uu;
                            done = true;
                            break;
                        }
                    }
                }
            }
            run = run + s;
        }
    }
    if (lane == 0) cand[blockIdx.x] = res;
}

}  // namespace fl
