// exact_fold.hpp -- the left fold s_{i+1} = s_i + v_i (s_0 = +0, each add
// rounded to nearest-even) computed by a whole workgroup, bit for bit the
// value a single thread adding in index order gets.
//
// The reference folds this way wherever it sums serially: k-means++'s
// totals, threshold searches and potentials (ferrolearn-cluster/src/
// kmeans.rs:200,207-213,219-228), the serial-path inertia (:310-322),
// MiniBatch's inertia (mini_batch_kmeans.rs:366-373) and each centroid's
// per-feature sum over its rows (kmeans.rs:370-386).  One lane doing the n
// dependent adds is latency bound; the fold is split by binades instead.
//
// Method.  While |s| lies in one binade [2^E, 2^(E+1)) (or below 2^-1021,
// where the spacing is 2^-1074 down to zero and across the sign), every
// value s can take there is a multiple of u = 2^le, le = E - 52 (or -1074),
// so s = m*u for an integer m.  For an addend x with |x| < 2^(le+53) write
// x/u = a + f (integer a = floor(x/u), 0 <= f < 1).  The exact sum is
// t = (m + a + f)*u; while t stays in the same binade its rounding is on
// the grid u, and unless f is exactly 1/2 it does not depend on m:
//     fl(s + x) = (m + a + [f > 1/2]) * u.
// So within a phase every increment a + [f > 1/2] is known up front, one
// block scan of them gives m before every element, and a walk checks that
// each exact sum stays in the binade (on m + a, which bounds t/u within 1).
// The first element that leaves the binade, has f == 1/2 (round half to
// even depends on m's parity), or is NaN, infinite or too large is added by
// the hardware, exactly as the serial loop would, and the next phase starts
// from the new binade.  After a crossing a few elements are added serially,
// which covers the short phases while the running sum is small.  Once s is
// infinite or NaN the rest follows in one parallel pass (s stays, or turns
// NaN at the first NaN or opposite infinity).  Every thread holds the same s
// and takes the same branches.
//
// The integers m, a and the increments are held in doubles: every one the
// method relies on is below 2^53 in magnitude (so exact), and the block scan
// of a phase's increments sums ranges whose totals are differences of two
// such m.
#pragma once

#include <climits>

#include <rocprim/block/block_scan.hpp>

#include "common.hpp"

namespace fl {

constexpr int EF_NT = 512;                 // threads per fold
constexpr int EF_NW = EF_NT / 64;          // waves per fold
constexpr int EF_P = 16;                   // window elements per thread
constexpr int EF_SERIAL = 32;              // serial adds after a binade crossing

// The reference assigns serially, folding the inertia in row order, when
// n * d < PARALLEL_WORK_THRESHOLD (ferrolearn-cluster/src/kmeans.rs:261,
// 286-290, 310-322); above it rayon sums per-chunk folds in an unspecified
// order.
constexpr int64_t KM_SERIAL_WORK = 100000;

// LDS position of window element i: one pad double after every 16, so that
// a thread's 16 contiguous elements load without bank conflicts.
__device__ __forceinline__ int ef_pad(int i) { return i + (i >> 4); }

// The fold of src[0..n) by the calling block of EF_NT threads (every
// thread gets the result), starting from s0 (+0 for a whole sequence, the
// running sum before src[0] for a piece of one); when pre is given, also the
// running sums pre[i] = s_{i+1}.  Thread t owns elements [t*P, t*P + P) of
// each window of EF_NT*P and keeps them in registers through the window's
// phases: P = 16 for long sequences, P = 4 for pieces of at most 2048
// elements (a phase's walk is then 4 elements long).
template <int P>
__device__ __forceinline__ double exact_fold_block_p(const double *__restrict__ src, int64_t n,
                                                     double *__restrict__ pre, double s0) {
    static_assert(P == 4 || P == 16, "window layout: P elements per thread, P | 16");
    constexpr int W = EF_NT * P;
    constexpr int WPAD = W + W / 16;
    using ScanD = rocprim::block_scan<double, EF_NT>;
    __shared__ double buf[2][WPAD];
    __shared__ typename ScanD::storage_type scan_st;
    __shared__ int sh_xc[EF_NW], sh_mc[EF_NW];
    __shared__ double m_at[EF_NT];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int rs = t * P;   // this thread's slice of every window
    // window 0 into buffer 0
#pragma unroll
    for (int u = 0; u < P; ++u) {
        const int i = u * EF_NT + t;
        if (i < n) buf[0][ef_pad(i)] = src[i];
    }
    __syncthreads();
    double s = s0;
    bool crossed = false;
    int cur = 0;
    for (int64_t w0 = 0; w0 < n; w0 += W, cur ^= 1) {
        const int wn = (int)(n - w0 < W ? n - w0 : W);
        // the next window's loads are in flight while this one is folded
        double nx[P];
        const int64_t w1 = w0 + W;
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int64_t i = w1 + u * EF_NT + t;
            nx[u] = i < n ? src[i] : 0.0;
        }
        const double *bw = buf[cur];
        double x[P];
#pragma unroll
        for (int u = 0; u < P; ++u) x[u] = bw[ef_pad(rs + u)];   // past wn: unused
        const int re = rs + P < wn ? rs + P : wn;
        double *pw = pre ? pre + w0 : nullptr;
        int j = 0;
        while (j < wn) {
            if (s != s || __builtin_isinf(s)) {
                // s stays until the first NaN or opposite infinity, then NaN
                int pz = s != s ? j : INT_MAX;
                if (s == s)
#pragma unroll
                    for (int u = 0; u < P; ++u) {
                        const int i = rs + u;
                        if (pz == INT_MAX && i >= j && i < re && (x[u] != x[u] || x[u] == -s))
                            pz = i;
                    }
                {
                    const uint64_t bz = __ballot(pz != INT_MAX);
                    const int wz = bz ? __shfl(pz, __builtin_ctzll(bz), 64) : INT_MAX;
                    if (lane == 0) sh_xc[wv] = wz;
                }
                __syncthreads();
                pz = INT_MAX;
#pragma unroll
                for (int i = 0; i < EF_NW; ++i)
                    if (pz == INT_MAX) pz = sh_xc[i];
                const double sn = s != s ? s : (pz < wn ? s + bw[ef_pad(pz)] : s);
                if (pw)
#pragma unroll
                    for (int u = 0; u < P; ++u) {
                        const int i = rs + u;
                        if (i >= j && i < re) pw[i] = i < pz ? s : sn;
                    }
                s = pz < wn ? sn : s;
                j = wn;
                __syncthreads();
                break;
            }
            if (crossed) {
                // a few serial adds (every thread, same values; thread 0 stores)
                const int lim = j + EF_SERIAL < wn ? j + EF_SERIAL : wn;
                while (j < lim) {
                    double b8[8];
                    const int cnt = lim - j < 8 ? lim - j : 8;
#pragma unroll
                    for (int u = 0; u < 8; ++u) b8[u] = u < cnt ? bw[ef_pad(j + u)] : 0.0;
                    bool nf = false;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (u < cnt && !nf) {
                            s = s + b8[u];
                            if (pw && t == 0) pw[j + u] = s;
                            if (s != s || __builtin_isinf(s)) {
                                nf = true;
                                j += u + 1;
                            }
                        }
                    if (nf) break;
                    j += cnt;
                }
                crossed = false;
                continue;
            }
            // one binade phase over [j, wn)
            const bool sub = !(__builtin_fabs(s) >= 0x1p-1021);
            const int le = sub ? -1074 : ilogb(s) - 52;
            const double U = ldexp(1.0, le + 53);   // addends at least this large cross
            const double m0 = ldexp(s, -le);
            // t = floor of the exact (m + x/u) stays in the binade when it
            // lies in [plo, phi] or [nlo, nhi] (t is m + a rounded: exact
            // when it does, and rounding is monotone)
            const double plo = sub ? -0x1p52 + 1.0 : 0x1p52, phi = sub ? 0x1p52 - 2.0 : 0x1p53 - 2.0;
            const double nlo = sub ? 1.0 : -0x1p53 + 1.0, nhi = sub ? 0.0 : -0x1p52 - 1.0;
            // pass 1: a = floor(x / u) and [f > 1/2]; the thread's first
            // element the phase cannot take (f == 1/2, NaN, infinite or too
            // large); the increments before it
            double a[P];
            int upm = 0, live = 0;
            bool open = true;
            double qs = 0.0;
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int i = rs + u;
                const double y = ldexp(x[u], -le);
                const double fa = floor(y);
                const double f = y - fa;
                const bool in = i >= j && i < re;
                const bool ok = open && in && __builtin_fabs(x[u]) < U && f != 0.5;
                open = open && (!in || ok);
                a[u] = fa;
                live |= (int)ok << u;
                upm |= (int)(f > 0.5) << u;
                qs += ok ? fa + (f > 0.5 ? 1.0 : 0.0) : 0.0;
            }
            // m before each thread's range
            double mx;
            ScanD().exclusive_scan(qs, mx, 0.0, scan_st,
                                   [](double p1, double p2) { return p1 + p2; });
            double m = m0 + mx;
            // pass 2: running sums (in registers) up to the first element the
            // phase cannot take or whose exact sum would leave the binade
            int mc = INT_MAX;
            double pv[P];
            bool go = true;
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int i = rs + u;
                const bool in = i >= j && i < re;
                const double tt = m + a[u];
                const bool st = (tt >= plo && tt <= phi) || (tt >= nlo && tt <= nhi);
                const bool adv = go && ((live >> u) & 1) && st;
                mc = go && in && !adv ? i : mc;
                go = go && (!in || adv);
                m = adv ? tt + (double)((upm >> u) & 1) : m;
                pv[u] = ldexp(m, le);
            }
            // the phase's end: threads hold increasing ranges, so the first
            // lane and wave that stopped hold it; each thread's m where it
            // stopped (the end's thread stopped at the end)
            {
                const uint64_t bm = __ballot(mc != INT_MAX);
                const int wmc = bm ? __shfl(mc, __builtin_ctzll(bm), 64) : INT_MAX;
                if (lane == 0) sh_mc[wv] = wmc;
                m_at[t] = m;
            }
            __syncthreads();
            int end = INT_MAX;
#pragma unroll
            for (int i = 0; i < EF_NW; ++i)
                if (end == INT_MAX) end = sh_mc[i];
            if (pw) {
#pragma unroll
                for (int u = 0; u < P; ++u) {
                    const int i = rs + u;
                    if (i >= j && i < re && i < end) pw[i] = pv[u];
                }
            }
            const int owner = (end < wn ? end : wn - 1) / P;
            const double se = ldexp(m_at[owner], le);
            if (end >= wn) {
                s = se;
                j = wn;
            } else {
                s = se + bw[ef_pad(end)];
                if (pw && t == 0) pw[end] = s;
                j = end + 1;
                crossed = true;
            }
        }
        // stage the next window (its buffer's last readers finished before
        // this window's phases began)
        if (w1 < n) {
            double *bn = buf[cur ^ 1];
#pragma unroll
            for (int u = 0; u < P; ++u) bn[ef_pad(u * EF_NT + t)] = nx[u];
        }
        __syncthreads();
    }
    return s;
}

__device__ __forceinline__ double exact_fold_block(const double *__restrict__ src, int64_t n,
                                                   double *__restrict__ pre, double s0 = 0.0) {
    return exact_fold_block_p<EF_P>(src, n, pre, s0);
}

// grid.x = number of folds: fold b reads v[b*ld .. b*ld + n), writes its
// result to sums[b] and, when prefix is given, the running sums
// prefix[b*ld + i] = s_{i+1}.  A non-zero *stop (when given) makes the
// launch a no-op.
static __global__ __attribute__((unused)) __launch_bounds__(EF_NT) void exact_fold_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, const int32_t *__restrict__ stop,
    double *__restrict__ sums, double *__restrict__ prefix) {
    if (stop && *stop) return;
    const double s = exact_fold_block(v + (int64_t)blockIdx.x * ld, n,
                                      prefix ? prefix + (int64_t)blockIdx.x * ld : nullptr);
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = s;
}

}  // namespace fl
