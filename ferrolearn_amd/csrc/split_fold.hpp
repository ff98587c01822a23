// split_fold.hpp -- the left fold of exact_fold.hpp (s_{i+1} = s_i + v_i
// from s_0 = +0, every add rounded to nearest-even; bit for bit the serial
// loop's result and running sums) spread over many compute units.
//
// The single-workgroup fold is bound by one compute unit's dependent-op
// latency (about 2 ns per element with running sums).  Here a sequence is cut
// into blocks of SF_L elements, and each block is done in parallel on the
// guess of the binade its running sum starts in:
//
//   sf_part_kernel    plain sum of each block (any order; only a guess)
//   sf_guess_kernel   per sequence, the exclusive scan of those sums: the
//                     approximate running sum at each block start, and from
//                     it the grid exponent le_b the block is done on
//   sf_spec_kernel    per element on grid u = 2^le_b (exact_fold.hpp's
//                     method): x/u = a + f, increment a + [f > 1/2]; an int64
//                     block scan gives the increments' running totals P_i,
//                     and the block keeps its total M and the least and
//                     greatest P_i + a_i (the binade test of every element).
//                     Elements the grid cannot take (f == 1/2, NaN, infinite,
//                     |x/u| >= 2^51) mark the block.
//   sf_combine_kernel one workgroup per sequence walks the blocks in order
//                     with the exact running sum s = m * 2^le (integer m).
//                     A block is taken whole when it is unmarked, le matches
//                     its guess and m + min(P_i + a_i), m + max(P_i + a_i)
//                     lie in the binade's range (the same test the serial
//                     method makes element by element): then s becomes
//                     (m + M) * 2^le exactly.  Any other block (a binade
//                     crossing, a wrong guess, a marked element, s not
//                     finite) is folded from the exact s by the workgroup
//                     fold (one window of 4 elements per thread), which also
//                     writes its running sums.
//   sf_prefix_kernel  running sums of the blocks taken whole:
//                     (m_start + P_i + increment_i) * 2^le.
//
// The integer work is exact: |increment| <= 2^51, so a block's running
// totals stay below 2^62 in magnitude, and the range tests compare
// min/max(P_i + a_i) against (bound - m) without overflow.  The guesses only
// decide how much work the combine redoes, never the result.
#pragma once

#include <rocprim/block/block_scan.hpp>

#include "exact_fold.hpp"

namespace fl {

constexpr int SF_NT = 256;             // threads per block in the parallel kernels
constexpr int SF_P = 8;                // elements per thread
constexpr int SF_L = SF_NT * SF_P;     // elements per block
constexpr int SF_NO_GUESS = INT_MIN;   // le of a block whose start is not finite
static_assert(SF_L == EF_NT * 4, "a block is one window of the 4-per-thread workgroup fold");

struct SfBlock {
    int64_t M;      // sum of the increments
    int64_t tmin;   // least P_i + a_i
    int64_t tmax;   // greatest P_i + a_i
    int32_t le;     // grid exponent the block was done on
    int32_t mark;   // an element the grid cannot take
};

struct SfStart {
    int64_t m;      // exact running sum before the block, in units of 2^le
    int32_t le;
    int32_t whole;  // taken whole by the combine (its running sums are still
                    // integers in prefix)
};

// Grid exponent of the binade of s (exact_fold.hpp's convention: the
// spacing 2^-1074 below 2^-1021).
__device__ __forceinline__ int sf_grid(double s) {
    return !(__builtin_fabs(s) >= 0x1p-1021) ? -1074 : ilogb(s) - 52;
}

// grid (nb, count): part[q * nb + b] = plain sum of block b of sequence q.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_part_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, int64_t nb,
    const int32_t *__restrict__ stop, double *__restrict__ part) {
    if (stop && *stop) return;
    const int64_t b = blockIdx.x, q = blockIdx.y;
    const double *src = v + q * ld + b * SF_L;
    const int64_t len = n - b * SF_L < SF_L ? n - b * SF_L : SF_L;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = u * SF_NT + (int)threadIdx.x;
        if (i < len) acc += src[i];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double w[SF_NT / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < SF_NT / 64; ++i) t += w[i];
        part[q * nb + b] = t;
    }
}

// grid (count), SF_NT threads: gle[q * nb + b] = grid exponent of the
// approximate running sum before block b.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_guess_kernel(
    const double *__restrict__ part, int64_t nb, const int32_t *__restrict__ stop,
    int32_t *__restrict__ gle) {
    if (stop && *stop) return;
    using Scan = rocprim::block_scan<double, SF_NT>;
    __shared__ typename Scan::storage_type st;
    const int64_t q = blockIdx.x;
    const double *p = part + q * nb;
    int32_t *g = gle + q * nb;
    const int64_t per = (nb + SF_NT - 1) / SF_NT;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    const int64_t b1 = b0 + per < nb ? b0 + per : nb;
    double loc = 0.0;
    for (int64_t b = b0; b < b1; ++b) loc += p[b];
    double run;
    Scan().exclusive_scan(loc, run, 0.0, st, [](double a, double c) { return a + c; });
    for (int64_t b = b0; b < b1; ++b) {
        g[b] = (run == run && !__builtin_isinf(run)) ? sf_grid(run) : SF_NO_GUESS;
        run += p[b];
    }
}

// grid (nb, count): each block's increments on its guessed grid.  With
// prefix, the increments' inclusive running totals go to prefix (as int64
// bits) for sf_prefix_kernel.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_spec_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, int64_t nb,
    const int32_t *__restrict__ stop, const int32_t *__restrict__ gle,
    SfBlock *__restrict__ blk, double *__restrict__ prefix) {
    if (stop && *stop) return;
    using Scan = rocprim::block_scan<int64_t, SF_NT>;
    __shared__ typename Scan::storage_type st;
    __shared__ int64_t red[2][SF_NT / 64];
    const int64_t b = blockIdx.x, q = blockIdx.y;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t base = q * ld + b * SF_L;
    const double *src = v + base;
    const int64_t len = n - b * SF_L < SF_L ? n - b * SF_L : SF_L;
    const int le = gle[q * nb + b];
    const bool guess = le != SF_NO_GUESS;
    // this thread's SF_P consecutive elements
    const int r0 = t * SF_P;
    int64_t a[SF_P], inc[SF_P];
    int mark = guess ? 0 : 1;
    int64_t tot = 0;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = r0 + u;
        a[u] = 0;
        inc[u] = 0;
        if (i < len && guess) {
            const double y = ldexp(src[i], -le);
            const bool ok = __builtin_fabs(y) < 0x1p51;   // false for NaN and infinities
            const double fa = floor(y);
            const double f = y - fa;
            if (!ok || f == 0.5) {
                mark = 1;
            } else {
                a[u] = (int64_t)fa;
                inc[u] = a[u] + (f > 0.5 ? 1 : 0);
            }
        }
        tot += inc[u];
    }
    int64_t run;
    Scan().exclusive_scan(tot, run, (int64_t)0, st,
                          [](int64_t x, int64_t y) { return x + y; });
    int64_t tmin = INT64_MAX, tmax = INT64_MIN;
    int64_t *pq = prefix ? reinterpret_cast<int64_t *>(prefix + base) : nullptr;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = r0 + u;
        if (i < len) {
            const int64_t tt = run + a[u];
            tmin = tt < tmin ? tt : tmin;
            tmax = tt > tmax ? tt : tmax;
            run += inc[u];
            if (pq) pq[i] = run;
        }
    }
    // block min / max / total / mark
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t mn = __shfl_xor(tmin, o, 64), mx = __shfl_xor(tmax, o, 64);
        tmin = mn < tmin ? mn : tmin;
        tmax = mx > tmax ? mx : tmax;
    }
    if (lane == 0) {
        red[0][wv] = tmin;
        red[1][wv] = tmax;
    }
    mark = __syncthreads_or(mark);
    if (t == SF_NT - 1) {
        // the last thread's running total after its elements is the block's M
        int64_t mn = INT64_MAX, mx = INT64_MIN;
#pragma unroll
        for (int w = 0; w < SF_NT / 64; ++w) {
            mn = red[0][w] < mn ? red[0][w] : mn;
            mx = red[1][w] > mx ? red[1][w] : mx;
        }
        SfBlock o;
        o.M = run;
        o.tmin = mn;
        o.tmax = mx;
        o.le = le;
        o.mark = mark;
        blk[q * nb + b] = o;
    }
}

// Exact running state s = m * 2^le of a finite s.
__device__ __forceinline__ void sf_state(double s, int &le, int64_t &m) {
    le = sf_grid(s);
    m = (int64_t)ldexp(s, -le);
}

// Whether a block done on grid le can be taken whole from the exact state
// (le, m): exact_fold.hpp's binade test (on m + a for every element).
__device__ __forceinline__ bool sf_fits(int le, int64_t m, const SfBlock &B) {
    int64_t lo, hi;
    if (le == -1074) {
        lo = -((int64_t)1 << 52) + 1;
        hi = ((int64_t)1 << 52) - 2;
    } else if (m > 0) {
        lo = (int64_t)1 << 52;
        hi = ((int64_t)1 << 53) - 2;
    } else {
        lo = -((int64_t)1 << 53) + 1;
        hi = -((int64_t)1 << 52) - 1;
    }
    return B.tmin >= lo - m && B.tmax <= hi - m;
}

// grid (count), EF_NT threads: the walk over the blocks (see the top).
static __global__ __launch_bounds__(EF_NT) __attribute__((unused)) void sf_combine_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, int64_t nb,
    const int32_t *__restrict__ stop, const SfBlock *__restrict__ blk,
    SfStart *__restrict__ starts, double *__restrict__ sums, double *__restrict__ prefix) {
    if (stop && *stop) return;
    constexpr int CH = 64;   // block summaries staged per round
    __shared__ SfBlock sb[CH];
    const int64_t q = blockIdx.x;
    const int t = threadIdx.x;
    const double *src = v + q * ld;
    double *pre = prefix ? prefix + q * ld : nullptr;
    const SfBlock *B = blk + q * nb;
    SfStart *S = starts + q * nb;
    double s = 0.0;
    bool fin = true;   // s finite: the state (le, m) is valid
    int le = -1074;
    int64_t m = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += CH) {
        const int cn = nb - c0 < CH ? (int)(nb - c0) : CH;
        if (t < cn) sb[t] = B[c0 + t];
        __syncthreads();
        for (int c = 0; c < cn; ++c) {
            const int64_t b = c0 + c;
            const SfBlock bk = sb[c];
            if (fin && !bk.mark && bk.le == le && sf_fits(le, m, bk)) {
                if (t == 0) {
                    SfStart o;
                    o.m = m;
                    o.le = le;
                    o.whole = 1;
                    S[b] = o;
                }
                m += bk.M;
                continue;
            }
            // fold the block from the exact s
            s = fin ? ldexp((double)m, le) : s;
            if (t == 0) {
                SfStart o;
                o.m = 0;
                o.le = 0;
                o.whole = 0;
                S[b] = o;
            }
            const int64_t len = n - b * SF_L < SF_L ? n - b * SF_L : SF_L;
            s = exact_fold_block_p<4>(src + b * SF_L, len, pre ? pre + b * SF_L : nullptr, s);
            fin = s == s && !__builtin_isinf(s);
            if (fin) sf_state(s, le, m);
        }
        __syncthreads();   // sb is staged again
    }
    if (fin) s = ldexp((double)m, le);
    if (t == 0 && sums) sums[q] = s;
}

// grid (nb, count): running sums of the blocks taken whole.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_prefix_kernel(
    int64_t n, int64_t ld, int64_t nb, const int32_t *__restrict__ stop,
    const SfStart *__restrict__ starts, double *__restrict__ prefix) {
    if (stop && *stop) return;
    const int64_t b = blockIdx.x, q = blockIdx.y;
    const SfStart S = starts[q * nb + b];
    if (!S.whole) return;
    const int64_t len = n - b * SF_L < SF_L ? n - b * SF_L : SF_L;
    double *p = prefix + q * ld + b * SF_L;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = u * SF_NT + (int)threadIdx.x;
        if (i < len) {
            const int64_t r = reinterpret_cast<const int64_t *>(p)[i];
            p[i] = ldexp((double)(S.m + r), S.le);
        }
    }
}

}  // namespace fl
