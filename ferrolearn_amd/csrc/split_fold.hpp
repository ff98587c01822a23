// split_fold.hpp -- the left fold of exact_fold.hpp (s_{i+1} = s_i + v_i
// from s_0 = +0, every add rounded to nearest-even; bit for bit the serial
// loop's result and running sums) spread over many compute units.
//
// The single-workgroup fold is bound by one compute unit's dependent-op
// latency (about 2 ns per element with running sums).  Here a sequence is cut
// into blocks of SF_L elements, and each block is done in parallel on a
// prediction of where its running sum goes:
//
//   sf_part_kernel    plain sum of each block (any order: only a prediction)
//   sf_guess_kernel   per sequence, the exclusive scan of those sums: an
//                     approximate running sum at each block start
//   sf_spec_kernel    per block, approximate running sums of its elements
//                     (a block scan) and from them each element's predicted
//                     grid exponent le (exact_fold.hpp's u = 2^le).  The
//                     block is cut into parts at every element predicted to
//                     leave its binade, to be an exact tie (x/u = a + 1/2),
//                     or too large for the grid (|x/u| >= 2^51, NaN,
//                     infinite): that element, the part's "cut", is added by
//                     the hardware.  Within a part every element is done on
//                     the part's grid: x/u = a + f, increment a + [f > 1/2],
//                     an int64 block scan gives the running totals P_i, and
//                     the part keeps its total M and the least and greatest
//                     P_i + a_i (exact_fold.hpp's binade test of each element).
//   sf_combine_kernel one workgroup per sequence walks the parts in order with
//                     the exact running sum s = m * 2^le (integer m).  A part
//                     is taken whole when none of its elements was refused,
//                     le matches its grid and m + min(P_i + a_i),
//                     m + max(P_i + a_i) lie in the binade's range: then s
//                     becomes (m + M) * 2^le exactly.  A cut is s + x in f64,
//                     the serial loop's own step.  Anything else (a wrong
//                     prediction, a refused element, s not finite) is folded
//                     from the exact s to the block's end by the workgroup
//                     fold, which also writes those running sums.
//   sf_prefix_kernel  running sums of the parts taken whole:
//                     (m_start + P_i + increment_i) * 2^le.
//
// The integer work is exact: |increment| <= 2^51, so a block's running
// totals stay below 2^62 in magnitude, and the range tests compare
// min/max(P_i + a_i) against (bound - m) without overflow.  The predictions
// only decide how much work the combine redoes, never the result.
#pragma once

#include <rocprim/block/block_scan.hpp>

#include "exact_fold.hpp"

namespace fl {

constexpr int SF_NT = 256;             // threads per block in the parallel kernels
constexpr int SF_P = 8;                // elements per thread
constexpr int SF_L = SF_NT * SF_P;     // elements per block
constexpr int SF_NP = 32;              // parts kept per block
constexpr int SF_NO_GUESS = INT_MIN;   // grid exponent of a running sum that is not finite
static_assert(SF_L == EF_NT * 4, "a block is one window of the 4-per-thread workgroup fold");

struct SfPart {
    int64_t M;          // sum of the increments of its grid elements
    int64_t tmin;       // least P_i + a_i
    int64_t tmax;       // greatest P_i + a_i
    double cut;         // the element after the part, added by the hardware
    int32_t start, end; // grid elements [start, end) of the block; the cut is `end`
    int32_t le;         // the part's grid exponent
    int32_t flags;      // SF_REFUSED, SF_HAS_CUT
};
constexpr int32_t SF_REFUSED = 1;   // an element the grid cannot take
constexpr int32_t SF_HAS_CUT = 2;

struct SfStart {
    int64_t m;          // exact running sum before the part, in units of 2^le
    int32_t le;
    int32_t whole;      // taken whole (its running sums are still integers in prefix)
};

// Grid exponent of the binade of s (exact_fold.hpp's convention: the
// spacing 2^-1074 below 2^-1021); SF_NO_GUESS for NaN and infinities.
__device__ __forceinline__ int sf_grid(double s) {
    if (s != s || __builtin_isinf(s)) return SF_NO_GUESS;
    return !(__builtin_fabs(s) >= 0x1p-1021) ? -1074 : ilogb(s) - 52;
}

// Whether the grid of exponent le cannot take x: an exact tie, too large,
// NaN or infinite.
__device__ __forceinline__ bool sf_refused(double x, int le) {
    if (le == SF_NO_GUESS) return true;
    const double y = ldexp(x, -le);
    if (!(__builtin_fabs(y) < 0x1p51)) return true;
    return y - floor(y) == 0.5;
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

// grid (count), SF_NT threads: apx[q * nb + b] = approximate running sum
// before block b.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_guess_kernel(
    const double *__restrict__ part, int64_t nb, const int32_t *__restrict__ stop,
    double *__restrict__ apx) {
    if (stop && *stop) return;
    using Scan = rocprim::block_scan<double, SF_NT>;
    __shared__ typename Scan::storage_type st;
    const int64_t q = blockIdx.x;
    const double *p = part + q * nb;
    double *g = apx + q * nb;
    const int64_t per = (nb + SF_NT - 1) / SF_NT;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    const int64_t b1 = b0 + per < nb ? b0 + per : nb;
    double loc = 0.0;
    for (int64_t b = b0; b < b1; ++b) loc += p[b];
    double run;
    Scan().exclusive_scan(loc, run, 0.0, st, [](double a, double c) { return a + c; });
    for (int64_t b = b0; b < b1; ++b) {
        g[b] = run;
        run += p[b];
    }
}

__device__ __forceinline__ unsigned long long sf_enc(int64_t x) {
    return (unsigned long long)x ^ 0x8000000000000000ull;   // order-preserving
}
__device__ __forceinline__ int64_t sf_dec(unsigned long long e) {
    return (int64_t)(e ^ 0x8000000000000000ull);
}

// grid (nb, count): the parts of each block (see the top).  parts holds
// SF_NP entries per block, npart their count.  With prefix, each grid
// element's running total within its part (increment included) goes to
// prefix as int64 bits, for sf_prefix_kernel.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_spec_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, int64_t nb,
    const int32_t *__restrict__ stop, const double *__restrict__ apx,
    SfPart *__restrict__ parts, int32_t *__restrict__ npart, double *__restrict__ prefix) {
    if (stop && *stop) return;
    using ScanD = rocprim::block_scan<double, SF_NT>;
    using ScanI = rocprim::block_scan<int, SF_NT>;
    using ScanL = rocprim::block_scan<int64_t, SF_NT>;
    __shared__ union {
        typename ScanD::storage_type d;
        typename ScanI::storage_type i;
        typename ScanL::storage_type l;
    } st;
    __shared__ int32_t p_start[SF_NP], p_end[SF_NP], p_le[SF_NP], p_flags[SF_NP];
    __shared__ double p_cut[SF_NP];
    __shared__ int64_t p_q0[SF_NP], p_qe[SF_NP];
    __shared__ unsigned long long p_min[SF_NP], p_max[SF_NP];
    __shared__ int sh_np;
    const int64_t b = blockIdx.x, q = blockIdx.y;
    const int t = threadIdx.x;
    const int64_t base = q * ld + b * SF_L;
    const double *src = v + base;
    const int len = (int)(n - b * SF_L < SF_L ? n - b * SF_L : SF_L);
    const int r0 = t * SF_P;   // this thread's SF_P consecutive elements
    if (t < SF_NP) {
        p_end[t] = len;
        p_flags[t] = 0;
        p_min[t] = ~0ull;
        p_max[t] = 0ull;
    }
    double x[SF_P];
    double loc = 0.0;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        x[u] = r0 + u < len ? src[r0 + u] : 0.0;
        loc += x[u];
    }
    // predicted grids and cuts from the approximate running sums
    double ex;
    ScanD().exclusive_scan(loc, ex, 0.0, st.d, [](double a, double c) { return a + c; });
    double run = apx[q * nb + b] + ex;
    unsigned cutm = 0;
    int gnext[SF_P];
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        gnext[u] = SF_NO_GUESS;
        if (r0 + u < len) {
            const int g = sf_grid(run);
            const double nx = run + x[u];
            gnext[u] = sf_grid(nx);
            if (gnext[u] != g || sf_refused(x[u], g)) cutm |= 1u << u;
            if (r0 + u == 0) p_le[0] = g;
            run = nx;
        }
    }
    __syncthreads();   // scan storage reused
    int cex;
    ScanI().exclusive_scan(__builtin_popcount(cutm), cex, 0, st.i,
                           [](int a, int c) { return a + c; });
    // part of each element (cuts past the last part's start are ignored);
    // a cut ends its part and starts the next
    int pid[SF_P];
    unsigned effm = 0;
    {
        int c = cex;
#pragma unroll
        for (int u = 0; u < SF_P; ++u) {
            const int i = r0 + u;
            pid[u] = c < SF_NP - 1 ? c : SF_NP - 1;
            if (i < len && ((cutm >> u) & 1) && c < SF_NP - 1) {
                effm |= 1u << u;
                p_end[pid[u]] = i;
                p_cut[pid[u]] = x[u];
                p_flags[pid[u]] = SF_HAS_CUT;
                if (i + 1 < len) {
                    p_start[pid[u] + 1] = i + 1;
                    p_le[pid[u] + 1] = gnext[u];
                }
            }
            if ((cutm >> u) & 1) ++c;
            if (i == len - 1) sh_np = pid[u] + 1;
        }
    }
    if (t == 0) p_start[0] = 0;
    __syncthreads();
    // increments on the part's grid
    int64_t a[SF_P], inc[SF_P];
    int64_t tot = 0;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = r0 + u;
        a[u] = 0;
        inc[u] = 0;
        if (i < len && !((effm >> u) & 1)) {
            const int le = p_le[pid[u]];
            if (sf_refused(x[u], le)) {
                atomicOr(&p_flags[pid[u]], SF_REFUSED);
            } else {
                const double y = ldexp(x[u], -le);
                const double fa = floor(y);
                a[u] = (int64_t)fa;
                inc[u] = a[u] + (y - fa > 0.5 ? 1 : 0);
            }
        }
        tot += inc[u];
    }
    int64_t qrun;
    ScanL().exclusive_scan(tot, qrun, (int64_t)0, st.l,
                           [](int64_t c, int64_t e) { return c + e; });
    int64_t qi[SF_P];
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = r0 + u;
        qi[u] = qrun;
        if (i < len) {
            if (i == p_start[pid[u]] || i == 0) p_q0[pid[u]] = qrun;
            if ((effm >> u) & 1) p_qe[pid[u]] = qrun;
            else if (i == len - 1) p_qe[pid[u]] = qrun + inc[u];
        }
        qrun += inc[u];
    }
    __syncthreads();
    int64_t *pq = prefix ? reinterpret_cast<int64_t *>(prefix + base) : nullptr;
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = r0 + u;
        if (i < len && !((effm >> u) & 1)) {
            const int64_t q0 = p_q0[pid[u]];
            const int64_t tt = qi[u] - q0 + a[u];
            atomicMin(&p_min[pid[u]], sf_enc(tt));
            atomicMax(&p_max[pid[u]], sf_enc(tt));
            if (pq) pq[i] = qi[u] + inc[u] - q0;
        }
    }
    __syncthreads();
    const int np = sh_np;
    if (t < np) {
        SfPart o;
        o.M = p_qe[t] - p_q0[t];
        o.tmin = sf_dec(p_min[t]);
        o.tmax = sf_dec(p_max[t]);
        o.cut = (p_flags[t] & SF_HAS_CUT) ? p_cut[t] : 0.0;
        o.start = p_start[t];
        o.end = p_end[t];
        o.le = p_le[t];
        o.flags = p_flags[t];
        parts[(q * nb + b) * SF_NP + t] = o;
    }
    if (t == 0) npart[q * nb + b] = np;
}

// Exact running state s = m * 2^le of a finite s.
__device__ __forceinline__ void sf_state(double s, int &le, int64_t &m) {
    le = !(__builtin_fabs(s) >= 0x1p-1021) ? -1074 : ilogb(s) - 52;
    m = (int64_t)ldexp(s, -le);
}

// Whether a part done on grid le can be taken whole from the exact state
// (le, m): exact_fold.hpp's binade test (on m + a for every element).
__device__ __forceinline__ bool sf_fits(int le, int64_t m, const SfPart &B) {
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

// grid (count), EF_NT threads: the walk over the parts (see the top).  The
// parts of SF_GB blocks are staged in LDS per round trip.
constexpr int SF_GB = EF_NT / SF_NP;
static __global__ __launch_bounds__(EF_NT) __attribute__((unused)) void sf_combine_kernel(
    const double *__restrict__ v, int64_t n, int64_t ld, int64_t nb,
    const int32_t *__restrict__ stop, const SfPart *__restrict__ parts,
    const int32_t *__restrict__ npart, SfStart *__restrict__ starts, double *__restrict__ sums,
    double *__restrict__ prefix) {
    if (stop && *stop) return;
    __shared__ SfPart sp[SF_GB][SF_NP];
    __shared__ int32_t snp[SF_GB];
    const int64_t q = blockIdx.x;
    const int t = threadIdx.x;
    const double *src = v + q * ld;
    double *pre = prefix ? prefix + q * ld : nullptr;
    double s = 0.0;
    bool fin = true;   // s finite: the state (le, m) holds it
    int le = -1074;
    int64_t m = 0;
    for (int64_t g0 = 0; g0 < nb; g0 += SF_GB) {
        const int gn = nb - g0 < SF_GB ? (int)(nb - g0) : SF_GB;
        {
            // every slot of the group's blocks (those past a block's count
            // are not read)
            const int gb = t / SF_NP, p = t % SF_NP;
            if (gb < gn) {
                sp[gb][p] = parts[((q * nb + g0 + gb) * SF_NP) + p];
                if (p == 0) snp[gb] = npart[q * nb + g0 + gb];
            }
        }
        __syncthreads();
        for (int gb = 0; gb < gn; ++gb) {
            const int64_t b = g0 + gb;
            const int64_t qb = q * nb + b;
            const int np = snp[gb];
            const int len = (int)(n - b * SF_L < SF_L ? n - b * SF_L : SF_L);
            const double *bs = src + b * SF_L;
            double *bp = pre ? pre + b * SF_L : nullptr;
            bool fell = false;
            for (int p = 0; p < np; ++p) {
                const SfPart P = sp[gb][p];
                SfStart o;
                o.m = m;
                o.le = le;
                o.whole = 0;
                if (fell) {
                    if (t == 0) starts[qb * SF_NP + p] = o;
                    continue;
                }
                if (P.start < P.end) {
                    if (fin && !(P.flags & SF_REFUSED) && P.le == le && sf_fits(le, m, P)) {
                        o.whole = 1;
                        if (t == 0) starts[qb * SF_NP + p] = o;
                        m += P.M;
                    } else {
                        // the rest of the block from the exact s
                        if (t == 0) starts[qb * SF_NP + p] = o;
                        s = fin ? ldexp((double)m, le) : s;
                        s = exact_fold_block_p<4>(bs + P.start, len - P.start,
                                                  bp ? bp + P.start : nullptr, s);
                        fin = s == s && !__builtin_isinf(s);
                        if (fin) sf_state(s, le, m);
                        fell = true;
                        continue;
                    }
                } else if (t == 0) {
                    starts[qb * SF_NP + p] = o;
                }
                if (P.flags & SF_HAS_CUT) {
                    s = fin ? ldexp((double)m, le) : s;
                    s = s + P.cut;
                    if (t == 0 && bp) bp[P.end] = s;
                    fin = s == s && !__builtin_isinf(s);
                    if (fin) sf_state(s, le, m);
                }
            }
        }
        __syncthreads();   // sp is staged again
    }
    if (fin) s = ldexp((double)m, le);
    if (t == 0 && sums) sums[q] = s;
}

// grid (nb, count): running sums of the parts taken whole.
static __global__ __launch_bounds__(SF_NT) __attribute__((unused)) void sf_prefix_kernel(
    int64_t n, int64_t ld, int64_t nb, const int32_t *__restrict__ stop,
    const SfPart *__restrict__ parts, const int32_t *__restrict__ npart,
    const SfStart *__restrict__ starts, double *__restrict__ prefix) {
    if (stop && *stop) return;
    __shared__ int32_t ps[SF_NP], pe[SF_NP], pl[SF_NP], pw[SF_NP];
    __shared__ int64_t pm[SF_NP];
    const int64_t b = blockIdx.x, q = blockIdx.y, qb = q * nb + b;
    const int t = threadIdx.x;
    const int np = npart[qb];
    if (t < np) {
        const SfPart P = parts[qb * SF_NP + t];
        const SfStart S = starts[qb * SF_NP + t];
        ps[t] = P.start;
        pe[t] = P.end;
        pl[t] = S.le;
        pw[t] = S.whole;
        pm[t] = S.m;
    }
    __syncthreads();
    const int len = (int)(n - b * SF_L < SF_L ? n - b * SF_L : SF_L);
    double *p = prefix + q * ld + b * SF_L;
    int64_t r[SF_P];
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = u * SF_NT + t;
        r[u] = i < len ? reinterpret_cast<const int64_t *>(p)[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < SF_P; ++u) {
        const int i = u * SF_NT + t;
        if (i >= len) continue;
        int k = -1;
        for (int c = 0; c < np; ++c)
            if (i >= ps[c] && i < pe[c]) k = c;
        if (k >= 0 && pw[k]) p[i] = ldexp((double)(pm[k] + r[u]), pl[k]);
    }
}

}  // namespace fl
