// dist_bf16.hpp -- BF16 MFMA candidate filter with a fused per-row top-M.
//
// Same contract as dist_topm.hpp (every kept key is a rigorous lower bound of
// the exact fold; the M smallest keys per row are kept; theta is the M-th
// key), but the dot product a.b runs on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation:
//   PASSES = 1 : a.b ~ ah.bh                    (ah = bf16(a))
//   PASSES = 3 : a.b ~ ah.bh + ah.bl + al.bh    (al = bf16(a - ah))
// Operands are pre-split into hi / lo bf16 arrays padded to NKC*32 features.
//
// Roles: the streamed rows (kNN database / KMeans centres) are the MFMA A
// operand, staged through LDS 64 rows at a time; the resident rows (kNN
// queries / KMeans samples) are the B operand and stay in VGPRs for the
// whole kernel, 32 per wave.
//
// Candidate lists live in registers: a lane's accumulators hold 2 resident
// rows x 16 streamed rows per tile, and across the stream the four lanes
// that share a resident row see disjoint quarters of the streamed rows.  Each
// lane keeps a private sorted list of up to MP pairs of its quarter for both
// of its resident rows, and all four lanes filter against a shared threshold
// T (shared_threshold below) that is an upper bound of the MP-th smallest
// key kept so far.  A pair that is not kept was filtered (key >= T at that
// time) or evicted from a full list (key >= that list's last key >= T), and
// T only decreases, so the final T bounds every dropped pair from below; the
// lists hold at least MP pairs with key <= T.  The output is the MP smallest
// pairs of the four lists, and its bound is their MP-th key.
//
// Key of a pair, in f32:  key' = fma(-2, acc, xn'_r),  xn'_r = (1 - b)|x_r|^2.
// The lower bound of the exact fold is key' + qoff_q with the per-row
// constant qoff_q = (1 - b)|q|^2 - c |q| Smax (Smax = the largest |x_r| of
// the streamed rows): a constant per resident row does not change the order,
// so the epilogue costs one FMA and one compare per pair, and qoff is added
// back (in f64) for the certificate.  c and b bound the bf16 representation
// error of the split, the FP32 accumulation (taken as 2^-23 per addition,
// whatever order MFMA uses), the f32 epilogue and the exact fold's own
// rounding (host.hpp, bound_bf16); c |q| Smax >= c |q||x_r| for every row.
#pragma once

#include "common.hpp"
#include "dist_topm.hpp"

namespace fl {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

constexpr int BF_WAVES = 4;
constexpr int BF_QW = 32;                 // resident rows per wave (two 16-col tiles)
constexpr int BF_Q = BF_WAVES * BF_QW;    // resident rows per block
constexpr int BF_T = 64;                  // streamed rows per tile (four 16-row tiles)
constexpr int BF_STAGE = 256;

// Output layout of v_mfma_f32_16x16x32_bf16: lane l holds D[i][l%16] with
// i = bf_acc_row(l, v) for accumulator register v = 0..3.
__device__ __forceinline__ int bf_acc_row(int lane, int v) { return 4 * (lane >> 4) + v; }

// f32 -> bf16, round to nearest even (finite inputs; NaN stays NaN).
__host__ __device__ __forceinline__ unsigned short bf16_bits_rne(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_bits_to_f32(unsigned short h) {
    return __builtin_bit_cast(float, (unsigned int)h << 16);
}

// Split rows into hi / lo bf16 (zero padded to dpad features).
template <typename T>
__global__ void bf16_split_kernel(const T *__restrict__ X, int64_t n, int d, int dpad,
                                  unsigned short *__restrict__ hi,
                                  unsigned short *__restrict__ lo) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dpad) return;
    const int64_t r = t / dpad;
    const int j = (int)(t % dpad);
    unsigned short h = 0, l = 0;
    if (j < d) {
        const double x = (double)X[r * d + j];
        h = bf16_bits_rne((float)x);
        const double rem = x - (double)bf16_bits_to_f32(h);   // exact (Sterbenz)
        l = bf16_bits_rne((float)rem);
    }
    hi[t] = h;
    lo[t] = l;
}

// Per streamed row: xn' = (1-b)|x|^2 (f32).
static __global__ __attribute__((unused)) void bf16_stream_terms_kernel(
    const double *__restrict__ n2, int64_t n, double b, float *__restrict__ xnp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xnp[i] = (float)((1.0 - b) * n2[i]);
}

// Per resident row: qoff = (1-b)|q|^2 - c|q| Smax - eta (f64); *smax is the
// largest streamed-row norm; eta = (d+1) 2^-100 (1 + |q| + Smax) covers bf16
// and f32 flushing of values below 2^-126.
static __global__ __attribute__((unused)) void bf16_resident_terms_kernel(
    const double *__restrict__ n2, const double *__restrict__ n1, int64_t n, int d, double c,
    double b, const double *__restrict__ smax, double *__restrict__ qoff) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sm = *smax;
    const double eta = (double)(d + 1) * 0x1.0p-100 * (1.0 + n1[i] + sm);
    qoff[i] = (1.0 - b) * n2[i] - c * n1[i] * sm - eta;
}

// *flag = 1 if any value is non-finite or larger than 2^56 in magnitude (the
// BF16 filter's bound assumes neither; such data uses the f64 filter).
template <typename T>
__global__ void bf16_unsafe_kernel(const T *__restrict__ X, int64_t count,
                                   unsigned int *__restrict__ flag) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)X[i];
        bad = bad || !(fabs(v) <= 0x1.0p56);
    }
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// *out = max(*out, max_i v[i]) for non-negative v; *out must start at 0.
static __global__ __attribute__((unused)) void max_nonneg_kernel(const double *__restrict__ v,
                                                                 int64_t n,
                                                                 unsigned long long *__restrict__ out) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, v[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

struct Bf16Args {
    int64_t nA, nB;          // resident rows, streamed rows
    int64_t b_per_split;
    const unsigned short *ah, *al;   // resident hi/lo [nA][dpad]
    const unsigned short *bh, *bl;   // streamed hi/lo [nB][dpad]
    const double *qoff;              // resident (1-b)|q|^2 - c|q|Smax
    const float *xnp;                // streamed (1-b)|x|^2
    Cand out;
};

template <int NKC>
struct Bf16Shared {
    static constexpr int ROWB = NKC * 64 + 16;   // bytes per staged row (hi or lo) + pad
    unsigned char tile[2][BF_T][ROWB];           // [hi/lo][row][bytes]
    float xnp[BF_T];
    float kbuf[BF_WAVES][64][33];                // a lane's 32 keys on the slow path
};

__device__ __forceinline__ bool pair_less_f(float ka, int32_t ia, float kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Sorted insert of (k, i) into a register list, dropping the last entry.
template <int MP>
__device__ __forceinline__ void plist_insert(float (&lk)[MP], int32_t (&li)[MP], float k,
                                             int32_t i) {
#pragma unroll
    for (int j = 0; j < MP; ++j) {
        const bool lt = pair_less_f(k, i, lk[j], li[j]);
        const float tk = lt ? lk[j] : k;
        const int32_t ti = lt ? li[j] : i;
        lk[j] = lt ? k : lk[j];
        li[j] = lt ? i : li[j];
        k = tk;
        i = ti;
    }
}

// Shared filter threshold of the four quarter lists of one resident row
// (lanes l, l^16, l^32, l^48):  T = min(max_i list_i[MP/4 - 1],
// min_i list_i[MP - 1]).  The four lists hold MP keys <= the first term, so
// it bounds the union's MP-th key from above; each quarter's MP-th key does
// too.  Both terms only decrease as lists fill, so T is monotone.
template <int MP>
__device__ __forceinline__ float shared_threshold(const float (&lk)[MP]) {
    float a = lk[MP / 4 - 1];
    a = fmaxf(a, __shfl_xor(a, 16));
    a = fmaxf(a, __shfl_xor(a, 32));
    float b = lk[MP - 1];
    b = fminf(b, __shfl_xor(b, 16));
    b = fminf(b, __shfl_xor(b, 32));
    return fminf(a, b);
}

template <int MP, int NKC, int PASSES>
__global__ __launch_bounds__(256) void dist_bf16_kernel(Bf16Args args) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef Bf16Shared<NKC> Sh;
    Sh &sh = *reinterpret_cast<Sh *>(smem_raw);
    constexpr int DPAD = NKC * 32;

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int g = lane >> 4;   // quarter of the streamed rows seen by this lane
    const int64_t nA = args.nA;
    const int64_t q0 = (int64_t)blockIdx.x * BF_Q + w * BF_QW;   // this wave's resident rows
    const int split = blockIdx.y;
    const int64_t bbeg = (int64_t)split * args.b_per_split;
    int64_t bend = bbeg + args.b_per_split;
    if (bend > args.nB) bend = args.nB;

    // Resident operand fragments: for column tile jt, lane l holds row
    // q0 + 16*jt + l%16, features kc*32 + 8*(l/16) .. +7.
    bf16x8_t qh[2][NKC], ql[2][NKC];
    bool qok[2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
        const int64_t q = q0 + 16 * jt + (lane & 15);
        qok[jt] = q < nA;
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            u16x8_t vh = (u16x8_t){0, 0, 0, 0, 0, 0, 0, 0}, vl = vh;
            if (qok[jt]) {
                const int64_t off = q * DPAD + kc * 32 + 8 * g;
                vh = *reinterpret_cast<const u16x8_t *>(args.ah + off);
                if (PASSES == 3) vl = *reinterpret_cast<const u16x8_t *>(args.al + off);
            }
            qh[jt][kc] = __builtin_bit_cast(bf16x8_t, vh);
            ql[jt][kc] = __builtin_bit_cast(bf16x8_t, vl);
        }
    }

    // Private candidate lists, ascending (key, index).
    float pk[2][MP];
    int32_t pi[2][MP];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int j = 0; j < MP; ++j) {
            pk[jt][j] = __builtin_huge_valf();
            pi[jt][j] = 0x7fffffff;
        }

    // Streamed tiles: thread t owns LOADS 16-byte chunks of each tile; the
    // next tile's chunks are fetched into registers while the current tile
    // is multiplied, then written to LDS after a barrier.
    constexpr int CH = DPAD / 8;                  // 16-byte chunks per row
    constexpr int CHUNKS = (PASSES == 3 ? 2 : 1) * BF_T * CH;
    constexpr int LOADS = (CHUNKS + 255) / 256;
    u16x8_t pf[LOADS];
    float pf_xn = 0.f;
    auto fetch = [&](int64_t t0) {
        // Branch-free: rows past the split end read the last valid row; the
        // values are zeroed when they are written to LDS, so that nothing
        // here consumes a load result (which would force a wait on it).
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const int t = threadIdx.x + 256 * u;
            const int tt = t < CHUNKS ? t : 0;
            const int part = tt / (BF_T * CH);
            const int rem = tt % (BF_T * CH);
            const int r = rem / CH, ch = rem % CH;
            const int64_t gb = t0 + r;
            const int64_t gc = gb < bend ? gb : bend - 1;
            const unsigned short *src = part ? args.bl : args.bh;
            pf[u] = *reinterpret_cast<const u16x8_t *>(src + gc * DPAD + ch * 8);
        }
        if (threadIdx.x < BF_T) {
            const int64_t gb = t0 + threadIdx.x;
            const int64_t gc = gb < bend ? gb : bend - 1;
            pf_xn = args.xnp[gc];
        }
    };
    wait_vm_loads();   // resident fragments are in registers before the loop
    if (bbeg < bend) fetch(bbeg);

    for (int64_t b0 = bbeg; b0 < bend; b0 += BF_T) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const int t = threadIdx.x + 256 * u;
            if (t < CHUNKS) {
                // Rows past the split end hold copies of the last row; their
                // keys are made +inf through xn' below.
                const int part = t / (BF_T * CH);
                const int rem = t % (BF_T * CH);
                *reinterpret_cast<u16x8_t *>(&sh.tile[part][rem / CH][(rem % CH) * 16]) = pf[u];
            }
        }
        if (threadIdx.x < BF_T) {
            // Rows past the split end get an infinite key: they never pass.
            const bool ok = b0 + threadIdx.x < bend;
            sh.xnp[threadIdx.x] = ok ? pf_xn : __builtin_huge_valf();
        }
        __syncthreads();
        if (b0 + BF_T < bend) fetch(b0 + BF_T);

        float4_t acc[4][2];
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) acc[it][jt] = (float4_t){0.f, 0.f, 0.f, 0.f};
        // For each feature chunk: read the four 16-row fragments (hi, lo),
        // then issue the products pass by pass so that eight independent
        // MFMAs separate two updates of one accumulator.
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            bf16x8_t xh[4], xl[4];
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const int r = 16 * it + (lane & 15);
                const int boff = kc * 64 + 16 * g;
                xh[it] = __builtin_bit_cast(bf16x8_t,
                                            *reinterpret_cast<const u16x8_t *>(&sh.tile[0][r][boff]));
                if (PASSES == 3)
                    xl[it] = __builtin_bit_cast(
                        bf16x8_t, *reinterpret_cast<const u16x8_t *>(&sh.tile[1][r][boff]));
            }
#pragma unroll
            for (int it = 0; it < 4; ++it)
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
                    acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[it], qh[jt][kc],
                                                                          acc[it][jt], 0, 0, 0);
            if (PASSES == 3) {
#pragma unroll
                for (int it = 0; it < 4; ++it)
#pragma unroll
                    for (int jt = 0; jt < 2; ++jt)
                        acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            xh[it], ql[jt][kc], acc[it][jt], 0, 0, 0);
#pragma unroll
                for (int it = 0; it < 4; ++it)
#pragma unroll
                    for (int jt = 0; jt < 2; ++jt)
                        acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            xl[it], qh[jt][kc], acc[it][jt], 0, 0, 0);
            }
        }

        // ---- epilogue ----
        // Fast path: one FMA and one compare per pair against the lane's own
        // list threshold; bit it*8 + v*2 + jt of pm marks a pair to insert.
        // Rows past the end have xn' = +inf; resident rows past nA never pass.
        float th[2];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) th[jt] = qok[jt] ? shared_threshold<MP>(pk[jt]) : -__builtin_huge_valf();
        float key[4][4][2];
        uint32_t pm = 0;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const float4_t xr =
                *reinterpret_cast<const float4_t *>(&sh.xnp[16 * it + bf_acc_row(lane, 0)]);
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int jt = 0; jt < 2; ++jt) {
                    const float kv = __builtin_fmaf(-2.f, acc[it][jt][v], xr[v]);
                    key[it][v][jt] = kv;
                    pm |= (kv < th[jt] ? 1u : 0u) << (it * 8 + v * 2 + jt);
                }
        }
        if (__ballot(pm != 0) == 0) continue;

        // Slow path: park the keys in this lane's LDS row, then insert the
        // marked pairs one at a time (a per-lane loop; the lists are private).
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                sh.kbuf[w][lane][it * 8 + v * 2] = key[it][v][0];
                sh.kbuf[w][lane][it * 8 + v * 2 + 1] = key[it][v][1];
            }
        wave_sync_lds();
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            uint32_t rem = pm & (jt ? 0xAAAAAAAAu : 0x55555555u);
            while (rem) {
                const int b = __builtin_ctz(rem);
                rem &= rem - 1u;
                const float kv = sh.kbuf[w][lane][b];
                const int32_t ix =
                    (int32_t)(b0 + 16 * (b >> 3) + bf_acc_row(lane, (b >> 1) & 3));
                if (pair_less_f(kv, ix, pk[jt][MP - 1], pi[jt][MP - 1]))
                    plist_insert<MP>(pk[jt], pi[jt], kv, ix);
            }
        }
    }

    // ---- output: the MP smallest pairs of the four quarter lists ----
    // The bound of every pair not output is the MP-th key of the union (the
    // lists hold at least MP keys <= T, so it is <= T); with fewer than MP
    // pairs in the union nothing was ever dropped and the bound is +inf.
    // Inputs are finite and bounded (the host routes other data to the f64
    // filter), so every streamed row got a finite key.
    __syncthreads();
    float *okey = &sh.kbuf[w][0][0];                              // [16][4*MP]
    int32_t *oidx = reinterpret_cast<int32_t *>(okey + 16 * 4 * MP);
    const int c16 = lane & 15;
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
        wave_sync_lds();
#pragma unroll
        for (int j = 0; j < MP; ++j) {
            okey[c16 * 4 * MP + g * MP + j] = pk[jt][j];
            oidx[c16 * 4 * MP + g * MP + j] = pi[jt][j];
        }
        wave_sync_lds();
        const int64_t q = q0 + 16 * jt + c16;
        int nv = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j) nv += pi[jt][j] != 0x7fffffff ? 1 : 0;
        nv += __shfl_xor(nv, 16);
        nv += __shfl_xor(nv, 32);
        if (q >= nA) continue;
        const int64_t o = (int64_t)split * nA + q;
        const double qo = args.qoff[q];
#pragma unroll
        for (int j = 0; j < MP; ++j) {
            if (pi[jt][j] == 0x7fffffff) continue;
            int r = 0;
            for (int e = 0; e < 4 * MP; ++e)
                r += pair_less_f(okey[c16 * 4 * MP + e], oidx[c16 * 4 * MP + e], pk[jt][j],
                                 pi[jt][j]) ? 1 : 0;
            if (r < MP) {
                args.out.key[o * MP + r] = (double)pk[jt][j] + qo;
                args.out.idx[o * MP + r] = pi[jt][j];
            }
            if (r == MP - 1) {
                // key' + qoff, rounded down by a few f32 ulps of each part so
                // that the bound stays below every key that was not output.
                const double kf = (double)pk[jt][j];
                args.out.theta[o] = (kf + qo) - (fabs(kf) + fabs(qo)) * 0x1.0p-22;
            }
        }
        if (g == 0) {
            args.out.cnt[o] = nv < MP ? nv : MP;
            args.out.theta_idx[o] = 0x7fffffff;
            if (nv < MP) args.out.theta[o] = __builtin_huge_val();
            for (int r = nv; r < MP; ++r) {
                args.out.key[o * MP + r] = __builtin_huge_val();
                args.out.idx[o * MP + r] = -1;
            }
        }
    }
}

}  // namespace fl
