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
// operand, staged through LDS 64 rows at a time by direct-to-LDS buffer
// loads (double buffered, shared by the 8 waves of a block); the resident
// rows (kNN queries / KMeans samples) are the B operand and stay in VGPRs for
// the whole kernel, 32 per wave.
//
// Candidate lists live in registers: each of the four lanes that share a
// resident row keeps a sorted list of MP/2 pairs of its quarter of the
// streamed rows, filtered against a threshold shared by the four lanes; the
// output is the MP smallest pairs of their union with a bound theta on every
// pair that was dropped (see dist_bf16_kernel).
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

#include <type_traits>

namespace fl {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));

constexpr int BF_WAVES = 8;
constexpr int BF_QW = 32;                 // resident rows per wave (two 16-col tiles)
constexpr int BF_Q = BF_WAVES * BF_QW;    // resident rows per block
constexpr int BF_T = 64;                  // streamed rows per tile (four 16-row tiles)

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
    const double *qoff;              // resident (1-b)|q|^2 - c|q|Smax - eta
    const float *xnp;                // streamed (1-b)|x|^2
    Cand out;
    // Argmin form with the decision fused (KMeans assignment, dec_X set):
    // the kernel folds the smallest-key candidate exactly against the
    // samples dec_X ([nA][d], f32 if dec_x_f32 else f64) and the f64 centres
    // dec_C, certifies it, and writes labels / best, or appends the row to
    // flags (count in *fcount) for the exact scan.
    const void *dec_X = nullptr;
    int dec_x_f32 = 0;
    int d = 0;
    const double *dec_C = nullptr;
    int32_t *labels = nullptr;
    double *best = nullptr;
    int32_t *flags = nullptr;
    int32_t *fcount = nullptr;
    // FP8 form (PASSES == 0): per resident row key factor -2 s_C s_x (f32,
    // indexed by source row); ah / bh hold E4M3 bytes, 128 per row.
    const float *afac = nullptr;
    // Resident rows gathered from a list (argmin form): launch row q is
    // source row rows[q], for q < *rows_count (device); qoff, afac, the
    // split rows, labels / best and flags use source rows.
    const int32_t *rows = nullptr;
    const int32_t *rows_count = nullptr;
};

// FP8 E4M3 (OCP: bias 7, largest finite 448, subnormal spacing 2^-9) of a
// value with |v| <= 448, rounded to nearest even.
__host__ __device__ __forceinline__ uint8_t fp8_e4m3_rne(double v) {
    const uint8_t sgn = v < 0.0 ? 0x80 : 0x00;
    const double a = fabs(v);
    if (a == 0.0) return sgn;
    int e = 0;
    (void)frexp(a, &e);   // a in [2^(e-1), 2^e)
    int E = e - 1;
    if (E < -6) {
        const int m = (int)rint(a * 512.0);   // multiples of 2^-9, exact scaling
        return m >= 8 ? (uint8_t)(sgn | 0x08) : (uint8_t)(sgn | m);
    }
    int m = (int)rint(ldexp(a, 3 - E));       // in [8, 16]
    if (m == 16) {
        m = 8;
        ++E;
    }
    return (uint8_t)(sgn | ((E + 7) << 3) | (m - 8));
}

// Power-of-two scale exponent of a row whose largest magnitude is mx: the
// smallest E with mx / 2^E < 448, kept in [-60, 60] so that products of two
// scales stay normal in f32.
__host__ __device__ __forceinline__ int fp8_scale_exp(double mx) {
    if (!(mx > 0.0)) return 0;
    int e = 0;
    (void)frexp(mx / 448.0, &e);   // mx/448 < 2^e
    return e < -60 ? -60 : (e > 60 ? 60 : e);
}

// Rows to E4M3 bytes (128 per row, zero padded) with per-row scale
// exponents (d <= 128): one wave per row, lane j converts elements 2j and
// 2j+1 after a wave-wide max of the row's magnitudes.  Grid: ceil(n / 4)
// blocks of 256.
template <typename T>
__global__ __launch_bounds__(256) void fp8_split_rows_kernel(const T *__restrict__ X, int64_t n,
                                                             int d, uint8_t *__restrict__ out,
                                                             int32_t *__restrict__ sexp) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;   // the whole wave
    const T *x = X + r * d;
    const int j0 = 2 * lane;
    const double a = j0 < d ? (double)x[j0] : 0.0;
    const double b = j0 + 1 < d ? (double)x[j0 + 1] : 0.0;
    double mx = fmax(fabs(a), fabs(b));
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    const int E = fp8_scale_exp(mx);
    const uint16_t v = (uint16_t)(fp8_e4m3_rne(ldexp(a, -E)) |
                                  ((uint16_t)fp8_e4m3_rne(ldexp(b, -E)) << 8));
    reinterpret_cast<uint16_t *>(out + r * 128)[lane] = v;
    if (lane == 0) sexp[r] = E;
}

// Rows to E4M3 bytes with one scale exponent for all of them (*sexp).
static __global__ __attribute__((unused)) void fp8_split_common_kernel(
    const double *__restrict__ X, int64_t n, int d, const int32_t *__restrict__ sexp,
    uint8_t *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 128) return;
    const int64_t r = t / 128;
    const int j = (int)(t % 128);
    out[t] = j < d ? fp8_e4m3_rne(ldexp(X[r * d + j], -*sexp)) : 0;
}

// *sexp = fp8_scale_exp of the largest magnitude among v[0..count), read
// from *mxbits (max of the non-negative magnitudes' bit patterns).
static __global__ __attribute__((unused)) void fp8_common_exp_kernel(
    const unsigned long long *__restrict__ mxbits, int32_t *__restrict__ sexp) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *sexp = fp8_scale_exp(__longlong_as_double((long long)*mxbits));
}

// *out = max(*out, bits of max_i |v[i]|) (zeroed before the launch).
static __global__ __attribute__((unused)) void max_abs_bits_kernel(
    const double *__restrict__ v, int64_t n, unsigned long long *__restrict__ out) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Per resident row of the FP8 form: qoff = (1-b)|q|^2 - c|q| Smax - eta with
// eta = (d+1) 2^-60 (1 + |q| + Smax): scale exponents are kept >= -60, so
// elements far below a row's largest one (flushed or subnormal in E4M3 at
// that scale) are off by at most 2^-70 each, and f32 flushing is covered too.
static __global__ __attribute__((unused)) void fp8_resident_terms_kernel(
    const double *__restrict__ n2, const double *__restrict__ n1, int64_t n, int d, double c,
    double b, const double *__restrict__ smax, double *__restrict__ qoff) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sm = *smax;
    const double eta = (double)(d + 1) * 0x1.0p-60 * (1.0 + n1[i] + sm);
    qoff[i] = (1.0 - b) * n2[i] - c * n1[i] * sm - eta;
}

// Per resident row: afac = -2 * 2^(eC + ex) (exact in f32).
static __global__ __attribute__((unused)) void fp8_resident_factor_kernel(
    const int32_t *__restrict__ sexp_rows, const int32_t *__restrict__ sexp_c, int64_t n,
    float *__restrict__ afac) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) afac[i] = ldexpf(-2.0f, sexp_rows[i] + *sexp_c);
}

// Raw buffer resource over [base, base + bytes): loads at offsets past the
// end return zero (the hardware range check), so a tile that runs past the
// end of its split needs no clamping.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bf_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// LDS layout of one staged tile: [part (hi, lo)][row 0..63][16-byte chunk],
// rows of NKC*64 bytes with the chunk order XOR-swizzled by the row
// (position = chunk ^ (row & SWZ)), so that the 16 rows read by one MFMA
// fragment load fall in distinct banks.  NST tiles are resident: one being
// multiplied while the next NST-1 are copied in by direct-to-LDS buffer
// loads.  Each wave also copies the tile's 64 streamed-row terms (xn') into
// a slot of its own.
template <int NKC, int PASSES>
struct Bf16Tile {
    static constexpr int CH = NKC * 4;                    // 16-byte chunks per row
    static constexpr int ROWB = CH * 16;
    static constexpr int SWZ = (CH < 16 ? CH : 16) - 1;
    static constexpr int PARTS = PASSES == 3 ? 2 : 1;
    static constexpr int BYTES = PARTS * BF_T * ROWB;
    static constexpr int NBLK = BYTES / 1024;             // 1 KiB copy blocks per tile
    static constexpr int BPW = (NBLK + BF_WAVES - 1) / BF_WAVES;
    static constexpr int XS = BF_WAVES * BF_T * 4;        // xn' slots per tile
    static constexpr int NST = 3 * (BYTES + XS) + 1024 <= 160 * 1024 ? 3 : 2;
    static constexpr int OPS = BPW + 1;                   // copies per wave per tile
    static constexpr size_t SHMEM = (size_t)NST * (BYTES + XS) + (NBLK % BF_WAVES ? 1024 : 0);
};

__device__ __forceinline__ bool pair_less_f(float ka, int32_t ia, float kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Sorted insert of (k, i) into a register list, dropping the last entry.
// A lane sees its pairs in increasing index order, so an equal key goes
// after the entries already there and the list stays sorted by (key, index).
// Every position is computed from the old list (no dependency chain).
template <int L>
__device__ __forceinline__ void plist_insert(float (&lk)[L], int32_t (&li)[L], float k,
                                             int32_t i) {
    bool lt[L];
#pragma unroll
    for (int j = 0; j < L; ++j) lt[j] = k < lk[j];
#pragma unroll
    for (int j = L - 1; j > 0; --j) {
        const float sk = lt[j - 1] ? lk[j - 1] : k;
        const int32_t si = lt[j - 1] ? li[j - 1] : i;
        lk[j] = lt[j] ? sk : lk[j];
        li[j] = lt[j] ? si : li[j];
    }
    lk[0] = lt[0] ? k : lk[0];
    li[0] = lt[0] ? i : li[0];
}

// Block barrier that orders LDS accesses only.  Copies into LDS that are
// still in flight (vmcnt) are waited for explicitly where needed, so that
// the copies of later tiles can stay in flight across it.
__device__ __forceinline__ void block_sync_lds() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Max / min over lanes l and l^16, then l and l^32, with the row swaps of
// v_permlane16_swap / v_permlane32_swap (no LDS round trip).  With both
// operands x, the two results are x with its rows exchanged pairwise; the
// second operand is passed through an opaque copy so that the compiler does
// not treat the two results as equal.
__device__ __forceinline__ unsigned opaque_copy(unsigned x) {
    unsigned y;
    asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
    return y;
}
template <bool MAX, int ROWS>
__device__ __forceinline__ float xred(float x) {
    const unsigned u = __builtin_bit_cast(unsigned, x);
    unsigned a, b;
    if (ROWS == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(u, opaque_copy(u), false, false);
        a = p[0];
        b = p[1];
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(u, opaque_copy(u), false, false);
        a = p[0];
        b = p[1];
    }
    const float fa = __builtin_bit_cast(float, a), fb = __builtin_bit_cast(float, b);
    return MAX ? fmaxf(fa, fb) : fminf(fa, fb);
}

// Filter threshold shared by the four lists of one resident row (lanes l,
// l^16, l^32, l^48), each of length L = MP/2:
//   T = min(max_i list_i[MP/4 - 1], min_i list_i[L - 1]).
// When the first term is finite the four lists hold at least MP keys <= it;
// the second is the smallest key any list could still evict.  Both terms
// only decrease.
template <int MP, int L>
__device__ __forceinline__ float shared_threshold(const float (&lk)[L]) {
    return fminf(xred<true, 32>(xred<true, 16>(lk[MP >= 4 ? MP / 4 - 1 : 0])),
                 xred<false, 32>(xred<false, 16>(lk[L - 1])));
}

// One block = 8 waves x 32 resident rows, all sharing each staged tile of
// 64 streamed rows.  Grid: (ceil(nA / 256), S splits of the streamed rows).
//
// Candidate lists: lane l's accumulators hold resident rows 16*jt + l%16
// (jt = 0, 1) against streamed rows 16*it + 4*(l/16) + v of each tile, so
// the four lanes of a resident row see disjoint quarters of the streamed
// rows.  Each lane keeps a sorted list of L = MP/2 pairs per resident row.
// A pair is dropped only when (a) its key is >= T at the start of its tile,
// (b) it is >= its lane's last entry, or (c) it is evicted (then it is >=
// the new last entry).  T and the last entries only decrease, so every
// dropped pair is >= min(T_end, min_i last_i) >= theta, where
//   theta = min(MP-th key of the union of the four lists, min_i last_i)
// (T_end >= the MP-th key of the union whenever T_end is finite).  The MP
// smallest pairs of the union are output with that theta; pairs of the
// union that are not output are >= its MP-th key.
template <int MP, int NKC, int PASSES>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu((MP == 0 && NKC <= 2) ? 4 : 1)))
void dist_bf16_kernel(Bf16Args args) {
    // The body is compiled for the device only: the host pass needs just the
    // launch stub, and it does not accept the buffer-to-LDS builtin with
    // per-lane offsets.
#if defined(__HIP_DEVICE_COMPILE__)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef Bf16Tile<NKC, PASSES> TL;
    // MP == 0: argmin form (KMeans assignment): each lane keeps its best pair
    // per resident row and the smallest key it dropped; the four lanes of a
    // row output their four bests and theta = the smallest dropped key.
    constexpr bool ARGMIN = MP == 0;
    // PASSES == 0: FP8 E4M3 operands (128 bytes per row, the layout of NKC = 2
    // bf16 rows), one v_mfma_scale_f32_16x16x128_f8f6f4 per 16 x 16 tile with
    // unit scales, key = fma(afac_row, acc, xn'); argmin form only.
    constexpr bool FP8 = PASSES == 0;
    static_assert(!FP8 || (ARGMIN && NKC == 2), "FP8 form: argmin, 128-byte rows");
    constexpr int L = ARGMIN ? 1 : MP / 2;
    constexpr int DPAD = NKC * 32;
    float *xs = reinterpret_cast<float *>(smem_raw + TL::NST * TL::BYTES);   // [NST][waves][BF_T]

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int g = lane >> 4;
    const int c16 = lane & 15;
    int64_t nA = args.nA;
    if (args.rows_count) {
        const int64_t c = *args.rows_count;
        nA = c < nA ? c : nA;
        if ((int64_t)blockIdx.x * BF_Q >= nA) return;   // the whole block is past the list
    }
    const int64_t q0 = (int64_t)blockIdx.x * BF_Q + w * BF_QW;
    const int split = blockIdx.y;
    const int64_t bbeg = (int64_t)split * args.b_per_split;
    int64_t bend = bbeg + args.b_per_split;
    if (bend > args.nB) bend = args.nB;

    // Resident operand fragments: for column tile jt, lane l holds row
    // q0 + 16*jt + l%16, features kc*32 + 8*(l/16) .. +7.
    bf16x8_t qh[2][NKC], ql[2][NKC];
    i32x8_t q8[2];
    bool qok[2];
    int64_t qsrc[2];   // source row of each of this lane's resident rows
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
        const int64_t q = q0 + 16 * jt + c16;
        qok[jt] = q < nA;
        qsrc[jt] = qok[jt] ? (args.rows ? (int64_t)args.rows[q] : q) : 0;
        if (FP8) {
            // bytes 32 g .. 32 g + 31 of the row: the k slice of lane group g
            i32x8_t v = (i32x8_t){0, 0, 0, 0, 0, 0, 0, 0};
            if (qok[jt])
                v = *reinterpret_cast<const i32x8_t *>(
                    reinterpret_cast<const unsigned char *>(args.ah) + qsrc[jt] * 128 + 32 * g);
            q8[jt] = v;
        }
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            u16x8_t vh = (u16x8_t){0, 0, 0, 0, 0, 0, 0, 0}, vl = vh;
            if (qok[jt] && !FP8) {
                const int64_t off = qsrc[jt] * DPAD + kc * 32 + 8 * g;
                vh = *reinterpret_cast<const u16x8_t *>(args.ah + off);
                if (PASSES == 3) vl = *reinterpret_cast<const u16x8_t *>(args.al + off);
            }
            qh[jt][kc] = __builtin_bit_cast(bf16x8_t, vh);
            ql[jt][kc] = __builtin_bit_cast(bf16x8_t, vl);
        }
    }
    float fx[2] = {-2.f, -2.f};   // key factor per resident row
    if (FP8)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) fx[jt] = qok[jt] ? args.afac[qsrc[jt]] : -2.f;

    float lk[2][L];
    int32_t li[2][L];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
        for (int j = 0; j < L; ++j) {
            lk[jt][j] = __builtin_huge_valf();
            li[jt][j] = 0x7fffffff;
        }
    float th[2];
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) th[jt] = qok[jt] ? __builtin_huge_valf() : -__builtin_huge_valf();
    float bk[2] = {__builtin_huge_valf(), __builtin_huge_valf()};   // argmin form
    float sec[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
    int32_t bi[2] = {0x7fffffff, 0x7fffffff};

    // Copy plan: 1 KiB block b of a tile is written by wave b % 8 (every
    // wave issues BPW block copies; a wave with no block of its own when the
    // tile has fewer than 8 blocks copies block b % NBLK into a scratch KiB,
    // so that every wave has the same count of copies in flight); lane l of
    // the block fills LDS slot 64*b + l, i.e. part, row and swizzled position
    // (row, pos), with chunk pos ^ (row & SWZ) of that streamed row.  Each
    // wave then copies the tile's 64 xn' terms into its own slot.
    const uint32_t split_bytes = (uint32_t)((bend - bbeg) * TL::ROWB);
    const __amdgpu_buffer_rsrc_t rh = bf_rsrc(args.bh + bbeg * DPAD, split_bytes);
    const __amdgpu_buffer_rsrc_t rl = bf_rsrc(args.bl + bbeg * DPAD, split_bytes);
    const __amdgpu_buffer_rsrc_t rx = bf_rsrc(args.xnp + bbeg, (uint32_t)((bend - bbeg) * 4));
    uint32_t voff[TL::BPW];
#pragma unroll
    for (int j = 0; j < TL::BPW; ++j) {
        const int s = 64 * ((j * BF_WAVES + w) % TL::NBLK) + lane;
        const int rem = s % (BF_T * TL::CH);
        const int row = rem / TL::CH, pos = rem % TL::CH;
        voff[j] = (uint32_t)(row * TL::ROWB + ((pos ^ (row & TL::SWZ)) * 16));
    }
    auto issue_data = [&](int buf, int64_t t0) {
        const uint32_t soff = (uint32_t)((t0 - bbeg) * TL::ROWB);
#pragma unroll
        for (int j = 0; j < TL::BPW; ++j) {
            const int b = j * BF_WAVES + w;   // a block never straddles the hi/lo parts
            const int bs = b % TL::NBLK;
            unsigned char *dst = b < TL::NBLK ? smem_raw + buf * TL::BYTES + 1024 * b
                                              : smem_raw + TL::NST * (TL::BYTES + TL::XS);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                bs >= TL::NBLK / TL::PARTS ? rl : rh, (__attribute__((address_space(3))) void *)dst,
                16, voff[j], soff, 0, 0);
        }
    };
    auto issue_xn = [&](int buf, int64_t t0) {
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rx, (__attribute__((address_space(3))) void *)(xs + (buf * BF_WAVES + w) * BF_T), 4,
            (uint32_t)lane * 4, (uint32_t)((t0 - bbeg) * 4), 0, 0);
    };

    // Tile t lives in buffer t % NST and is issued NST-1 tiles ahead.  The
    // epilogue is software-pipelined: while the products of tile t are
    // issued, the keys of tile t-1 (kept in the other accumulator set) are
    // filtered, so that the compares overlap the matrix cores; the inserts of
    // tile t-1 follow.  At the end of tile t each wave waits for its own
    // copies of tile t+1 (those of tile t+2 may stay in flight), then the
    // block synchronises.
    const int64_t ntiles = (bend - bbeg + BF_T - 1) / BF_T;
#pragma unroll
    for (int t = 0; t < TL::NST - 1; ++t)
        if (t < ntiles) {
            issue_data(t, bbeg + (int64_t)t * BF_T);
            issue_xn(t, bbeg + (int64_t)t * BF_T);
        }
    if (TL::NST == 3 && ntiles > 1)
        __builtin_amdgcn_s_waitcnt(0x0F70 | (TL::OPS & 15) | ((TL::OPS >> 4) << 14));
    else
        wait_vm_loads();
    block_sync_lds();

    // Products of one tile into acc.
    auto products = [&](float4_t (&acc)[4][2], int buf) {
        const unsigned char *tb = smem_raw + buf * TL::BYTES;
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) acc[it][jt] = (float4_t){0.f, 0.f, 0.f, 0.f};
        if (FP8) {
            // lane group g reads chunks 2g and 2g+1 (bytes 32g..32g+31) of
            // row 16 it + c16, the same k slice as its resident fragment
            const int p0 = (2 * g) ^ (c16 & TL::SWZ), p1 = (2 * g + 1) ^ (c16 & TL::SWZ);
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const unsigned char *rw = tb + (16 * it + c16) * TL::ROWB;
                const i32x4_t lo = *reinterpret_cast<const i32x4_t *>(rw + p0 * 16);
                const i32x4_t hi = *reinterpret_cast<const i32x4_t *>(rw + p1 * 16);
                const i32x8_t x8 = (i32x8_t){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
                    acc[it][jt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        x8, q8[jt], acc[it][jt], 0, 0, 0, 0, 0, 0);
            }
            return;
        }
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            const int pos = (kc * 4 + g) ^ (c16 & TL::SWZ);
            const unsigned char *fr = tb + c16 * TL::ROWB + pos * 16;
            bf16x8_t xh[4], xl[4];
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                xh[it] = __builtin_bit_cast(
                    bf16x8_t, *reinterpret_cast<const u16x8_t *>(fr + it * 16 * TL::ROWB));
                if (PASSES == 3)
                    xl[it] = __builtin_bit_cast(
                        bf16x8_t, *reinterpret_cast<const u16x8_t *>(
                                      fr + BF_T * TL::ROWB + it * 16 * TL::ROWB));
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
    };

    // Key of a pair: one FMA, with xn' read from this wave's slot of the
    // tile's buffer.  Streamed rows past the split end (only in the last
    // tile; copied as zeros) get +inf.  Resident rows past nA have T = -inf.
    auto load_xr = [&](float4_t (&xr)[4], int b) {
        const float *xt = xs + (b * BF_WAVES + w) * BF_T;
#pragma unroll
        for (int it = 0; it < 4; ++it)
            xr[it] = *reinterpret_cast<const float4_t *>(&xt[16 * it + 4 * g]);
    };
    auto keyof = [&](const float4_t (&acc)[4][2], const float4_t (&xr)[4], int it, int v, int jt,
                     int nvalid) {
        const float k = __builtin_fmaf(-2.f, acc[it][jt][v], xr[it][v]);
        return (nvalid < BF_T && 16 * it + 4 * g + v >= nvalid) ? __builtin_huge_valf() : k;
    };
    // Filter of one tile: gany[jt][it] marks a pair of that group below T.
    auto filter = [&](const float4_t (&acc)[4][2], int b, int nvalid, bool (&gany)[2][4]) {
        float4_t xr[4];
        load_xr(xr, b);
        bool any = false;
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                bool a = false;
#pragma unroll
                for (int v = 0; v < 4; ++v) a = a || keyof(acc, xr, it, v, jt, nvalid) < th[jt];
                gany[jt][it] = a;
                any = any || a;
            }
        return any;
    };
    // Inserts of one tile: a group of four pairs (one it, one jt) is entered
    // only when some lane has a pair of it below T.
    auto inserts = [&](const float4_t (&acc)[4][2], int b, int nvalid, const bool (&gany)[2][4],
                       int64_t b0) {
        float4_t xr[4];
        load_xr(xr, b);
        const int32_t rb = (int32_t)b0 + 4 * g;
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            if (__ballot(gany[jt][0] || gany[jt][1] || gany[jt][2] || gany[jt][3]) == 0) continue;
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                if (!gany[jt][it]) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float kv = keyof(acc, xr, it, v, jt, nvalid);
                    if (kv < th[jt] && kv < lk[jt][L - 1])
                        plist_insert<L>(lk[jt], li[jt], kv, rb + 16 * it + v);
                }
            }
            th[jt] = qok[jt] ? shared_threshold<MP, L>(lk[jt]) : -__builtin_huge_valf();
        }
    };

    // Argmin form: per lane and row the smallest key bk (with its index bi)
    // and the second smallest sec, which is the smallest key dropped so far:
    // bk' = min(bk, k), sec' = med3(bk, sec, k).  A lane sees its pairs in
    // increasing index order, so strict < keeps the lowest index among equal
    // keys (and an equal key becomes sec).  MASK: the last tile, whose rows
    // past the split end get +inf.
    auto argmin_update = [&](const float4_t (&acc)[4][2], int b, int nvalid, int64_t b0,
                             auto mask_tag) {
        constexpr bool MASK = decltype(mask_tag)::value;
        float4_t xr[4];
        load_xr(xr, b);
        const int32_t rb = (int32_t)b0 + 4 * g;
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int it = 0; it < 4; ++it)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float kv = __builtin_fmaf(fx[jt], acc[it][jt][v], xr[it][v]);
                    if (MASK) kv = 16 * it + 4 * g + v < nvalid ? kv : __builtin_huge_valf();
                    const bool lt = kv < bk[jt];
                    sec[jt] = __builtin_amdgcn_fmed3f(bk[jt], sec[jt], kv);
                    bk[jt] = fminf(bk[jt], kv);
                    bi[jt] = lt ? rb + 16 * it + v : bi[jt];
                }
    };
    using NoMask = std::integral_constant<bool, false>;
    using Mask = std::integral_constant<bool, true>;

    float4_t accA[4][2], accB[4][2];
    // One step: products of tile t into cur_acc while tile t-1 (prev_acc,
    // buffer pbuf) is filtered, then the inserts of tile t-1.  The tile data
    // of tile t+NST-1 is copied into the buffer of tile t-1 (read by no one
    // since the last block barrier), its xn' terms only after this wave's
    // own reads of that slot.
    auto step = [&](int64_t t, int buf, float4_t (&cur_acc)[4][2],
                    const float4_t (&prev_acc)[4][2], auto first_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;   // t == 0: no tile t-1
        const int64_t b0 = bbeg + t * BF_T;
        const int pbuf = buf == 0 ? TL::NST - 1 : buf - 1;
        const bool more = t + TL::NST - 1 < ntiles;
        const int64_t tn = b0 + (int64_t)(TL::NST - 1) * BF_T;
        if (more) issue_data(TL::NST == 3 ? pbuf : (buf ^ 1), tn);
        // Tile t-1 is never the last tile here.  At t = 0 there is no tile
        // t-1: the filter runs on unset sums and its result is discarded (no
        // branch, so that filter and products form one scheduling region).
        {
            bool gany[2][4];
            const bool any = filter(prev_acc, pbuf, BF_T, gany) && !FIRST;
            products(cur_acc, buf);
            // Interleave the filter's compares with the matrix products:
            // three VALU instructions after every three MFMAs.
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
            }
            if (__ballot(any) != 0) inserts(prev_acc, pbuf, BF_T, gany, b0 - BF_T);
        }
        if (more) issue_xn(TL::NST == 3 ? pbuf : (buf ^ 1), tn);
        if (t + 1 < ntiles) {
            if (TL::NST == 3 && t + 2 < ntiles)
                __builtin_amdgcn_s_waitcnt(0x0F70 | (TL::OPS & 15) | ((TL::OPS >> 4) << 14));
            else
                wait_vm_loads();
        }
        block_sync_lds();
    };

    int buf = 0;
    int64_t t = 0;
    if constexpr (ARGMIN) {
        // Argmin form: the update is branch-free VALU work that the other
        // waves' products cover, so tiles are simply taken one at a time.
        for (; t < ntiles; ++t) {
            const int64_t b0 = bbeg + t * BF_T;
            const int nb = TL::NST == 3 ? (buf == 0 ? TL::NST - 1 : buf - 1) : (buf ^ 1);
            if (t + TL::NST - 1 < ntiles) {
                issue_data(nb, b0 + (int64_t)(TL::NST - 1) * BF_T);
                issue_xn(nb, b0 + (int64_t)(TL::NST - 1) * BF_T);
            }
            products(accA, buf);
            const int nvalid = (int)(bend - b0 < BF_T ? bend - b0 : BF_T);
            if (nvalid < BF_T)
                argmin_update(accA, buf, nvalid, b0, Mask{});
            else
                argmin_update(accA, buf, BF_T, b0, NoMask{});
            if (t + 1 < ntiles) {
                if (TL::NST == 3 && t + 2 < ntiles)
                    __builtin_amdgcn_s_waitcnt(0x0F70 | (TL::OPS & 15) | ((TL::OPS >> 4) << 14));
                else
                    wait_vm_loads();
            }
            block_sync_lds();
            buf = buf + 1 == TL::NST ? 0 : buf + 1;
        }
    } else if (ntiles > 0) {
        step(0, buf, accA, accB, Mask{});
        buf = buf + 1 == TL::NST ? 0 : buf + 1;
        t = 1;
    }
    if constexpr (!ARGMIN) {
        for (; t + 1 < ntiles; t += 2) {
            step(t, buf, accB, accA, NoMask{});
            buf = buf + 1 == TL::NST ? 0 : buf + 1;
            step(t + 1, buf, accA, accB, NoMask{});
            buf = buf + 1 == TL::NST ? 0 : buf + 1;
        }
        if (t < ntiles) {
            step(t, buf, accB, accA, NoMask{});
            buf = buf + 1 == TL::NST ? 0 : buf + 1;
            t += 1;
        }
    }
    // The last tile's sums are in accA when ntiles is odd, else in accB.
    if (!ARGMIN && ntiles > 0) {
        const int64_t bl0 = bbeg + (ntiles - 1) * BF_T;
        const int nvalid = (int)(bend - bl0);
        const int lbuf = buf == 0 ? TL::NST - 1 : buf - 1;   // buffer of the last tile
        {
            bool gany[2][4];
            if (ntiles & 1) {
                if (__ballot(filter(accA, lbuf, nvalid, gany)) != 0)
                    inserts(accA, lbuf, nvalid, gany, bl0);
            } else {
                if (__ballot(filter(accB, lbuf, nvalid, gany)) != 0)
                    inserts(accB, lbuf, nvalid, gany, bl0);
            }
        }
    }

    if constexpr (ARGMIN) {
        if (args.dec_X) {
            // Decision in place: the smallest (key, index) of the four lanes
            // is folded exactly; every other centre has a key >= the
            // smallest of the other lanes' bests and every lane's dropped
            // keys (theta), so theta > fold certifies the label (an equal
            // key fails it and the row takes the exact scan).
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) {
                const int64_t qi = q0 + 16 * jt + c16;   // launch row
                const int64_t q = qsrc[jt];              // source row
                const float kmin = xred<false, 32>(xred<false, 16>(bk[jt]));
                // indices < 2^24 (k <= 8192) are exact in f32
                const float icand = bk[jt] == kmin ? (float)bi[jt] : 0x1.0p30f;
                const float imf = xred<false, 32>(xred<false, 16>(icand));
                const int32_t imin = imf >= 0x1.0p30f ? -1 : (int32_t)imf;
                const bool mine = bk[jt] == kmin && bi[jt] == imin;
                const float other = mine ? sec[jt] : fminf(sec[jt], bk[jt]);
                const float thk = xred<false, 32>(xred<false, 16>(other));
                if (qi >= nA || g != jt) continue;
                const int d = args.d;
                bool ok = imin >= 0;
                double bf = 0.0;
                if (ok) {
                    bf = args.dec_x_f32
                             ? fold_exact((const float *)args.dec_X + q * d,
                                          args.dec_C + (int64_t)imin * d, d)
                             : fold_exact((const double *)args.dec_X + q * d,
                                          args.dec_C + (int64_t)imin * d, d);
                    if (thk != __builtin_huge_valf()) {
                        const double kf = (double)thk, qo = args.qoff[q];
                        ok = (kf + qo) - (fabs(kf) + fabs(qo)) * 0x1.0p-22 > bf;
                    }
                }
                if (ok) {
                    args.labels[q] = imin;
                    args.best[q] = bf;
                } else {
                    args.flags[atomicAdd(args.fcount, 1)] = (int32_t)q;
                }
            }
            return;
        }
        // Four candidates per resident row (one per lane, valid ones first)
        // and theta = the smallest key any lane dropped.
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int64_t q = q0 + 16 * jt + c16;
            const bool valid = bi[jt] != 0x7fffffff;
            int pos = 0, cnt = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int vj = __shfl(valid ? 1 : 0, c16 + 16 * j);
                pos += j < g ? vj : 0;
                cnt += vj;
            }
            const float thk = xred<false, 32>(xred<false, 16>(sec[jt]));
            if (q >= nA) continue;
            const double qo = args.qoff[q];
            if (valid) {
                args.out.key[q * 4 + pos] = (double)bk[jt] + qo;
                args.out.idx[q * 4 + pos] = bi[jt];
            }
            if (g == 0) {
                args.out.cnt[q] = cnt;
                args.out.theta_idx[q] = 0x7fffffff;
                if (thk == __builtin_huge_valf()) {
                    args.out.theta[q] = __builtin_huge_val();
                } else {
                    const double kf = (double)thk;
                    args.out.theta[q] = (kf + qo) - (fabs(kf) + fabs(qo)) * 0x1.0p-22;
                }
            }
        }
        return;
    }
    // ---- output: the MP smallest pairs of the union of the four lists ----
    float *okey = reinterpret_cast<float *>(smem_raw) + w * (16 * 4 * L * 2);   // [16][4L]
    int32_t *oidx = reinterpret_cast<int32_t *>(okey + 16 * 4 * L);
#pragma unroll
    for (int jt = 0; jt < 2; ++jt) {
        wave_sync_lds();
#pragma unroll
        for (int j = 0; j < L; ++j) {
            okey[c16 * 4 * L + g * L + j] = lk[jt][j];
            oidx[c16 * 4 * L + g * L + j] = li[jt][j];
        }
        wave_sync_lds();
        int nv = 0;
#pragma unroll
        for (int j = 0; j < L; ++j) nv += li[jt][j] != 0x7fffffff ? 1 : 0;
        nv += __shfl_xor(nv, 16);
        nv += __shfl_xor(nv, 32);
        float lastmin = lk[jt][L - 1];
        lastmin = fminf(lastmin, __shfl_xor(lastmin, 16));
        lastmin = fminf(lastmin, __shfl_xor(lastmin, 32));
        const int64_t q = q0 + 16 * jt + c16;
        const bool live = q < nA;
        const int64_t o = (int64_t)split * nA + q;
        const double qo = live ? args.qoff[q] : 0.0;
        float ump = __builtin_huge_valf();
#pragma unroll
        for (int j = 0; j < L; ++j) {
            if (li[jt][j] == 0x7fffffff) continue;
            int r = 0;
            for (int e = 0; e < 4 * L; ++e)
                r += pair_less_f(okey[c16 * 4 * L + e], oidx[c16 * 4 * L + e], lk[jt][j],
                                 li[jt][j]) ? 1 : 0;
            if (live && r < MP) {
                args.out.key[o * MP + r] = (double)lk[jt][j] + qo;
                args.out.idx[o * MP + r] = li[jt][j];
            }
            if (r == MP - 1) ump = lk[jt][j];
        }
        ump = fminf(ump, __shfl_xor(ump, 16));
        ump = fminf(ump, __shfl_xor(ump, 32));
        const float thk = fminf(ump, lastmin);
        if (live && g == 0) {
            const int nout = nv < MP ? nv : MP;
            args.out.cnt[o] = nout;
            args.out.theta_idx[o] = 0x7fffffff;
            if (thk == __builtin_huge_valf()) {
                args.out.theta[o] = __builtin_huge_val();
            } else {
                // key' + qoff, rounded down by a few f32 ulps of each part so
                // that the bound stays below every key that was not output.
                const double kf = (double)thk;
                args.out.theta[o] = (kf + qo) - (fabs(kf) + fabs(qo)) * 0x1.0p-22;
            }
            for (int r = nout; r < MP; ++r) {
                args.out.key[o * MP + r] = __builtin_huge_val();
                args.out.idx[o * MP + r] = -1;
            }
        }
    }
#endif
}

}  // namespace fl
