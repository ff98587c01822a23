// dist_f64.hpp -- the plain FP64 form of the candidate filter: f64 MFMA
// (v_mfma_f64_16x16x4f64) dot products with a fused per-row top-M (kNN) or
// argmin (KMeans), for up to 128 features and up to 24 neighbours.
//
// Same contract and the same structure as dist_bf16.hpp (see there): the
// streamed rows (kNN database / KMeans centres, stored row-major with DP
// features per row, DP = 4*D4 >= d, zero padded) are the MFMA A operand,
// copied into LDS 64 rows at a time by direct-to-LDS buffer loads (NST tiles
// resident, shared by the 8 waves of a block); the resident rows (queries /
// samples, f64) are the B operand and stay in VGPRs, 16 per wave; each of
// the four lanes of a resident row keeps a private list of its quarter of
// the streamed rows (or, in the argmin form, its best pair and smallest
// dropped key).
//
// Key of a pair, in f64: key = fma(-2, acc, xn'_r), xn'_r = (1 - b)|x_r|^2,
// and the per-row constant qoff_q = (1 - b)|q|^2 - c |q| Smax - eta (Smax =
// the largest |x_r|) is added back for the output: key + qoff <= fold, with
// c, b from bound_fp64 (which covers the f64 MFMA dot product, the norms, the
// key's roundings and the exact fold's own error, times 2), and eta =
// (d+2) 2^-1021 (1 + |q| + Smax) covering products below the normal range.
// The host uses this kernel only for finite data with |x| <= 2^250, so no
// key overflows.
#pragma once

#include "common.hpp"
#include "dist_bf16.hpp"

namespace fl {

constexpr int F64_WAVES = 8;
constexpr int F64_QW = 16;                    // resident rows per wave (one 16-col tile)
constexpr int F64_Q = F64_WAVES * F64_QW;     // resident rows per block
constexpr int F64_T = 64;                     // streamed rows per tile

struct F64Args {
    int64_t nA, nB;
    int d;
    int64_t b_per_split;
    const void *A;             // resident rows [nA][d] (TA)
    const void *B;             // streamed rows [nB][DP] (TB), zero padded
    const double *qoff;        // resident (1-b)|q|^2 - c|q|Smax - eta
    const double *xnp;         // streamed (1-b)|x|^2
    Cand out;
};

// LDS layout of one staged tile: [row 0..63][16-byte chunk] of TB values,
// DP values per row, the chunk order XOR-swizzled by the row so that the 16
// rows read by one fragment load fall in distinct banks.
template <int D4, typename TB>
struct F64Tile {
    static constexpr int DP = 4 * D4;
    static constexpr int ROWB = DP * (int)sizeof(TB);
    static constexpr int CH = ROWB / 16;                 // 16-byte chunks per row
    static constexpr int EPC = 16 / (int)sizeof(TB);     // values per chunk
    static constexpr int SWZ = (CH < 16 ? CH : 16) - 1;
    static constexpr int BYTES = F64_T * ROWB;
    static constexpr int NBLK = BYTES / 1024;
    static constexpr int BPW = (NBLK + F64_WAVES - 1) / F64_WAVES;
    static constexpr int XS = F64_WAVES * F64_T * 8;     // xn' slots per tile (f64)
    static constexpr int NST = 3 * (BYTES + XS) + 1024 <= 160 * 1024 ? 3 : 2;
    static constexpr int OPS = BPW + 2;                   // copies per wave per tile
    static constexpr size_t SHMEM = (size_t)NST * (BYTES + XS) + (NBLK % F64_WAVES ? 1024 : 0);
    static_assert(BYTES % 1024 == 0, "tile must be whole KiB blocks");
};

template <int L>
__device__ __forceinline__ void plist_insert_d(double (&lk)[L], int32_t (&li)[L], double k,
                                               int32_t i) {
    bool lt[L];
#pragma unroll
    for (int j = 0; j < L; ++j) lt[j] = k < lk[j];
#pragma unroll
    for (int j = L - 1; j > 0; --j) {
        const double sk = lt[j - 1] ? lk[j - 1] : k;
        const int32_t si = lt[j - 1] ? li[j - 1] : i;
        lk[j] = lt[j] ? sk : lk[j];
        li[j] = lt[j] ? si : li[j];
    }
    lk[0] = lt[0] ? k : lk[0];
    li[0] = lt[0] ? i : li[0];
}

// f64 max / min over lanes l and l^16, then l^32 (the halves of the double
// go through the row swaps separately).
template <bool MAX, int ROWS>
__device__ __forceinline__ double xred_d(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    unsigned alo, blo, ahi, bhi;
    if (ROWS == 16) {
        const auto pl = __builtin_amdgcn_permlane16_swap(lo, opaque_copy(lo), false, false);
        const auto ph = __builtin_amdgcn_permlane16_swap(hi, opaque_copy(hi), false, false);
        alo = pl[0], blo = pl[1], ahi = ph[0], bhi = ph[1];
    } else {
        const auto pl = __builtin_amdgcn_permlane32_swap(lo, opaque_copy(lo), false, false);
        const auto ph = __builtin_amdgcn_permlane32_swap(hi, opaque_copy(hi), false, false);
        alo = pl[0], blo = pl[1], ahi = ph[0], bhi = ph[1];
    }
    const double a = __builtin_bit_cast(double, ((uint64_t)ahi << 32) | alo);
    const double b = __builtin_bit_cast(double, ((uint64_t)bhi << 32) | blo);
    return MAX ? fmax(a, b) : fmin(a, b);
}

template <int MP, int L>
__device__ __forceinline__ double shared_threshold_d(const double (&lk)[L]) {
    return fmin(xred_d<true, 32>(xred_d<true, 16>(lk[MP >= 4 ? MP / 4 - 1 : 0])),
                xred_d<false, 32>(xred_d<false, 16>(lk[L - 1])));
}

// Grid: (ceil(nA / 128), S splits of the streamed rows), 512 threads.
// MP == 0: argmin form (four candidates per resident row, KMeans); else the
// MP smallest pairs per resident row and split (lists of MP/2 per lane).
template <int MP, int D4, typename TA, typename TB>
__global__ __launch_bounds__(512) void dist_f64_kernel(F64Args args) {
#if defined(__HIP_DEVICE_COMPILE__)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef F64Tile<D4, TB> TL;
    constexpr bool ARGMIN = MP == 0;
    constexpr int L = ARGMIN ? 1 : MP / 2;
    double *xs = reinterpret_cast<double *>(smem_raw + TL::NST * TL::BYTES);   // [NST][waves][64]

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4;
    const int c16 = lane & 15;
    const int64_t nA = args.nA;
    const int d = args.d;
    const int64_t q0 = (int64_t)blockIdx.x * F64_Q + w * F64_QW;
    const int split = blockIdx.y;
    const int64_t bbeg = (int64_t)split * args.b_per_split;
    int64_t bend = bbeg + args.b_per_split;
    if (bend > args.nB) bend = args.nB;

    // Resident fragments: lane l holds features 4*kq + l/16 of row q0 + l%16.
    const int64_t q = q0 + c16;
    const bool qok = q < nA;
    double qf[D4];
#pragma unroll
    for (int kq = 0; kq < D4; ++kq) {
        const int f = 4 * kq + g;
        qf[kq] = (qok && f < d) ? (double)reinterpret_cast<const TA *>(args.A)[q * d + f] : 0.0;
    }

    double lk[L];
    int32_t li[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        lk[j] = __builtin_huge_val();
        li[j] = 0x7fffffff;
    }
    double th = qok ? __builtin_huge_val() : -__builtin_huge_val();
    double bk = __builtin_huge_val(), sec = __builtin_huge_val();   // argmin form
    int32_t bi = 0x7fffffff;

    // Copy plan as in dist_bf16_kernel: 1 KiB block b of a tile is written by
    // wave b % 8 (a scratch KiB when the tile has fewer than 8 blocks); lane
    // l of the block fills LDS slot 64*b + l = (row, pos), with chunk
    // pos ^ (row & SWZ) of that streamed row.  Each wave also copies the
    // tile's 64 xn' terms (512 bytes, two dword copies) into its own slot.
    const uint32_t split_bytes = (uint32_t)((bend - bbeg) * TL::ROWB);
    const __amdgpu_buffer_rsrc_t rb = bf_rsrc(
        reinterpret_cast<const unsigned char *>(args.B) + bbeg * TL::ROWB, split_bytes);
    const __amdgpu_buffer_rsrc_t rx = bf_rsrc(args.xnp + bbeg, (uint32_t)((bend - bbeg) * 8));
    uint32_t voff[TL::BPW];
#pragma unroll
    for (int j = 0; j < TL::BPW; ++j) {
        const int s = 64 * ((j * F64_WAVES + w) % TL::NBLK) + lane;
        const int row = s / TL::CH, pos = s % TL::CH;
        voff[j] = (uint32_t)(row * TL::ROWB + ((pos ^ (row & TL::SWZ)) * 16));
    }
    auto issue_data = [&](int buf, int64_t t0) {
        const uint32_t soff = (uint32_t)((t0 - bbeg) * TL::ROWB);
#pragma unroll
        for (int j = 0; j < TL::BPW; ++j) {
            const int b = j * F64_WAVES + w;
            unsigned char *dst = b < TL::NBLK ? smem_raw + buf * TL::BYTES + 1024 * b
                                              : smem_raw + TL::NST * (TL::BYTES + TL::XS);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rb, (__attribute__((address_space(3))) void *)dst, 16, voff[j], soff, 0, 0);
        }
    };
    auto issue_xn = [&](int buf, int64_t t0) {
        unsigned char *dst = reinterpret_cast<unsigned char *>(xs + (buf * F64_WAVES + w) * F64_T);
        const uint32_t soff = (uint32_t)((t0 - bbeg) * 8);
#pragma unroll
        for (int h = 0; h < 2; ++h)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rx, (__attribute__((address_space(3))) void *)(dst + 256 * h), 4,
                (uint32_t)(256 * h + 4 * lane), soff, 0, 0);
    };

    const int64_t ntiles = (bend - bbeg + F64_T - 1) / F64_T;
#pragma unroll
    for (int t = 0; t < TL::NST - 1; ++t)
        if (t < ntiles) {
            issue_data(t, bbeg + (int64_t)t * F64_T);
            issue_xn(t, bbeg + (int64_t)t * F64_T);
        }
    if (TL::NST == 3 && ntiles > 1)
        __builtin_amdgcn_s_waitcnt(0x0F70 | (TL::OPS & 15) | ((TL::OPS >> 4) << 14));
    else
        wait_vm_loads();
    block_sync_lds();

    // Streamed value (row r of the tile, feature f) from the swizzled layout.
    auto tval = [&](const unsigned char *tb, int r, int f) -> double {
        const int c = f / TL::EPC, e = f % TL::EPC;
        const int pos = c ^ (r & TL::SWZ);
        return (double)*reinterpret_cast<const TB *>(tb + r * TL::ROWB + pos * 16 +
                                                     e * (int)sizeof(TB));
    };

    int buf = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        const int64_t b0 = bbeg + t * F64_T;
        const int pbuf = buf == 0 ? TL::NST - 1 : buf - 1;
        const bool more = t + TL::NST - 1 < ntiles;
        const int64_t tn = b0 + (int64_t)(TL::NST - 1) * F64_T;
        if (more) issue_data(TL::NST == 3 ? pbuf : (buf ^ 1), tn);
        const unsigned char *tb = smem_raw + buf * TL::BYTES;

        double4_t acc[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) acc[it] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kq = 0; kq < D4; ++kq) {
            double a[4];
#pragma unroll
            for (int it = 0; it < 4; ++it) a[it] = tval(tb, 16 * it + c16, 4 * kq + g);
#pragma unroll
            for (int it = 0; it < 4; ++it)
                acc[it] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[it], qf[kq], acc[it], 0, 0, 0);
            // Keep the fragment reads of later steps from all being hoisted
            // here (they would need more registers than the wave has).
            if ((kq & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }

        // Epilogue.  acc[it][v] is streamed row 16*it + g + 4*v against this
        // lane's resident row; a lane sees its rows in increasing order.
        const double *xt = xs + (buf * F64_WAVES + w) * F64_T;
        const int nvalid = (int)(bend - b0 < F64_T ? bend - b0 : F64_T);
        double key[4][4];
        bool any = false;
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * it + g + 4 * v;
                double kv = __builtin_fma(-2.0, acc[it][v], xt[r]);
                kv = r < nvalid ? kv : __builtin_huge_val();
                key[it][v] = kv;
                any = any || kv < th;
            }
        const int32_t rb0 = (int32_t)b0 + g;
        if constexpr (ARGMIN) {
#pragma unroll
            for (int it = 0; it < 4; ++it)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double kv = key[it][v];
                    const bool lt = kv < bk;
                    sec = fmin(sec, lt ? bk : kv);
                    bk = lt ? kv : bk;
                    bi = lt ? rb0 + 16 * it + 4 * v : bi;
                }
        } else {
            if (__ballot(any) != 0) {
#pragma unroll
                for (int it = 0; it < 4; ++it)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double kv = key[it][v];
                        if (kv < th && kv < lk[L - 1])
                            plist_insert_d<L>(lk, li, kv, rb0 + 16 * it + 4 * v);
                    }
                th = qok ? shared_threshold_d<MP, L>(lk) : -__builtin_huge_val();
            }
        }
        if (more) issue_xn(TL::NST == 3 ? pbuf : (buf ^ 1), tn);
        if (t + 1 < ntiles) {
            if (TL::NST == 3 && t + 2 < ntiles)
                __builtin_amdgcn_s_waitcnt(0x0F70 | (TL::OPS & 15) | ((TL::OPS >> 4) << 14));
            else
                wait_vm_loads();
        }
        block_sync_lds();
        buf = buf + 1 == TL::NST ? 0 : buf + 1;
    }

    if constexpr (ARGMIN) {
        const bool valid = bi != 0x7fffffff;
        int pos = 0, cnt = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int vj = __shfl(valid ? 1 : 0, c16 + 16 * j);
            pos += j < g ? vj : 0;
            cnt += vj;
        }
        const double thk = xred_d<false, 32>(xred_d<false, 16>(sec));
        if (!qok) return;
        const double qo = args.qoff[q];
        if (valid) {
            args.out.key[q * 4 + pos] = bk + qo;
            args.out.idx[q * 4 + pos] = bi;
        }
        if (g == 0) {
            args.out.cnt[q] = cnt;
            args.out.theta_idx[q] = 0x7fffffff;
            // key + qoff, rounded down by a few ulps of each part.
            args.out.theta[q] = thk == __builtin_huge_val()
                                    ? __builtin_huge_val()
                                    : (thk + qo) - (fabs(thk) + fabs(qo)) * 0x1.0p-50;
        }
        return;
    }

    // ---- output: the MP smallest pairs of the union of the four lists ----
    double *okey = reinterpret_cast<double *>(smem_raw) + w * (16 * 4 * L * 2);   // [16][4L]
    int32_t *oidx = reinterpret_cast<int32_t *>(okey + 16 * 4 * L);
    wave_sync_lds();
#pragma unroll
    for (int j = 0; j < L; ++j) {
        okey[c16 * 4 * L + g * L + j] = lk[j];
        oidx[c16 * 4 * L + g * L + j] = li[j];
    }
    wave_sync_lds();
    int nv = 0;
#pragma unroll
    for (int j = 0; j < L; ++j) nv += li[j] != 0x7fffffff ? 1 : 0;
    nv += __shfl_xor(nv, 16);
    nv += __shfl_xor(nv, 32);
    const double lastmin = xred_d<false, 32>(xred_d<false, 16>(lk[L - 1]));
    const int64_t o = (int64_t)split * nA + q;
    const double qo = qok ? args.qoff[q] : 0.0;
    double ump = __builtin_huge_val();
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if (li[j] == 0x7fffffff) continue;
        int r = 0;
        for (int e = 0; e < 4 * L; ++e)
            r += pair_less(okey[c16 * 4 * L + e], oidx[c16 * 4 * L + e], lk[j], li[j]) ? 1 : 0;
        if (qok && r < MP) {
            args.out.key[o * MP + r] = lk[j] + qo;
            args.out.idx[o * MP + r] = li[j];
        }
        if (r == MP - 1) ump = lk[j];
    }
    ump = xred_d<false, 32>(xred_d<false, 16>(ump));
    const double thk = fmin(ump, lastmin);
    if (qok && g == 0) {
        const int nout = nv < MP ? nv : MP;
        args.out.cnt[o] = nout;
        args.out.theta_idx[o] = 0x7fffffff;
        args.out.theta[o] = thk == __builtin_huge_val()
                                ? __builtin_huge_val()
                                : (thk + qo) - (fabs(thk) + fabs(qo)) * 0x1.0p-50;
        for (int r = nout; r < MP; ++r) {
            args.out.key[o * MP + r] = __builtin_huge_val();
            args.out.idx[o * MP + r] = -1;
        }
    }
#endif
}

// *flags |= 1 if any value is non-finite or above 2^56 in magnitude (the
// BF16 filters' range), |= 2 if non-finite or above 2^250 (the f64 filter's).
template <typename T>
__global__ void range_flags_kernel(const T *__restrict__ X, int64_t count,
                                   unsigned int *__restrict__ flags) {
    bool b16 = false, b64 = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = fabs((double)X[i]);
        b16 = b16 || !(v <= 0x1.0p56);
        b64 = b64 || !(v <= 0x1.0p250);
    }
    const unsigned f = (__ballot(b16) != 0 ? 1u : 0u) | (__ballot(b64) != 0 ? 2u : 0u);
    if (f && (threadIdx.x & 63) == 0) atomicOr(flags, f);
}

// Per resident row for the f64 filter: qoff = (1-b)|q|^2 - c|q|Smax - eta.
static __global__ __attribute__((unused)) void f64_resident_terms_kernel(
    const double *__restrict__ n2, const double *__restrict__ n1, int64_t n, int d, double c,
    double b, const double *__restrict__ smax, double *__restrict__ qoff) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sm = *smax;
    const double eta = (double)(d + 2) * 0x1.0p-1021 * (1.0 + n1[i] + sm);
    qoff[i] = (1.0 - b) * n2[i] - c * n1[i] * sm - eta;
}

// Per streamed row: xn' = (1-b)|x|^2 (f64).
static __global__ __attribute__((unused)) void f64_stream_terms_kernel(
    const double *__restrict__ n2, int64_t n, double b, double *__restrict__ xnp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xnp[i] = (1.0 - b) * n2[i];
}

// Copy rows of d values into rows of dp values, zero padded.
template <typename T>
__global__ void pad_rows_kernel(const T *__restrict__ X, int64_t n, int d, int dp,
                                T *__restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dp) return;
    const int64_t r = t / dp;
    const int j = (int)(t % dp);
    Y[t] = j < d ? X[r * d + j] : (T)0;
}

}  // namespace fl
