// dist_bf16_res.hpp -- BF16x3 KMeans argmin filter with the centres resident
// in LDS, the decision fused.
//
// Same contract as the argmin form of dist_bf16_kernel with PASSES = 3 and
// the decision fused (dec_X set): a.c ~ ah.ch + ah.cl + al.ch on
// v_mfma_f32_16x16x32_bf16, key = fma(-2, acc, xn'_c) in f32, and key + qoff
// is a lower bound of the exact fold.  The layout follows
// dist_fp8_res_kernel: one block per compute unit holds a chunk of the
// centres (hi and lo rows, 16-byte chunks XOR-swizzled by the row) and their
// xn' terms in LDS, and its 16 waves take groups of 32 samples in turn with
// no block barrier, so that one wave's loads and decision overlap the other
// waves' matrix products.  A chunk holds KR = 1024 / NKC centres (128 KiB of
// hi / lo rows); with more centres the block passes over its samples once per
// chunk and carries each sample's state between chunks in global memory:
// the best packed key so far, its centre, and the smallest other key.
//
// Each lane keeps the smallest and second smallest key of its quarter of a
// chunk's centres with the centre's slot packed into the key's low 6
// mantissa bits (as in dist_fp8_res_kernel, under 64 ulp, < 2^-17
// relative), and the tile of the smallest.  After a chunk the four lanes of
// a sample combine: the smallest packed key (the lowest g among equal ones)
// names the chunk's candidate, which replaces the carried one only when
// strictly smaller (an equal key keeps the earlier, lower centre, and then
// counts as "other", so the certificate fails); every other key of the
// chunk and of earlier chunks goes into "other".  After the last chunk the
// candidate is folded exactly and certified against other + qoff, widened
// by 2^-16 relative for the packing and the f32 roundings; rows that fail
// are listed for the exact scan.
#pragma once

#include "dist_bf16.hpp"

namespace fl {

constexpr int BR_WAVES = 16;

// Per-sample state carried between centre chunks.
struct BfResArgs {
    Bf16Args a;           // argmin form, PASSES = 3, rows == nullptr, dec_X set
    float *pk;            // best packed key so far
    float *psec;          // smallest other key so far
    int32_t *pidx;        // centre of pk (-1: none yet)
};

template <int NKC>
struct BfResShape {
    static constexpr int DPAD = NKC * 32;
    static constexpr int ROWB = DPAD * 2;                 // bytes of one part of a centre row
    static constexpr int CH = ROWB / 16;                  // 16-byte chunks per row
    static constexpr int SWZ = (CH < 16 ? CH : 16) - 1;
    static constexpr int KR = 1024 / NKC;                 // centres per LDS chunk
    static constexpr size_t SHMEM = (size_t)KR * ROWB * 2 + (size_t)KR * 4;
};

template <int NKC>
__global__ __launch_bounds__(BR_WAVES * 64)
__attribute__((amdgpu_waves_per_eu(4))) void dist_bf16_res_kernel(BfResArgs ra) {
#if defined(__HIP_DEVICE_COMPILE__)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef BfResShape<NKC> S;
    const Bf16Args &args = ra.a;
    unsigned char *chi = smem_raw;                                        // [KR][ROWB]
    unsigned char *clo = smem_raw + (size_t)S::KR * S::ROWB;              // [KR][ROWB]
    float *xl = reinterpret_cast<float *>(smem_raw + (size_t)2 * S::KR * S::ROWB);   // [KR]
    const int64_t k = args.nB, nA = args.nA;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4;
    const int c16 = lane & 15;
    const int64_t ngroups = ceil_div(nA, 32);
    const int64_t wstep = (int64_t)gridDim.x * BR_WAVES;
    const int nchunks = (int)ceil_div(k, S::KR);
    // -inf in a register the compiler cannot see through, so that
    // med3(x, y, ninf) is not folded back into a canonicalising min
    const float ninf = __builtin_bit_cast(float, opaque_copy(0xff800000u));
    const unsigned char *bh = reinterpret_cast<const unsigned char *>(args.bh);
    const unsigned char *bl = reinterpret_cast<const unsigned char *>(args.bl);

    for (int ch = 0; ch < nchunks; ++ch) {
        const int64_t c0 = (int64_t)ch * S::KR;
        const int kc_n = (int)(k - c0 < S::KR ? k - c0 : S::KR);
        const int kp = (int)ceil_div(kc_n, 64) * 64;
        const bool last = ch + 1 == nchunks;
        // the chunk's centres in: 16-byte chunk p of row r goes to position
        // p ^ (r & SWZ); rows past the chunk are zero, and their xn' terms
        // FLT_MAX (finite keys that lose to every real one)
        __syncthreads();   // the previous chunk's readers are done
        for (int e = threadIdx.x; e < kp * S::CH; e += blockDim.x) {
            const int r = e / S::CH, p = e % S::CH;
            i32x4_t vh = (i32x4_t){0, 0, 0, 0}, vl = vh;
            if (r < kc_n) {
                vh = *reinterpret_cast<const i32x4_t *>(bh + (c0 + r) * S::ROWB + p * 16);
                vl = *reinterpret_cast<const i32x4_t *>(bl + (c0 + r) * S::ROWB + p * 16);
            }
            const int pos = p ^ (r & S::SWZ);
            *reinterpret_cast<i32x4_t *>(chi + r * S::ROWB + pos * 16) = vh;
            *reinterpret_cast<i32x4_t *>(clo + r * S::ROWB + pos * 16) = vl;
        }
        for (int r = threadIdx.x; r < kp; r += blockDim.x)
            xl[r] = r < kc_n ? args.xnp[c0 + r] : 0x1.fffffep127f;
        __syncthreads();
        const int ntiles = kp / 64;

        for (int64_t grp = (int64_t)blockIdx.x * BR_WAVES + w; grp < ngroups; grp += wstep) {
            const int64_t q0 = grp * 32;
            // resident fragments: lane l holds row q0 + 16 jt + l % 16,
            // features kc * 32 + 8 (l / 16) .. +7
            bf16x8_t qh[2][NKC], ql[2][NKC];
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) {
                const int64_t q = q0 + 16 * jt + c16;
#pragma unroll
                for (int kc = 0; kc < NKC; ++kc) {
                    u16x8_t vh = (u16x8_t){0, 0, 0, 0, 0, 0, 0, 0}, vl = vh;
                    if (q < nA) {
                        const int64_t off = q * S::DPAD + kc * 32 + 8 * g;
                        vh = *reinterpret_cast<const u16x8_t *>(args.ah + off);
                        vl = *reinterpret_cast<const u16x8_t *>(args.al + off);
                    }
                    qh[jt][kc] = __builtin_bit_cast(bf16x8_t, vh);
                    ql[jt][kc] = __builtin_bit_cast(bf16x8_t, vl);
                }
            }
            float bk[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
            float sec[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
            int32_t bt[2] = {-1, -1};   // global tile (64 centres) of bk
            for (int t = 0; t < ntiles; ++t) {
                float4_t acc[4][2];
#pragma unroll
                for (int it = 0; it < 4; ++it)
#pragma unroll
                    for (int jt = 0; jt < 2; ++jt) acc[it][jt] = (float4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kc = 0; kc < NKC; ++kc) {
                    const int pos = (kc * 4 + g) ^ (c16 & S::SWZ);
                    bf16x8_t xh[4], xlo[4];
#pragma unroll
                    for (int it = 0; it < 4; ++it) {
                        const int row = 64 * t + 16 * it + c16;
                        xh[it] = __builtin_bit_cast(
                            bf16x8_t, *reinterpret_cast<const u16x8_t *>(chi + row * S::ROWB + pos * 16));
                        xlo[it] = __builtin_bit_cast(
                            bf16x8_t, *reinterpret_cast<const u16x8_t *>(clo + row * S::ROWB + pos * 16));
                    }
#pragma unroll
                    for (int it = 0; it < 4; ++it)
#pragma unroll
                        for (int jt = 0; jt < 2; ++jt) {
                            acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                xh[it], qh[jt][kc], acc[it][jt], 0, 0, 0);
                            acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                xh[it], ql[jt][kc], acc[it][jt], 0, 0, 0);
                            acc[it][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                xlo[it], qh[jt][kc], acc[it][jt], 0, 0, 0);
                        }
                }
                // lane l's accumulators: centres 64 t + 16 it + 4 g + v of
                // the chunk
                const float b0[2] = {bk[0], bk[1]};
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    const float4_t xr =
                        *reinterpret_cast<const float4_t *>(&xl[64 * t + 16 * it + 4 * g]);
#pragma unroll
                    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const float kv = __builtin_fmaf(-2.f, acc[it][jt][v], xr[v]);
                            const float kpk = __builtin_bit_cast(
                                float, (__builtin_bit_cast(uint32_t, kv) & ~63u) | (uint32_t)(16 * it + v));
                            // med3 with -inf is the min, without the
                            // canonicalising max that fminf puts on a
                            // bit-cast operand
                            sec[jt] = __builtin_amdgcn_fmed3f(bk[jt], sec[jt], kpk);
                            bk[jt] = __builtin_amdgcn_fmed3f(bk[jt], kpk, ninf);
                        }
                }
                const int32_t tg = (int32_t)(c0 / 64) + t;
#pragma unroll
                for (int jt = 0; jt < 2; ++jt) bt[jt] = bk[jt] != b0[jt] ? tg : bt[jt];
            }
            // combine the row's four lanes with the carried state
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) {
                const int64_t q = q0 + 16 * jt + c16;
                const float kmin = xred<false, 32>(xred<false, 16>(bk[jt]));
                const uint64_t eq = __ballot(bk[jt] == kmin);
                const uint64_t eqr = (eq >> c16) & 0x0001000100010001ull;   // bit 16 g' per lane g'
                const int gw = eqr ? (__builtin_ctzll(eqr) >> 4) : 4;
                const bool mine = g == gw;
                const uint32_t code = __builtin_bit_cast(uint32_t, bk[jt]) & 63u;
                const int32_t cand = bt[jt] < 0 ? -1
                                                : 64 * bt[jt] + 16 * (int32_t)(code >> 4) + 4 * g +
                                                      (int32_t)(code & 3u);
                const int32_t imin = __shfl(cand, c16 + 16 * (gw & 3));
                const float other = mine ? sec[jt] : fminf(sec[jt], bk[jt]);
                const float thk = xred<false, 32>(xred<false, 16>(other));
                if (q >= nA || g != jt) continue;
                float pk = __builtin_huge_valf(), ps = __builtin_huge_valf();
                int32_t pi = -1;
                if (ch > 0) {
                    pk = ra.pk[q];
                    ps = ra.psec[q];
                    pi = ra.pidx[q];
                }
                const bool newwin = gw < 4 && imin >= 0 && kmin < pk;
                const float wk = newwin ? kmin : pk;
                const int32_t wi = newwin ? imin : pi;
                const float wo = newwin ? fminf(thk, fminf(pk, ps)) : fminf(ps, kmin);
                if (!last) {
                    ra.pk[q] = wk;
                    ra.psec[q] = wo;
                    ra.pidx[q] = wi;
                    continue;
                }
                const int d = args.d;
                bool ok = wi >= 0 && wi < k;
                double bf = 0.0;
                if (ok) {
                    bf = args.dec_x_f32 ? fold_exact((const float *)args.dec_X + q * d,
                                                     args.dec_C + (int64_t)wi * d, d)
                                        : fold_exact((const double *)args.dec_X + q * d,
                                                     args.dec_C + (int64_t)wi * d, d);
                    if (wo != __builtin_huge_valf()) {
                        const double kf = (double)wo, qo = args.qoff[q];
                        ok = (kf + qo) - ((fabs(kf) + fabs(qo)) * 0x1.0p-16 + 0x1.0p-140) > bf;
                    }
                }
                if (ok) {
                    args.labels[q] = wi;
                    args.best[q] = bf;
                } else {
                    args.flags[atomicAdd(args.fcount, 1)] = (int32_t)q;
                }
            }
        }
    }
#endif
}

}  // namespace fl
