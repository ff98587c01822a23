// dist_fp8.hpp -- FP8 E4M3 KMeans argmin filter with every centre resident
// in LDS (k <= 1024, d <= 128), the decision fused.
//
// Same contract as the FP8 form of dist_bf16_kernel (argmin, PASSES == 0):
// a.b on v_mfma_scale_f32_16x16x128_f8f6f4 with unit scales, key =
// fma(afac_row, acc, xn'_c) in f32.  Each lane keeps the smallest and second
// smallest key with the centre's slot packed into the low mantissa bits
// (three VALU operations per key besides the fma, no index compare), and the
// tile of the smallest; the smallest of a row's four lanes is folded exactly
// and certified against every other key, widened by the packing's slack;
// rows that fail are listed for the next filter.
//
// Layout: one block per compute unit holds all k centres (128 bytes each,
// 16-byte chunks XOR-swizzled by row & 7 so that the 16 rows of one MFMA
// fragment load spread over the banks) and their xn' terms in LDS, loaded
// once.  Its 16 waves then take groups of 32 samples (two 16-column MFMA
// tiles) in turn with no further block barrier: every wave runs the full
// centre scan for its group from LDS and decides, so one wave's loads and
// decision overlap the others' matrix products.
#pragma once

#include "dist_bf16.hpp"

namespace fl {

constexpr int FR_WAVES = 16;
constexpr int FR_MAXK = 1024;

// Dynamic LDS of the kernel for k centres.
inline size_t fp8_res_shmem(int64_t k) {
    const int64_t kp = ceil_div(k, 64) * 64;
    return (size_t)kp * 128 + (size_t)kp * 4;
}

static __global__ __launch_bounds__(FR_WAVES * 64)
__attribute__((amdgpu_waves_per_eu(4), unused)) void dist_fp8_res_kernel(Bf16Args args) {
#if defined(__HIP_DEVICE_COMPILE__)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int64_t k = args.nB;
    const int64_t kp = ceil_div(k, 64) * 64;
    unsigned char *cl = smem_raw;                                  // [kp][128], swizzled
    float *xl = reinterpret_cast<float *>(smem_raw + kp * 128);    // [kp]
    const unsigned char *c8 = reinterpret_cast<const unsigned char *>(args.bh);
    // centres in: 16-byte chunk p of row r goes to position p ^ (r & 7);
    // rows past k are zero, and their xn' terms FLT_MAX (finite keys that lose
    // to every real one and never pass the candidate check below)
    for (int64_t e = threadIdx.x; e < kp * 8; e += blockDim.x) {
        const int64_t r = e >> 3;
        const int p = (int)(e & 7);
        i32x4_t v = (i32x4_t){0, 0, 0, 0};
        if (r < k) v = *reinterpret_cast<const i32x4_t *>(c8 + r * 128 + p * 16);
        *reinterpret_cast<i32x4_t *>(cl + r * 128 + ((p ^ (int)(r & 7)) * 16)) = v;
    }
    for (int64_t r = threadIdx.x; r < kp; r += blockDim.x)
        xl[r] = r < k ? args.xnp[r] : 0x1.fffffep127f;
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4;
    const int c16 = lane & 15;
    const int64_t nA = args.nA;
    const int64_t ngroups = ceil_div(nA, 32);
    const int64_t wstep = (int64_t)gridDim.x * FR_WAVES;
    const int ntiles = (int)(kp / 64);
    const int p0 = (2 * g) ^ (c16 & 7), p1 = (2 * g + 1) ^ (c16 & 7);

    for (int64_t grp = (int64_t)blockIdx.x * FR_WAVES + w; grp < ngroups; grp += wstep) {
        const int64_t q0 = grp * 32;
        i32x8_t q8[2];
        bool qok[2];
        float fx[2];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int64_t q = q0 + 16 * jt + c16;
            qok[jt] = q < nA;
            i32x8_t v = (i32x8_t){0, 0, 0, 0, 0, 0, 0, 0};
            if (qok[jt])
                v = *reinterpret_cast<const i32x8_t *>(
                    reinterpret_cast<const unsigned char *>(args.ah) + q * 128 + 32 * g);
            q8[jt] = v;
            fx[jt] = qok[jt] ? args.afac[q] : -2.f;
        }
        // Running (smallest, second smallest) packed key per lane and row,
        // and the tile of the smallest.  A packed key is the f32 key with
        // its low 6 mantissa bits replaced by the centre's slot in the tile
        // as seen from the lane (16 it + v; the lane's own g supplies the
        // rest of the index), which moves it by under 64 ulp (< 2^-17
        // relative) and keeps one lane's keys in one tile distinct.
        float bk[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
        float sec[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
        int32_t bt[2] = {-1, -1};
        // -inf in a register the compiler cannot see through, so that
        // med3(x, y, ninf) is not folded back into a canonicalising min
        const float ninf = __builtin_bit_cast(float, opaque_copy(0xff800000u));
        for (int t = 0; t < ntiles; ++t) {
            float4_t acc[4][2];
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const unsigned char *rw = cl + (int64_t)(64 * t + 16 * it + c16) * 128;
                const i32x4_t lo = *reinterpret_cast<const i32x4_t *>(rw + p0 * 16);
                const i32x4_t hi = *reinterpret_cast<const i32x4_t *>(rw + p1 * 16);
                const i32x8_t x8 = (i32x8_t){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
                    acc[it][jt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        x8, q8[jt], (float4_t){0.f, 0.f, 0.f, 0.f}, 0, 0, 0, 0, 0, 0);
            }
            // lane l's accumulators: centres 64 t + 16 it + 4 g + v; padded
            // centres have xn' = FLT_MAX, so their keys stay finite
            const float b0[2] = {bk[0], bk[1]};
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const float4_t xr = *reinterpret_cast<const float4_t *>(&xl[64 * t + 16 * it + 4 * g]);
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float kv = __builtin_fmaf(fx[jt], acc[it][jt][v], xr[v]);
                        const float kp = __builtin_bit_cast(
                            float, (__builtin_bit_cast(uint32_t, kv) & ~63u) | (uint32_t)(16 * it + v));
                        // med3 with -inf is the min, without the canonicalising
                        // max that fminf puts on a bit-cast operand
                        sec[jt] = __builtin_amdgcn_fmed3f(bk[jt], sec[jt], kp);
                        bk[jt] = __builtin_amdgcn_fmed3f(bk[jt], kp, ninf);
                    }
            }
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) bt[jt] = bk[jt] != b0[jt] ? t : bt[jt];
        }
        // decision: the smallest packed key of the row's four lanes (the
        // lowest g among equal ones) names the candidate; every other key,
        // widened by the packing's 2^-16 relative slack, must exceed its
        // exact distance
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
            const int d = args.d;
            bool ok = gw < 4 && imin >= 0 && imin < k;
            double bf = 0.0;
            if (ok) {
                bf = args.dec_x_f32 ? fold_exact((const float *)args.dec_X + q * d,
                                                 args.dec_C + (int64_t)imin * d, d)
                                    : fold_exact((const double *)args.dec_X + q * d,
                                                 args.dec_C + (int64_t)imin * d, d);
                if (thk != __builtin_huge_valf()) {
                    const double kf = (double)thk, qo = args.qoff[q];
                    ok = (kf + qo) - ((fabs(kf) + fabs(qo)) * 0x1.0p-16 + 0x1.0p-140) > bf;
                }
            }
            if (ok) {
                args.labels[q] = imin;
                args.best[q] = bf;
            } else {
                args.flags[atomicAdd(args.fcount, 1)] = (int32_t)q;
            }
        }
    }
#endif
}

}  // namespace fl
