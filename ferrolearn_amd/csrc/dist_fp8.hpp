// dist_fp8.hpp -- FP8 E4M3 KMeans argmin filter with every centre resident
// in LDS (k <= 1024, d <= 128), the decision fused.
//
// Same contract and arithmetic as the FP8 form of dist_bf16_kernel (argmin,
// PASSES == 0): a.b on v_mfma_scale_f32_16x16x128_f8f6f4 with unit scales,
// key = fma(afac_row, acc, xn'_c) in f32, per lane the best key, its index
// and the second key; the smallest (key, index) of a row's four lanes is
// folded exactly and certified against every other key; rows that fail are
// listed for the next filter.
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
    // rows past k are zero, and their keys are masked below
    for (int64_t e = threadIdx.x; e < kp * 8; e += blockDim.x) {
        const int64_t r = e >> 3;
        const int p = (int)(e & 7);
        i32x4_t v = (i32x4_t){0, 0, 0, 0};
        if (r < k) v = *reinterpret_cast<const i32x4_t *>(c8 + r * 128 + p * 16);
        *reinterpret_cast<i32x4_t *>(cl + r * 128 + ((p ^ (int)(r & 7)) * 16)) = v;
    }
    for (int64_t r = threadIdx.x; r < kp; r += blockDim.x)
        xl[r] = r < k ? args.xnp[r] : __builtin_huge_valf();
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
        float bk[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
        float sec[2] = {__builtin_huge_valf(), __builtin_huge_valf()};
        int32_t bi[2] = {0x7fffffff, 0x7fffffff};
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
            // lane l's accumulators: centres 64 t + 16 it + 4 g + v; a lane
            // sees its centres in increasing order, so strict < keeps the
            // lowest index among equal keys; centres past k have xn' = +inf
            const int32_t rb = 64 * t + 4 * g;
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const float4_t xr = *reinterpret_cast<const float4_t *>(&xl[64 * t + 16 * it + 4 * g]);
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float kv = __builtin_fmaf(fx[jt], acc[it][jt][v], xr[v]);
                        const bool lt = kv < bk[jt];
                        sec[jt] = __builtin_amdgcn_fmed3f(bk[jt], sec[jt], kv);
                        bk[jt] = fminf(bk[jt], kv);
                        bi[jt] = lt ? rb + 16 * it + v : bi[jt];
                    }
            }
        }
        // decision: as dist_bf16_kernel's fused argmin decision
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int64_t q = q0 + 16 * jt + c16;
            const float kmin = xred<false, 32>(xred<false, 16>(bk[jt]));
            const float icand = bk[jt] == kmin ? (float)bi[jt] : 0x1.0p30f;
            const float imf = xred<false, 32>(xred<false, 16>(icand));
            const int32_t imin = imf >= 0x1.0p30f ? -1 : (int32_t)imf;
            const bool mine = bk[jt] == kmin && bi[jt] == imin;
            const float other = mine ? sec[jt] : fminf(sec[jt], bk[jt]);
            const float thk = xred<false, 32>(xred<false, 16>(other));
            if (q >= nA || g != jt) continue;
            const int d = args.d;
            bool ok = imin >= 0 && imin < k;
            double bf = 0.0;
            if (ok) {
                bf = args.dec_x_f32 ? fold_exact((const float *)args.dec_X + q * d,
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
    }
#endif
}

}  // namespace fl
