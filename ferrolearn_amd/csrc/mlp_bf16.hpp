// mlp_bf16.hpp -- kernels of the cell router: a d -> H -> H -> C perceptron
// with ReLU hidden layers and a softmax output, trained with softmax
// cross-entropy and Adam, all in BF16 MFMA with FP32 accumulation.
//
// The network follows ferrolearn-neural/src/mlp.rs: z = a @ W + b with W
// stored fan_in x fan_out (mlp.rs:229-235), ReLU hidden activations
// (:69-74) whose derivative is z > 0 (:86-89), softmax rows (:106-127),
// loss -sum(y ln max(p, 1e-15)) / batch and output delta p - y
// (:1326-1340), grad_w = a^T delta / batch + alpha W, grad_b = mean(delta)
// (:316-325), delta_prev = (delta @ W^T) * relu'(z) (:329-337) and Adam with
// bias correction (:372-409).  Activations, deltas and the weights' working
// copies are BF16; master weights, Adam moments, logits and the softmax are
// FP32.
#pragma once

#include "common.hpp"
#include "dist_bf16.hpp"

namespace fl {

constexpr int MG_BM = 64;   // C rows per block
constexpr int MG_BN = 64;   // C columns per block
constexpr int MG_BK = 32;   // K per MFMA step (v_mfma_f32_16x16x32_bf16)

// C[M x N] = scale * op(A)[M x K] op(B)[K x N] (+ addscale * addf) (+ bias[n]),
// then ReLU and / or a ReLU-derivative mask, stored as f32 and / or bf16.
// K must be a multiple of 32, and K-contiguous operands must hold zeros in
// any padding up to K.
//   A_KCONT: A(m, k) = A[m * lda + k]   else A(m, k) = A[k * lda + m]
//   B_KCONT: B(k, n) = B[n * ldb + k]   else B(k, n) = B[k * ldb + n]
struct MlpGemmArgs {
    const uint16_t *A;
    int64_t lda;
    const uint16_t *B;
    int64_t ldb;
    int M, N, K;
    float scale;
    const float *addf;  // optional
    int64_t ldadd;
    float addscale;
    const float *bias;  // optional, per column
    int relu;
    const uint16_t *mask;  // optional: keep where mask(m, n) > 0
    int64_t ldmask;
    float *Cf;  // optional f32 output
    int64_t ldcf;
    uint16_t *Cb;  // optional bf16 output
    int64_t ldcb;
};

// 16-byte chunk position of (row, chunk) in a 64-byte LDS row: rows 4 apart
// land on different banks, so the 16 rows one MFMA fragment reads are
// conflict-free.
__device__ __forceinline__ int mg_swz(int row, int chunk) { return chunk ^ ((row >> 2) & 3); }

// Stage one 64 x 32 operand tile (rows = M or N index, columns = K) into LDS.
template <bool KCONT>
__device__ __forceinline__ void mg_stage(uint16_t *__restrict__ S, const uint16_t *__restrict__ P,
                                         int64_t ld, int r0, int rows, int k0, int t) {
    if (KCONT) {
        const int r = t >> 2, ch = t & 3;
        u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
        if (r0 + r < rows)
            v = *reinterpret_cast<const u16x8_t *>(P + (int64_t)(r0 + r) * ld + k0 + ch * 8);
        *reinterpret_cast<u16x8_t *>(S + r * MG_BK + mg_swz(r, ch) * 8) = v;
    } else {
        const int kk = t >> 3, rc = (t & 7) * 8;
        const uint16_t *src = P + (int64_t)(k0 + kk) * ld + r0 + rc;
        uint16_t v[8];
        if ((ld & 7) == 0 && r0 + rc + 8 <= rows) {
            // eight contiguous, 16-byte aligned elements: one vector load
            const u16x8_t w = *reinterpret_cast<const u16x8_t *>(src);
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = w[j];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = r0 + rc + j < rows ? src[j] : (uint16_t)0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int r = rc + j;
            S[r * MG_BK + mg_swz(r, kk >> 3) * 8 + (kk & 7)] = v[j];
        }
    }
}

// 256 threads: four waves, each owning a 32 x 32 quarter of the 64 x 64 tile
// as 2 x 2 MFMA tiles.
template <bool A_KCONT, bool B_KCONT>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(MlpGemmArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t As[MG_BM * MG_BK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[MG_BN * MG_BK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * MG_BM, n0 = blockIdx.x * MG_BN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < a.K; k0 += MG_BK) {
        mg_stage<A_KCONT>(As, a.A, a.lda, m0, a.M, k0, t);
        mg_stage<B_KCONT>(Bs, a.B, a.ldb, n0, a.N, k0, t);
        __syncthreads();
        bf16x8_t af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int ra = wm + i * 16 + (lane & 15), rb = wn + i * 16 + (lane & 15);
            af[i] = __builtin_bit_cast(
                bf16x8_t, *reinterpret_cast<const u16x8_t *>(As + ra * MG_BK +
                                                             mg_swz(ra, lane >> 4) * 8));
            bfr[i] = __builtin_bit_cast(
                bf16x8_t, *reinterpret_cast<const u16x8_t *>(Bs + rb * MG_BK +
                                                             mg_swz(rb, lane >> 4) * 8));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0,
                                                                    0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + i * 16 + bf_acc_row(lane, v);
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= a.M || n >= a.N) continue;
                float x = acc[i][j][v] * a.scale;
                if (a.addf) x += a.addscale * a.addf[(int64_t)m * a.ldadd + n];
                if (a.bias) x += a.bias[n];
                if (a.relu) x = x > 0.f ? x : 0.f;
                if (a.mask && !(bf16_bits_to_f32(a.mask[(int64_t)m * a.ldmask + n]) > 0.f))
                    x = 0.f;
                if (a.Cf) a.Cf[(int64_t)m * a.ldcf + n] = x;
                if (a.Cb) a.Cb[(int64_t)m * a.ldcb + n] = bf16_bits_rne(x);
            }
}

// Row-wise softmax and cross-entropy against the soft target "fraction of
// the row's valid target cells equal to c" (mlp.rs:1326-1340; a row lists k
// target cells, -1 for none): writes the output delta p - y as bf16 (zero in
// padded columns, in rows >= nb and in rows without targets) and the row's
// loss.  One 256-thread block per row; LDS holds the row.
__global__ __launch_bounds__(256) void mlp_softmax_xent_kernel(
    const float *__restrict__ Z, int64_t ldz, int C, int Cp, const int32_t *__restrict__ cells,
    int k, int nb, uint16_t *__restrict__ G, int64_t ldg, double *__restrict__ row_loss) {
    extern __shared__ float p[];
    __shared__ float red[256];
    const int b = blockIdx.x, t = threadIdx.x;
    uint16_t *g = G + (int64_t)b * ldg;
    int nvalid = 0;
    if (b < nb)
        for (int j = 0; j < k; ++j) nvalid += cells[(int64_t)b * k + j] >= 0 ? 1 : 0;
    if (b >= nb || nvalid == 0) {
        for (int c = t; c < Cp; c += 256) g[c] = 0;
        if (b < nb && t == 0) row_loss[b] = 0.0;
        return;
    }
    const float *z = Z + (int64_t)b * ldz;
    float mx = -__builtin_huge_valf();
    for (int c = t; c < C; c += 256) mx = fmaxf(mx, z[c]);
    red[t] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] = fmaxf(red[t], red[t + o]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float s = 0.f;
    for (int c = t; c < C; c += 256) {
        const float e = expf(z[c] - mx);
        p[c] = e;
        s += e;
    }
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    s = red[0];
    if (s > 0.f)
        for (int c = t; c < C; c += 256) p[c] = p[c] / s;
    __syncthreads();
    if (t == 0) {
        const int32_t *cb = cells + (int64_t)b * k;
        double loss = 0.0;
        for (int j = 0; j < k; ++j) {
            const int c = cb[j];
            if (c >= 0 && c < C) loss -= log((double)fmaxf(p[c], 1e-15f)) / nvalid;
        }
        row_loss[b] = loss;
        const float inv = 1.0f / (float)nvalid;
        for (int j = 0; j < k; ++j) {
            const int c = cb[j];
            if (c >= 0 && c < C) p[c] -= inv;
        }
    }
    __syncthreads();
    for (int c = t; c < Cp; c += 256) g[c] = c < C ? bf16_bits_rne(p[c]) : (uint16_t)0;
}

// Adam step over one parameter block (mlp.rs:372-409); optionally refresh
// the bf16 working copy.
__global__ void mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ G,
                                float *__restrict__ Mo, float *__restrict__ V, int64_t n, float lr,
                                float b1, float b2, float eps, float bc1, float bc2,
                                uint16_t *__restrict__ Pb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = b1 * Mo[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    Mo[i] = m;
    V[i] = v;
    const float p = P[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    P[i] = p;
    if (Pb) Pb[i] = bf16_bits_rne(p);
}

// f32 -> bf16 over a block of parameters.
__global__ void mlp_to_bf16_kernel(const float *__restrict__ P, int64_t n,
                                   uint16_t *__restrict__ Pb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Pb[i] = bf16_bits_rne(P[i]);
}

// Input rows of a batch as bf16, zero padded to dp features and to the
// batch's padded row count: out[b] = Q[rows ? rows[b] : first + b].
template <typename TQ>
__global__ void mlp_gather_rows_kernel(const TQ *__restrict__ Q, int d, const int32_t *__restrict__ rows,
                                       int64_t first, int nb, int nbp, int dp,
                                       uint16_t *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nbp * dp) return;
    const int b = (int)(t / dp), j = (int)(t % dp);
    uint16_t v = 0;
    if (b < nb && j < d) {
        const int64_t r = rows ? (int64_t)rows[b] : first + b;
        v = bf16_bits_rne((float)Q[r * d + j]);
    }
    out[t] = v;
}

// Target cells of a batch: out[b][j] = cells[rows[b]][j].
__global__ void mlp_gather_cells_kernel(const int32_t *__restrict__ cells, int k,
                                        const int32_t *__restrict__ rows, int nb,
                                        int32_t *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * k) return;
    out[t] = cells[(int64_t)rows[t / k] * k + t % k];
}

}  // namespace fl
