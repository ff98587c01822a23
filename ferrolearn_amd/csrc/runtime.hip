// runtime.hip -- error reporting, library info and Backend::gemm.
#include <cstring>
#include <string>

#include "host.hpp"

namespace fl {

static thread_local std::string g_last_error;

void set_error(fl_status, const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
fl_status fail(fl_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

size_t dtype_size(int32_t dtype) { return dtype == FL_F32 ? 4 : 8; }

// C = A . B in f64 (row-major m x k times k x n) on v_mfma_f64_16x16x4f64.
// One block of 4 waves per 64 x 64 tile of C; wave w owns the 32 x 32
// quarter (w / 2, w % 2) as 2 x 2 MFMA tiles.  A and B go through LDS 16
// columns / rows of k at a time (zero filled past the edges), so every
// product the tile needs is formed exactly once per k.  The MFMA operands:
// lane l supplies A[i][kk] with i = l % 16, kk = l / 16, and B[kk][j] with
// j = l % 16; it receives D[l / 16 + 4 v][l % 16], v = 0..3.
constexpr int GM_T = 64;    // C tile edge
constexpr int GM_K = 16;    // k per LDS stage
constexpr int GM_AP = GM_K + 1;    // A row stride in LDS (odd: 16 rows hit distinct banks)
constexpr int GM_BP = GM_T + 2;    // B row stride in LDS

__global__ __launch_bounds__(256) void gemm_f64_kernel(const double *__restrict__ A,
                                                       const double *__restrict__ B, int64_t m,
                                                       int64_t k, int64_t n,
                                                       double *__restrict__ C) {
    __shared__ double As[GM_T * GM_AP];
    __shared__ double Bs[GM_K * GM_BP];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int64_t bm = (int64_t)blockIdx.y * GM_T, bn = (int64_t)blockIdx.x * GM_T;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    double4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    // staging: thread t copies A[bm + t / 4][k0 + 4 (t % 4) + e] and
    // B[k0 + t / 16][bn + 4 (t % 16) + e], e = 0..3
    const int ar = t >> 2, ac = (t & 3) * 4;
    const int br = t >> 4, bc = (t & 15) * 4;
    for (int64_t k0 = 0; k0 < k; k0 += GM_K) {
        double av[4], bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t ai = bm + ar, ak = k0 + ac + e;
            av[e] = ai < m && ak < k ? A[ai * k + ak] : 0.0;
            const int64_t bk = k0 + br, bj = bn + bc + e;
            bv[e] = bk < k && bj < n ? B[bk * n + bj] : 0.0;
        }
        __syncthreads();   // the previous stage's reads are done
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            As[ar * GM_AP + ac + e] = av[e];
            Bs[br * GM_BP + bc + e] = bv[e];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GM_K; ks += 4) {
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[(wr + 16 * i + c16) * GM_AP + ks + g];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[(ks + g) * GM_BP + wc + 16 * j + c16];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = bm + wr + 16 * i + g + 4 * v, c = bn + wc + 16 * j + c16;
                if (r < m && c < n) C[r * n + c] = acc[i][j][v];
            }
}

}  // namespace fl

using namespace fl;

extern "C" {

const char *fl_last_error(void) { return g_last_error.c_str(); }

int fl_version(void) { return 100; }

int fl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

fl_status fl_device_malloc(void **ptr, uint64_t bytes) {
    return guarded([&]() -> fl_status {
        if (!ptr) return fail(FL_INVALID_PARAMETER, "InvalidParameter: ptr: must not be NULL");
        *ptr = nullptr;
        FL_HIP(hipMalloc(ptr, bytes ? bytes : 1));
        return FL_OK;
    });
}

fl_status fl_device_free(void *ptr) {
    return guarded([&]() -> fl_status {
        if (ptr) FL_HIP(hipFree(ptr));
        return FL_OK;
    });
}

fl_status fl_memcpy_to_device(void *dst, const void *src, uint64_t bytes) {
    return guarded([&]() -> fl_status {
        if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return FL_OK;
    });
}

fl_status fl_memcpy_to_host(void *dst, const void *src, uint64_t bytes) {
    return guarded([&]() -> fl_status {
        if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

fl_status fl_memcpy_device(void *dst, const void *src, uint64_t bytes) {
    return guarded([&]() -> fl_status {
        if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
        return FL_OK;
    });
}

fl_status fl_device_synchronize(void) {
    return guarded([&]() -> fl_status {
        FL_HIP(hipDeviceSynchronize());
        return FL_OK;
    });
}

// Backend::gemm (ferrolearn-core/src/backend.rs:41-44,
// backend_faer.rs:61-78): ShapeMismatch (expected A's shape, actual B's)
// when the inner dimensions differ; otherwise C = A . B.  The inputs go to
// the device on one stream and C comes back before the call returns.
fl_status fl_backend_gemm_f64(const double *A, int64_t m, int64_t k, const double *B, int64_t k2,
                              int64_t n, double *C) {
    return guarded([&]() -> fl_status {
        if (k != k2) {
            const std::string a = std::to_string(m) + "x" + std::to_string(k);
            const std::string b = std::to_string(k2) + "x" + std::to_string(n);
            return fail(FL_SHAPE_MISMATCH,
                        "ShapeMismatch: expected [" + std::to_string(m) + ", " +
                            std::to_string(k) + "], actual [" + std::to_string(k2) + ", " +
                            std::to_string(n) + "]: gemm: A is " + a + " but B is " + b +
                            " (inner dimensions " + std::to_string(k) + " != " +
                            std::to_string(k2) + ")");
        }
        if (m < 0 || k < 0 || n < 0)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: gemm: negative dimension");
        if (m == 0 || n == 0) return FL_OK;
        if (ceil_div(m, GM_T) > 65535)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: gemm: m above 65535 * 64");
        DevBuf a, b, c;
        double *dA = (double *)a.get((size_t)(m * k) * 8);
        double *dB = (double *)b.get((size_t)(k * n) * 8);
        double *dC = (double *)c.get((size_t)(m * n) * 8);
        hipStream_t st = nullptr;
        FL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{st};
        if (m * k) FL_HIP(hipMemcpyAsync(dA, A, (size_t)(m * k) * 8, hipMemcpyHostToDevice, st));
        if (k * n) FL_HIP(hipMemcpyAsync(dB, B, (size_t)(k * n) * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(gemm_f64_kernel, dim3((unsigned)ceil_div(n, GM_T), (unsigned)ceil_div(m, GM_T)),
                           dim3(256), 0, st, dA, dB, m, k, n, dC);
        check_launch();
        FL_HIP(hipMemcpyAsync(C, dC, (size_t)(m * n) * 8, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        return FL_OK;
    });
}

}  // extern "C"
