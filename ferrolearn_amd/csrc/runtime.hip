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

// C = A . B in f64; one lane per output element, k ascending.
__global__ void gemm_f64_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                int64_t m, int64_t k, int64_t n, double *__restrict__ C) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * n) return;
    const int64_t i = t / n, j = t % n;
    double acc = 0.0;
    for (int64_t p = 0; p < k; ++p) acc = acc + A[i * k + p] * B[p * n + j];
    C[t] = acc;
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
// backend_faer.rs:61-78): ShapeMismatch when the inner dimensions differ.
fl_status fl_backend_gemm_f64(const double *A, int64_t m, int64_t k, const double *B, int64_t k2,
                              int64_t n, double *C) {
    return guarded([&]() -> fl_status {
        if (k != k2)
            return fail(FL_SHAPE_MISMATCH,
                        "ShapeMismatch: expected [" + std::to_string(k) + "], actual [" +
                            std::to_string(k2) + "]: gemm inner dimensions must match");
        if (m == 0 || n == 0) return FL_OK;
        DevBuf a, b, c;
        double *dA = (double *)a.get((size_t)(m * k) * 8);
        double *dB = (double *)b.get((size_t)(k * n) * 8);
        double *dC = (double *)c.get((size_t)(m * n) * 8);
        if (m * k) FL_HIP(hipMemcpy(dA, A, (size_t)(m * k) * 8, hipMemcpyHostToDevice));
        if (k * n) FL_HIP(hipMemcpy(dB, B, (size_t)(k * n) * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(gemm_f64_kernel, dim3(grid1(m * n, 256)), dim3(256), 0, 0, dA, dB, m,
                           k, n, dC);
        check_launch();
        FL_HIP(hipMemcpy(C, dC, (size_t)(m * n) * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

}  // extern "C"
