// common.hpp -- shared host/device helpers for the gfx950 distance path.
//
// Every translation unit is compiled with -ffp-contract=off, so `a + b * c`
// is two rounded ops, never an FMA: the exact folds below reproduce the
// reference's Rust arithmetic (ferrolearn-cluster/src/kmeans.rs:159-163).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/ferrolearn_hip.h"

namespace fl {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors --
struct Error {
    fl_status code;
    std::string msg;
};

void set_error(fl_status code, const std::string &msg);
void clear_error();
fl_status fail(fl_status code, const std::string &msg);

#define FL_HIP(call)                                                         \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            throw ::fl::DeviceError(std::string(#call) + ": " +             \
                                    hipGetErrorString(e_));                  \
    } while (0)

struct DeviceError {
    std::string msg;
    explicit DeviceError(std::string m) : msg(std::move(m)) {}
};

// A FerroError-style failure detected after work has started (e.g. a NaN
// distance found on the device).
struct StatusError {
    fl_status code;
    std::string msg;
};

// Wait until every outstanding vector-memory load of this wave has landed
// (s_waitcnt vmcnt(0), other counters untouched).  Used after a kernel's
// prologue loads so that the compiler does not have to guard their first
// use inside the main loop with a wait that would also drain the loop's own
// prefetches.
__device__ __forceinline__ void wait_vm_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ------------------------------------------------------------ device math --
template <typename T>
__device__ __forceinline__ double ld64(const T *p) {
    return (double)(*p);
}

// Exact reference fold acc = acc + (a_j - b_j)^2, left to right from +0.
template <typename TA, typename TB>
__device__ __forceinline__ double fold_exact(const TA *a, const TB *b, int d) {
    // Loads are issued 16 at a time so that their latencies overlap; the sum
    // is still the left fold in feature order.
    double acc = 0.0;
    int j = 0;
    for (; j + 16 <= d; j += 16) {
        double av[16], bv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            av[u] = (double)a[j + u];
            bv[u] = (double)b[j + u];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double t = av[u] - bv[u];
            acc = acc + t * t;
        }
    }
    for (; j < d; ++j) {
        const double t = (double)a[j] - (double)b[j];
        acc = acc + t * t;
    }
    return acc;
}

// Lane u's double, to every lane of the wave (u wave-uniform): two
// v_readlane_b32, which cost far less than a shuffle through LDS.
__device__ __forceinline__ double wave_bcast(double v, int u) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, u);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), u);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// (key, idx) lexicographic order; NaN keys compare false like Rust's `<`.
__device__ __forceinline__ bool pair_less(double ka, int64_t ia, double kb,
                                          int64_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

__host__ __device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) {
    return (a + b - 1) / b;
}

}  // namespace fl
