// host.hpp -- host-side runtime pieces shared by the KMeans and kNN entry
// points: device buffers, a growable workspace, stream-ordered event timing,
// error-bound constants and the filter-kernel launcher.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "dist_bf16.hpp"
#include "dist_fp8.hpp"
#include "dist_bf16_res.hpp"
#include "rand09.hpp"
#include "split_fold.hpp"
#include "dist_f64.hpp"
#include "dist_topm.hpp"

namespace fl {

// Device buffer owned by the host object (hipMalloc / hipFree).
class DevBuf {
  public:
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void *get(size_t bytes) {
        if (bytes > cap_) {
            release();
            size_t want = bytes < 256 ? 256 : bytes;
            FL_HIP(hipMalloc(&p_, want));
            cap_ = want;
        }
        return p_;
    }
    void *ptr() const { return p_; }
    size_t cap() const { return cap_; }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        cap_ = 0;
    }

  private:
    void *p_ = nullptr;
    size_t cap_ = 0;
};

// Streams and named-buffer sets are kept per device after a call and handed
// to the next one: a new stream costs about a millisecond and each hipMalloc
// tens of microseconds, against a few milliseconds for a whole small fit.
// Only sets of at most kPoolKeepBytes are kept (larger ones are freed, so a
// large fit gives its memory back), and at most kPoolMax of each per device.
// The pool lives until the process ends.
using BufSet = std::map<std::string, DevBuf>;

class DevicePool {
  public:
    static constexpr size_t kPoolKeepBytes = (size_t)64 << 20;
    static constexpr size_t kPoolMax = 16;

    static DevicePool &get() {
        static DevicePool *p = new DevicePool;   // never destroyed: no hipFree at exit
        return *p;
    }
    hipStream_t take_stream(int dev) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto &v = streams_[dev];
            if (!v.empty()) {
                hipStream_t s = v.back();
                v.pop_back();
                return s;
            }
        }
        hipStream_t s = nullptr;
        FL_HIP(hipStreamCreate(&s));
        return s;
    }
    void give_stream(int dev, hipStream_t s) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto &v = streams_[dev];
            if (v.size() < kPoolMax) {
                v.push_back(s);
                return;
            }
        }
        (void)hipStreamDestroy(s);
    }
    BufSet *take_bufs(int dev) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto &v = bufs_[dev];
            if (!v.empty()) {
                BufSet *b = v.back();
                v.pop_back();
                return b;
            }
        }
        return new BufSet;
    }
    // The caller's work on these buffers may still be in flight on a stream
    // this object does not know: the device is synchronised before the set
    // goes back (hipFree would have synchronised too).
    void give_bufs(int dev, BufSet *b) {
        size_t bytes = 0;
        for (auto &kv : *b) bytes += kv.second.cap();
        if (bytes <= kPoolKeepBytes && hipDeviceSynchronize() == hipSuccess) {
            std::lock_guard<std::mutex> g(mu_);
            auto &v = bufs_[dev];
            if (v.size() < kPoolMax) {
                v.push_back(b);
                return;
            }
        }
        delete b;
    }

  private:
    std::mutex mu_;
    std::map<int, std::vector<hipStream_t>> streams_;
    std::map<int, std::vector<BufSet *>> bufs_;
};

inline int current_device() {
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    return dev;
}

// Named scratch buffers that only grow; reused across calls on one object,
// and the set itself taken from and returned to the device's pool.
// Arena form (constructed over caller memory): get() hands out 256-byte
// aligned slices of that memory by name and never allocates; a slot keeps
// its first size, and running past the end throws instead of allocating.
class Workspace {
  public:
    Workspace() : dev_(current_device()), bufs_(DevicePool::get().take_bufs(dev_)) {}
    Workspace(void *base, size_t bytes) : arena_((unsigned char *)base), arena_cap_(bytes) {
        if (!arena_) {
            dev_ = current_device();
            bufs_ = DevicePool::get().take_bufs(dev_);
        }
    }
    ~Workspace() {
        if (bufs_) DevicePool::get().give_bufs(dev_, bufs_);
    }
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;

    template <typename T>
    T *get(const std::string &name, int64_t count) {
        const size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(T);
        if (arena_) return static_cast<T *>(arena_get(name.c_str(), bytes));
        return static_cast<T *>((*bufs_)[name].get(bytes));
    }
    void clear() {
        if (bufs_) bufs_->clear();
    }
    size_t arena_used() const { return arena_used_; }

  private:
    static constexpr int kSlots = 48;
    struct Slot {
        char name[24];
        size_t off, bytes;
    };
    void *arena_get(const char *name, size_t bytes) {
        for (int i = 0; i < nslots_; ++i)
            if (std::strncmp(slots_[i].name, name, sizeof(slots_[i].name)) == 0) {
                if (bytes > slots_[i].bytes)
                    throw DeviceError(std::string("internal: scratch slot '") + name + "' grew");
                return arena_ + slots_[i].off;
            }
        const size_t off = (arena_used_ + 255) & ~(size_t)255;
        if (nslots_ == kSlots || std::strlen(name) >= sizeof(slots_[0].name) ||
            off + bytes > arena_cap_)
            throw DeviceError("internal: the fixed scratch buffer is too small for this call");
        Slot &s = slots_[nslots_++];
        std::strncpy(s.name, name, sizeof(s.name));
        s.off = off;
        s.bytes = bytes;
        arena_used_ = off + bytes;
        return arena_ + off;
    }
    int dev_ = 0;
    BufSet *bufs_ = nullptr;
    unsigned char *arena_ = nullptr;
    size_t arena_cap_ = 0, arena_used_ = 0;
    Slot slots_[kSlots];
    int nslots_ = 0;
};

// Accumulates kernel time between two events recorded on one stream.  The
// event pairs are pooled: release(id) returns a pair once its time has been
// read (or is no longer wanted), so a loop of phases creates its events once.
class PhaseTimer {
  public:
    explicit PhaseTimer(hipStream_t s) : s_(s) {}
    ~PhaseTimer() {
        for (auto &p : pairs_) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
    int begin() {
        int id;
        if (free_.empty()) {
            hipEvent_t a, b;
            FL_HIP(hipEventCreate(&a));
            FL_HIP(hipEventCreate(&b));
            pairs_.push_back({a, b});
            id = (int)pairs_.size() - 1;
        } else {
            id = free_.back();
            free_.pop_back();
        }
        FL_HIP(hipEventRecord(pairs_[id].first, s_));
        return id;
    }
    void end(int id) { FL_HIP(hipEventRecord(pairs_[id].second, s_)); }
    double ms(int id) {
        float t = 0.f;
        FL_HIP(hipEventSynchronize(pairs_[id].second));
        FL_HIP(hipEventElapsedTime(&t, pairs_[id].first, pairs_[id].second));
        return (double)t;
    }
    void release(int id) { free_.push_back(id); }

  private:
    hipStream_t s_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs_;
    std::vector<int> free_;
};

// A stream for one call: taken from the device's pool and returned to it
// (work still queued on it stays ordered before the next user's).
struct StreamGuard {
    hipStream_t s = nullptr;
    int dev = 0;
    StreamGuard() : dev(current_device()) { s = DevicePool::get().take_stream(dev); }
    ~StreamGuard() {
        if (s) DevicePool::get().give_stream(dev, s);
    }
    StreamGuard(const StreamGuard &) = delete;
    StreamGuard &operator=(const StreamGuard &) = delete;
};

// Zeroed page-locked host array (asynchronous device-to-host copies).
template <typename T>
class PinnedHost {
  public:
    explicit PinnedHost(size_t n) {
        FL_HIP(hipHostMalloc((void **)&p_, n * sizeof(T), hipHostMallocDefault));
        std::memset((void *)p_, 0, n * sizeof(T));
    }
    ~PinnedHost() { (void)hipHostFree(p_); }
    PinnedHost(const PinnedHost &) = delete;
    PinnedHost &operator=(const PinnedHost &) = delete;
    T &operator[](size_t i) { return p_[i]; }

  private:
    T *p_ = nullptr;
};

// Two events, destroyed with the object.
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    EventPair() {
        FL_HIP(hipEventCreate(&e[0]));
        FL_HIP(hipEventCreate(&e[1]));
    }
    ~EventPair() {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
};

struct WallClock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

// Rigorous bound coefficients for the f64 GEMM form (the filter in
// dist_topm.hpp): |value - fold| <= alpha*|a||b| + beta*(|a|^2 + |b|^2),
// covering the MFMA dot product, the row norms, the combination, the
// reference fold's own rounding and the key's rounding, times 2 for margin.
struct Bound {
    double alpha, beta;
};
inline double gamma_n(double m, double u) { return m * u / (1.0 - m * u); }
inline Bound bound_fp64(int d) {
    const double u = std::ldexp(1.0, -53);
    const double gd = gamma_n(d, u), gd2 = gamma_n(d + 2, u);
    return {2.0 * (2.0 * gd + 6.0 * u + 2.0 * gd2), 2.0 * (gd + 4.0 * u + gd2)};
}

// Same role for the BF16 forms (dist_bf16.hpp): the key is
// fma(-2, acc, fma(-c|q|, |x|, (1-b)|x|^2)) + (1-b)|q|^2 <= fold, with
//   acc error <= cdot |a||b|: split representation (3.2 ub^2 for the 3-product
//   form, 2 ub + ub^2 for one product, ub = 2^-8 plus the f64->f32 step) and
//   FP32 accumulation of PASSES*d exact products at <= 2^-23 per addition;
//   f32 epilogue roundings (u32 per op) and the exact fold's own error;
// everything doubled for margin.
inline Bound bound_bf16(int d, int passes) {
    const double u64 = std::ldexp(1.0, -53), u32 = std::ldexp(1.0, -24);
    const double ub = std::ldexp(1.0, -8) + std::ldexp(1.0, -23);
    const double uacc = std::ldexp(1.0, -23);
    const double gacc = gamma_n((double)passes * d, uacc);
    const double rep = passes == 3 ? 3.2 * ub * ub : 2.0 * ub + ub * ub;
    const double cdot = rep + (passes == 3 ? 1.03 : 1.01) * gacc;
    const double gf = gamma_n(d + 2, u64);
    return {2.0 * (2.0 * cdot + 3.0 * u32 + 2.0 * gf), 2.0 * (2.0 * u32 + 2.0 * gf)};
}

// Same role for the FP8 E4M3 form: rows x~ = s_x q(x / s_x) with per-row
// power-of-two scales (centres share one scale s_C).  Per element |x - x~| <=
// u8 |x| + s 2^-10 (u8 = 2^-4; 2^-10 is half the subnormal spacing times s),
// and s <= 2 max|x_j| / 448, so |x - x~| <= (u8 + rho)|x| with rho =
// 2 sqrt(d) / (448 * 1024) for a sample row, and the centres' term is
// bounded with Smax, the largest centre norm.  E4M3 x E4M3 products are
// exact in f32; the MFMA's accumulation of them is taken as <= eacc times
// the sum of |products| (<= |x~||c~|).  eacc = 2^-8: the instruction's
// accumulation error was measured at <= 2.1e-4 of that sum over 400 x 256
// dot products of mixed-magnitude E4M3 data, and 2^-8 leaves a factor 16.
// As for BF16, the f32 epilogue roundings and the exact fold's own error
// are added and everything is doubled.
inline Bound bound_fp8(int d) {
    const double u64 = std::ldexp(1.0, -53), u32 = std::ldexp(1.0, -24);
    const double u8 = std::ldexp(1.0, -4);
    const double rho = 2.0 * std::sqrt((double)d) / (448.0 * 1024.0);
    const double eacc = std::ldexp(1.0, -8);
    const double rx = u8 + rho;
    const double rep = rx + (1.0 + rx) * (u8 + rho);
    const double cdot = rep + eacc * (1.0 + rx) * (1.0 + u8 + rho) * 1.01;
    const double gf = gamma_n(d + 2, u64);
    return {2.0 * (2.0 * cdot + 3.0 * u32 + 2.0 * gf), 2.0 * (2.0 * u32 + 2.0 * gf)};
}

// The library's generator (SplitMix64 seeding xoshiro256**), identical to
// the oracle's restatement; kmeans.rs:178,203,213 draw from rand's StdRng,
// whose stream is not reproduced.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t out = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    double f64() { return (double)(next() >> 11) * 0x1.0p-53; }
    int64_t range(int64_t n) {
        return (int64_t)(((unsigned __int128)next() * (uint64_t)n) >> 64);
    }
    int64_t range_from(int64_t lo, int64_t hi) { return lo + range(hi - lo); }
};

inline void check_launch() { FL_HIP(hipGetLastError()); }

template <int MP, int NKC, int PASSES>
void launch_bf16_t(const Bf16Args &a, int S, hipStream_t st) {
    // Staged tiles during the stream; the output merge reuses the same LDS
    // (per wave 16 rows x 4 lists x MP/2 pairs of 8 bytes).
    const size_t merge = MP == 0 ? 0 : (size_t)BF_WAVES * 16 * 4 * (MP / 2) * 8;
    const size_t shmem = std::max(Bf16Tile<NKC, PASSES>::SHMEM, merge);
    auto kern = dist_bf16_kernel<MP, NKC, PASSES>;
    FL_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shmem));
    dim3 grid((unsigned)ceil_div(a.nA, BF_Q), (unsigned)S);
    hipLaunchKernelGGL(kern, grid, dim3(BF_WAVES * 64), shmem, st, a);
    check_launch();
}

template <int MP, int PASSES>
void launch_bf16_m(int nkc, const Bf16Args &a, int S, hipStream_t st) {
    switch (nkc) {
        case 1: return launch_bf16_t<MP, 1, PASSES>(a, S, st);
        case 2: return launch_bf16_t<MP, 2, PASSES>(a, S, st);
        case 4: return launch_bf16_t<MP, 4, PASSES>(a, S, st);
        case 8: return launch_bf16_t<MP, 8, PASSES>(a, S, st);
    }
    throw DeviceError("internal: unsupported feature chunk count");
}

// NKC (32-feature chunks, a power of two for the LDS swizzle) used for d
// features by the BF16 kernels; 0 when d is too large for the
// register-resident layout (d > 256).
inline int bf16_nkc(int64_t d) {
    const int64_t c = (d + 31) / 32;
    if (c <= 2) return (int)c;
    if (c <= 4) return 4;
    if (c <= 8) return 8;
    return 0;
}

// MP: candidates output per resident row and split (lists of MP/2 per lane).
template <int PASSES>
void launch_bf16_p(int MP, int nkc, const Bf16Args &a, int S, hipStream_t st) {
    switch (MP) {
        case 0: return launch_bf16_m<0, PASSES>(nkc, a, S, st);   // argmin form (4 per row)
        case 8: return launch_bf16_m<8, PASSES>(nkc, a, S, st);
        case 16: return launch_bf16_m<16, PASSES>(nkc, a, S, st);
    }
    throw DeviceError("internal: unsupported candidate list size");
}

inline void launch_bf16(int passes, int MP, int nkc, const Bf16Args &a, int S, hipStream_t st) {
    if (passes == 3) return launch_bf16_p<3>(MP, nkc, a, S, st);
    if (passes == 0) {
        // FP8 form: argmin with the decision fused, 128-byte rows
        if (MP != 0 || nkc != 2) throw DeviceError("internal: FP8 form is argmin over 128-byte rows");
        return launch_bf16_t<0, 2, 0>(a, S, st);
    }
    return launch_bf16_p<1>(MP, nkc, a, S, st);
}

// f64 filter: D4 = 4-feature steps (a power of two for the LDS swizzle), 0
// when d > 128 (the resident fragments would not fit in registers).
inline int f64_d4(int64_t d) {
    const int64_t q = (d + 3) / 4;
    for (int v : {4, 8, 16, 32})
        if (q <= v) return v;
    return 0;
}

template <int MP, int D4, typename TA, typename TB>
void launch_f64_t(const F64Args &a, int S, hipStream_t st) {
    const size_t merge = MP == 0 ? 0 : (size_t)F64_WAVES * 16 * 4 * (MP / 2) * 16;
    const size_t shmem = std::max(F64Tile<D4, TB>::SHMEM, merge);
    auto kern = dist_f64_kernel<MP, D4, TA, TB>;
    FL_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shmem));
    dim3 grid((unsigned)ceil_div(a.nA, F64_Q), (unsigned)S);
    hipLaunchKernelGGL(kern, grid, dim3(F64_WAVES * 64), shmem, st, a);
    check_launch();
}

template <int MP, typename TA, typename TB>
void launch_f64_m(int d4, const F64Args &a, int S, hipStream_t st) {
    switch (d4) {
        case 4: return launch_f64_t<MP, 4, TA, TB>(a, S, st);
        case 8: return launch_f64_t<MP, 8, TA, TB>(a, S, st);
        case 16: return launch_f64_t<MP, 16, TA, TB>(a, S, st);
        case 32: return launch_f64_t<MP, 32, TA, TB>(a, S, st);
    }
    throw DeviceError("internal: unsupported f64 feature steps");
}

// kNN: f64 queries against a database of TB; MP candidates per row and split.
template <typename TB>
void launch_f64_knn(int MP, int d4, const F64Args &a, int S, hipStream_t st) {
    if (MP == 8) return launch_f64_m<8, double, TB>(d4, a, S, st);
    if (MP == 16) return launch_f64_m<16, double, TB>(d4, a, S, st);
    throw DeviceError("internal: unsupported candidate list size");
}

// KMeans: samples of TA against f64 centres; argmin form.
template <typename TA>
void launch_f64_km(int d4, const F64Args &a, hipStream_t st) {
    launch_f64_m<0, TA, double>(d4, a, 1, st);
}

template <typename TA, typename TB, int M, int MODE>
void launch_topm_t(const TA *A, const TB *B, const TopmArgs &a, int S, hipStream_t st) {
    const size_t shmem = sizeof(TopmShared<M, MODE>);
    auto kern = dist_topm_kernel<TA, TB, M, MODE>;
    FL_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shmem));
    dim3 grid((unsigned)ceil_div(a.nA, TM_ROWS), (unsigned)S);
    hipLaunchKernelGGL(kern, grid, dim3(256), shmem, st, A, B, a);
    check_launch();
}

template <typename TA, typename TB>
void launch_topm(int M, int mode, const TA *A, const TB *B, const TopmArgs &a, int S,
                 hipStream_t st) {
    if (mode == MODE_MFMA64) {
        switch (M) {
            case 4: return launch_topm_t<TA, TB, 4, MODE_MFMA64>(A, B, a, S, st);
            case 16: return launch_topm_t<TA, TB, 16, MODE_MFMA64>(A, B, a, S, st);
            case 32: return launch_topm_t<TA, TB, 32, MODE_MFMA64>(A, B, a, S, st);
            case 64: return launch_topm_t<TA, TB, 64, MODE_MFMA64>(A, B, a, S, st);
        }
    } else {
        switch (M) {
            case 16: return launch_topm_t<TA, TB, 16, MODE_EXACT>(A, B, a, S, st);
            case 32: return launch_topm_t<TA, TB, 32, MODE_EXACT>(A, B, a, S, st);
            case 64: return launch_topm_t<TA, TB, 64, MODE_EXACT>(A, B, a, S, st);
        }
    }
    throw DeviceError("internal: unsupported candidate list size");
}

// Compute units of the current device (cached per device).
inline int device_cu_count() {
    static int cache[64] = {0};
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        int n = 0;
        FL_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = n > 0 ? n : 256;
    }
    return cache[dev];
}

// FP8 argmin with every centre in LDS (k <= FR_MAXK): one block per compute
// unit (its LDS holds the centres), at most one wave-group per 32 samples.
inline void launch_fp8_res(const Bf16Args &a, hipStream_t st) {
    if (a.nB > FR_MAXK) throw DeviceError("internal: too many centres for the LDS-resident filter");
    const size_t shmem = fp8_res_shmem(a.nB);
    FL_HIP(hipFuncSetAttribute((const void *)dist_fp8_res_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    const int64_t blocks =
        std::max<int64_t>(1, std::min<int64_t>(device_cu_count(), ceil_div(ceil_div(a.nA, 32), FR_WAVES)));
    hipLaunchKernelGGL(dist_fp8_res_kernel, dim3((unsigned)blocks), dim3(FR_WAVES * 64), shmem, st,
                       a);
    check_launch();
}

// The BF16x3 argmin filter with the centres resident in LDS (chunks of
// 1024 / nkc centres), one block per compute unit.
template <int NKC>
void launch_bf16_res_t(const BfResArgs &ra, hipStream_t st) {
    const size_t shmem = BfResShape<NKC>::SHMEM;
    FL_HIP(hipFuncSetAttribute((const void *)dist_bf16_res_kernel<NKC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    const int64_t blocks = std::max<int64_t>(
        1, std::min<int64_t>(device_cu_count(), ceil_div(ceil_div(ra.a.nA, 32), BR_WAVES)));
    hipLaunchKernelGGL(dist_bf16_res_kernel<NKC>, dim3((unsigned)blocks), dim3(BR_WAVES * 64),
                       shmem, st, ra);
    check_launch();
}

inline void launch_bf16_res(int nkc, const BfResArgs &ra, hipStream_t st) {
    switch (nkc) {
        case 1: return launch_bf16_res_t<1>(ra, st);
        case 2: return launch_bf16_res_t<2>(ra, st);
        case 4: return launch_bf16_res_t<4>(ra, st);
    }
    throw DeviceError("internal: unsupported feature chunk count for the resident filter");
}

// Number of splits of the streamed rows for the BF16 filter.  A block holds
// one compute unit (8 waves); time ~ rounds of blocks x tiles per split x
// (1 + 0.5 x the share of tiles whose epilogue inserts), where that share
// follows the list-filling transient: about 512 (1 + ln(T / 512)) of the T
// tiles of a split.
inline int64_t bf16_splits(int64_t nblk, int64_t nB, int64_t smax) {
    const double cus = device_cu_count();
    int64_t best = 1;
    double best_cost = 0.0;
    for (int64_t S = 1; S <= std::max<int64_t>(1, smax); ++S) {
        const double tiles = std::ceil((double)nB / (double)S / 64.0);
        const double rounds = std::ceil((double)(nblk * S) / cus);
        const double slow = tiles <= 512.0 ? 1.0 : 512.0 * (1.0 + std::log(tiles / 512.0)) / tiles;
        const double cost = rounds * tiles * (1.0 + 0.5 * slow);
        if (S == 1 || cost < best_cost) {
            best = S;
            best_cost = cost;
        }
    }
    return best;
}

inline unsigned grid1(int64_t n, int block) { return (unsigned)ceil_div(n, block); }

// Convert to f64 (exact widening) when needed; returns the f64 view.
template <typename T>
const double *as_f64(const T *X, int64_t count, Workspace &ws, const char *slot,
                     hipStream_t st);

template <>
inline const double *as_f64<double>(const double *X, int64_t, Workspace &, const char *,
                                    hipStream_t) {
    return X;
}
template <>
inline const double *as_f64<float>(const float *X, int64_t count, Workspace &ws, const char *slot,
                                   hipStream_t st) {
    double *Y = ws.get<double>(slot, count);
    if (count == 0) return Y;   // no launch with an empty grid
    hipLaunchKernelGGL(to_f64_kernel<float>, dim3(grid1(count, 256)), dim3(256), 0, st, X, count,
                       Y);
    check_launch();
    return Y;
}

// Library entry guard: turns exceptions into status codes.
template <typename F>
fl_status guarded(F &&f) {
    try {
        clear_error();
        return f();
    } catch (const DeviceError &e) {
        return fail(FL_DEVICE_ERROR, e.msg);
    } catch (const StatusError &e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(FL_DEVICE_ERROR, "host allocation failed");
    }
}

size_t dtype_size(int32_t dtype);

// Sequences at least this long are folded over many compute units
// (split_fold.hpp); shorter ones by one workgroup each.
constexpr int64_t kSplitFoldMin = 4 * SF_L;

// The left fold of count sequences v[q*ld .. q*ld + n) into sums[q] and,
// when prefix is given, the running sums prefix[q*ld + i] = s_{i+1}: bit for
// bit the serial loop, in stream order on st.  A non-zero *stop (when given)
// makes the whole fold a no-op.
inline void left_fold(const double *v, int64_t n, int64_t ld, int64_t count,
                      const int32_t *stop, double *sums, double *prefix, Workspace &ws,
                      hipStream_t st) {
    if (count <= 0) return;
    if (n < kSplitFoldMin) {
        hipLaunchKernelGGL(exact_fold_kernel, dim3((unsigned)count), dim3(EF_NT), 0, st, v, n,
                           ld, stop, sums, prefix);
        check_launch();
        return;
    }
    const int64_t nb = ceil_div(n, (int64_t)SF_L);
    const int64_t cc = count < 65535 ? count : 65535;   // sequences per pass (grid.y)
    double *bsum = ws.get<double>("sf_bsum", nb * cc);
    double *apx = ws.get<double>("sf_apx", nb * cc);
    SfPart *parts = ws.get<SfPart>("sf_parts", nb * cc * SF_NP);
    int32_t *npart = ws.get<int32_t>("sf_npart", nb * cc);
    SfStart *starts = ws.get<SfStart>("sf_starts", nb * cc * SF_NP);
    for (int64_t q0 = 0; q0 < count; q0 += cc) {
        const int64_t c = count - q0 < cc ? count - q0 : cc;
        const double *vq = v + q0 * ld;
        double *sq = sums ? sums + q0 : nullptr;
        double *pq = prefix ? prefix + q0 * ld : nullptr;
        const dim3 g2((unsigned)nb, (unsigned)c);
        hipLaunchKernelGGL(sf_part_kernel, g2, dim3(SF_NT), 0, st, vq, n, ld, nb, stop, bsum);
        hipLaunchKernelGGL(sf_guess_kernel, dim3((unsigned)c), dim3(SF_NT), 0, st, bsum, nb,
                           stop, apx);
        hipLaunchKernelGGL(sf_spec_kernel, g2, dim3(SF_NT), 0, st, vq, n, ld, nb, stop, apx, parts,
                           npart, pq);
        hipLaunchKernelGGL(sf_combine_kernel, dim3((unsigned)c), dim3(EF_NT), 0, st, vq, n, ld,
                           nb, stop, parts, npart, starts, sq, pq);
        if (pq)
            hipLaunchKernelGGL(sf_prefix_kernel, g2, dim3(SF_NT), 0, st, n, ld, nb, stop, parts,
                               npart, starts, pq);
        check_launch();
    }
}

}  // namespace fl
