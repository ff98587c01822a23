// metrics.hip -- euclidean_distances and pairwise_distances_argmin(_min)
// with Metric::Euclidean on gfx950.
//
// Reference: ferrolearn-metrics/src/pairwise.rs:157-189 and :564-629, whose Euclidean
// distances are the expansion of euclidean_distances (:157-185):
//   x_sq = fold(x*x), y_sq = fold(y*y), dot = fold(x*y)   (left folds)
//   dist = sqrt(max((x_sq + y_sq) - 2*dot, 0))            (max: the non-NaN operand)
// and the index the first j of the smallest distance (best starts at j = 0,
// then strict <).  This is not the squared_euclidean fold of the KMeans and
// kNN paths -- the expansion cancels -- so it has its own exact kernel: every
// pair is evaluated with these operations in this order, one IEEE op each
// (no contraction), and the result is bit-exact.
#include <algorithm>
#include <string>

#include "host.hpp"

using namespace fl;

namespace {

constexpr int PW_ROWS = 256;   // X rows per block, one per thread
constexpr int PW_J = 32;       // Y rows per tile (accumulators per thread)
constexpr int PW_K = 32;       // features staged per step

// Each thread folds its X row against PW_J staged Y rows at a time; features
// arrive PW_K at a time through LDS (X transposed, one column per thread),
// so every accumulator still adds its products in feature order.  FULL:
// every distance goes to dist_out (row-major n x m) instead of the argmin.
template <typename T, bool FULL>
__global__ __launch_bounds__(PW_ROWS) void pw_argmin_kernel(
    const T *__restrict__ X, int64_t n, const T *__restrict__ Y, int64_t m, int d,
    const double *__restrict__ xsq, const double *__restrict__ ysq, int64_t *__restrict__ idx_out,
    double *__restrict__ min_out, double *__restrict__ dist_out) {
    __shared__ double xs[PW_K][PW_ROWS + 1];
    __shared__ double ys[PW_K][PW_J];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * PW_ROWS;
    const int64_t i = r0 + t;
    const double xq = i < n ? xsq[i] : 0.0;
    double bv = 0.0;
    int64_t bi = 0;
    for (int64_t j0 = 0; j0 < m; j0 += PW_J) {
        double acc[PW_J];
#pragma unroll
        for (int jj = 0; jj < PW_J; ++jj) acc[jj] = 0.0;
        for (int k0 = 0; k0 < d; k0 += PW_K) {
            const int kw = d - k0 < PW_K ? d - k0 : PW_K;
            __syncthreads();
            for (int e = t; e < PW_ROWS * PW_K; e += PW_ROWS) {
                const int r = e / PW_K, c = e % PW_K;
                double v = 0.0;
                if (r0 + r < n && c < kw) v = (double)X[(r0 + r) * d + k0 + c];
                xs[c][r] = v;
            }
            for (int e = t; e < PW_J * PW_K; e += PW_ROWS) {
                const int jj = e / PW_K, c = e % PW_K;
                double v = 0.0;
                if (j0 + jj < m && c < kw) v = (double)Y[(j0 + jj) * d + k0 + c];
                ys[c][jj] = v;
            }
            __syncthreads();
            for (int c = 0; c < kw; ++c) {
                const double x = xs[c][t];
#pragma unroll
                for (int jj = 0; jj < PW_J; ++jj) acc[jj] = acc[jj] + x * ys[c][jj];
            }
        }
        const int64_t jw = m - j0 < PW_J ? m - j0 : PW_J;
        for (int jj = 0; jj < jw; ++jj) {
            const int64_t j = j0 + jj;
            const double v = sqrt(fmax((xq + ysq[j]) - 2.0 * acc[jj], 0.0));
            if (FULL) {
                if (i < n) dist_out[i * m + j] = v;
            } else if (j == 0 || v < bv) {
                bv = v;
                bi = j;
            }
        }
    }
    if (!FULL && i < n) {
        idx_out[i] = bi;
        if (min_out) min_out[i] = bv;
    }
}

template <typename T>
void pw_argmin_run(const T *X, int64_t n, const T *Y, int64_t m, int d, int64_t *idx,
                   double *mins, double *dist, Workspace &ws, hipStream_t st) {
    double *xsq = ws.get<double>("pw_xsq", n);
    double *xs1 = ws.get<double>("pw_xs1", n);
    double *ysq = ws.get<double>("pw_ysq", m);
    double *ys1 = ws.get<double>("pw_ys1", m);
    hipLaunchKernelGGL(row_norms_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, st, X, n, d, xsq,
                       xs1);
    hipLaunchKernelGGL(row_norms_kernel<T>, dim3(grid1(m, 256)), dim3(256), 0, st, Y, m, d, ysq,
                       ys1);
    if (dist)
        hipLaunchKernelGGL((pw_argmin_kernel<T, true>), dim3(grid1(n, PW_ROWS)), dim3(PW_ROWS), 0,
                           st, X, n, Y, m, d, xsq, ysq, idx, mins, dist);
    else
        hipLaunchKernelGGL((pw_argmin_kernel<T, false>), dim3(grid1(n, PW_ROWS)), dim3(PW_ROWS), 0,
                           st, X, n, Y, m, d, xsq, ysq, idx, mins, dist);
    check_launch();
}

// pairwise.rs:77-117 (check_feature_dim, then check_non_empty), as reached
// through euclidean_distances.
fl_status pw_validate(int32_t dtype, int64_t n, int64_t dx, int64_t m, int64_t dy) {
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    if (n < 0 || m < 0 || dx < 0 || dy < 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: shapes: negative");
    if (dx != dy)
        return fail(FL_SHAPE_MISMATCH,
                    "ShapeMismatch: expected [" + std::to_string(n) + ", " + std::to_string(dx) +
                        "], actual [" + std::to_string(m) + ", " + std::to_string(dy) +
                        "]: euclidean_distances: feature dimensions must match (X has " +
                        std::to_string(dx) + " cols, Y has " + std::to_string(dy) + " cols)");
    if (n == 0 || m == 0)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required 1, actual 0: euclidean_distances: both X and "
                    "Y must have at least one row");
    if (dx == 0)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required 1, actual 0: euclidean_distances: feature "
                    "dimension must be at least 1");
    if (dx > (1 << 20))
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_features: must be in 1..2^20");
    return FL_OK;
}

fl_status pw_device(const void *dX, int32_t dtype, int64_t n, const void *dY, int64_t m, int d,
                    int64_t *d_idx, double *d_min, double *d_dist = nullptr) {
    StreamGuard sg;
    Workspace ws;
    if (dtype == FL_F32)
        pw_argmin_run<float>((const float *)dX, n, (const float *)dY, m, d, d_idx, d_min, d_dist,
                             ws, sg.s);
    else
        pw_argmin_run<double>((const double *)dX, n, (const double *)dY, m, d, d_idx, d_min,
                              d_dist, ws, sg.s);
    FL_HIP(hipStreamSynchronize(sg.s));
    return FL_OK;
}

}  // namespace

extern "C" {

fl_status fl_pairwise_distances_argmin_min_device(const void *dX, int32_t dtype, int64_t n,
                                                  int64_t dx, const void *dY, int64_t m,
                                                  int64_t dy, int64_t *d_idx_out,
                                                  double *d_min_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = pw_validate(dtype, n, dx, m, dy)) return s;
        if (!dX || !dY || !d_idx_out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        return pw_device(dX, dtype, n, dY, m, (int)dx, d_idx_out, d_min_out);
    });
}

fl_status fl_pairwise_distances_argmin_min(const void *X, int32_t dtype, int64_t n, int64_t dx,
                                           const void *Y, int64_t m, int64_t dy,
                                           int64_t *idx_out, double *min_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = pw_validate(dtype, n, dx, m, dy)) return s;
        if (!X || !Y || !idx_out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        const size_t es = dtype_size(dtype);
        DevBuf bx, by, bi, bm;
        void *dX = bx.get((size_t)(n * dx) * es);
        void *dY = by.get((size_t)(m * dy) * es);
        int64_t *di = (int64_t *)bi.get((size_t)n * 8);
        double *dm = min_out ? (double *)bm.get((size_t)n * 8) : nullptr;
        FL_HIP(hipMemcpy(dX, X, (size_t)(n * dx) * es, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(dY, Y, (size_t)(m * dy) * es, hipMemcpyHostToDevice));
        if (fl_status s = pw_device(dX, dtype, n, dY, m, (int)dx, di, dm)) return s;
        FL_HIP(hipMemcpy(idx_out, di, (size_t)n * 8, hipMemcpyDeviceToHost));
        if (min_out) FL_HIP(hipMemcpy(min_out, dm, (size_t)n * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

// euclidean_distances (pairwise.rs:157-189): the same validation, then
// every distance.
fl_status fl_euclidean_distances_device(const void *dX, int32_t dtype, int64_t n, int64_t dx,
                                        const void *dY, int64_t m, int64_t dy, double *d_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = pw_validate(dtype, n, dx, m, dy)) return s;
        if (!dX || !dY || !d_out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        return pw_device(dX, dtype, n, dY, m, (int)dx, nullptr, nullptr, d_out);
    });
}

fl_status fl_euclidean_distances(const void *X, int32_t dtype, int64_t n, int64_t dx,
                                 const void *Y, int64_t m, int64_t dy, double *out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = pw_validate(dtype, n, dx, m, dy)) return s;
        if (!X || !Y || !out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        const size_t es = dtype_size(dtype);
        DevBuf bx, by, bo;
        void *dX = bx.get((size_t)(n * dx) * es);
        void *dY = by.get((size_t)(m * dy) * es);
        double *dO = (double *)bo.get((size_t)(n * m) * 8);
        FL_HIP(hipMemcpy(dX, X, (size_t)(n * dx) * es, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(dY, Y, (size_t)(m * dy) * es, hipMemcpyHostToDevice));
        if (fl_status s = pw_device(dX, dtype, n, dY, m, (int)dx, nullptr, nullptr, dO)) return s;
        FL_HIP(hipMemcpy(out, dO, (size_t)(n * m) * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

}  // extern "C"
