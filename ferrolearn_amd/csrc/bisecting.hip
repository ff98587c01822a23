// bisecting.hip -- BisectingKMeans (ferrolearn-cluster/src/bisecting_kmeans.rs)
// on gfx950.
//
// The reference's loop runs on the host: pick the cluster to bisect (the
// largest, or the largest SSE; the last of equal candidates, as Rust's
// max_by_key / max_by return), run n_init 2-means bisections with StdRng
// seeded base + split * 1000 + run, keep the strictly smallest SSE, replace
// the cluster by its non-empty children.  The work over the samples runs
// on the device, every sum in the reference's order:
//   - assignment: per member, the two folds to the centres (acc + (a-b)^2,
//     left to right) and d0 <= d1 -> 0 else 1, with a changed flag;
//   - centres: each label's members in index order (a stable select), their
//     values gathered per feature, and the left fold of each sequence by the
//     exact workgroup fold, divided by the member count;
//   - SSE: the members' folds to their centre, left-folded in member order.
// So the labels, centres and inertia equal the reference's bit for bit
// (f32 samples widened to f64, as everywhere in this library).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "host.hpp"
#include "kmeans_kernels.hpp"

using namespace fl;

namespace {

// c[j] = X[row * d + j] widened.
template <typename TX>
__global__ void bk2_row_kernel(const TX *__restrict__ X, int64_t row, int d, double *__restrict__ c) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d) c[j] = (double)X[row * d + j];
}

// Assignment of the members idx[0..n): d0 <= d1 -> 0 else 1 (NaN -> 1).
template <typename TX>
__global__ void bk2_assign_kernel(const TX *__restrict__ X, int d, const int32_t *__restrict__ idx,
                                  int64_t n, const double *__restrict__ c0,
                                  const double *__restrict__ c1, uint8_t *__restrict__ lab,
                                  uint8_t *__restrict__ flag0, int32_t *__restrict__ changed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ch = false;
    if (i < n) {
        const TX *x = X + (int64_t)idx[i] * d;
        const double d0 = fold_exact(x, c0, d), d1 = fold_exact(x, c1, d);
        const uint8_t nl = d0 <= d1 ? 0 : 1;
        ch = nl != lab[i];
        lab[i] = nl;
        flag0[i] = nl == 0;
    }
    if (__ballot(ch) != 0 && (threadIdx.x & 63) == 0) atomicOr(changed, 1);
}

// buf[j * ld + k] = X[sel[k] * d + j] for k < *cnt: the members' values per
// feature, in member order.
template <typename TX>
__global__ void bk2_gather_kernel(const TX *__restrict__ X, int d, const int32_t *__restrict__ sel,
                                  const int32_t *__restrict__ cnt, int64_t ld,
                                  double *__restrict__ buf) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *cnt) return;
    const TX *x = X + (int64_t)sel[k] * d;
    for (int j = blockIdx.y; j < d; j += gridDim.y) buf[(int64_t)j * ld + k] = (double)x[j];
}

// flag[i] = (lab[i] == 0).
__global__ void bk2_flip_kernel(const uint8_t *__restrict__ lab, int64_t n,
                                uint8_t *__restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = lab[i] == 0;
}

// c[j] = sums[j] / count when count > 0 (a centre without members keeps its
// value, bisecting_kmeans.rs:292-303).
__global__ void bk2_centre_kernel(const double *__restrict__ sums, const int32_t *__restrict__ cnt,
                                  int d, double *__restrict__ c) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t n = *cnt;
    if (j < d && n > 0) c[j] = sums[j] / (double)n;
}

// dd[k] = fold(X[sel[k]], c) for k < *cnt (or every k < n when cnt is null).
template <typename TX>
__global__ void bk2_dist_kernel(const TX *__restrict__ X, int d, const int32_t *__restrict__ sel,
                                const int32_t *__restrict__ cnt, int64_t n,
                                const uint8_t *__restrict__ lab, const double *__restrict__ c0,
                                const double *__restrict__ c1, double *__restrict__ dd) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t m = cnt ? (int64_t)*cnt : n;
    if (k >= m) return;
    const TX *x = X + (int64_t)sel[k] * d;
    dd[k] = fold_exact(x, (lab && lab[k]) ? c1 : c0, d);
}

template <typename TX>
class Bisect {
  public:
    Bisect(const TX *X, int64_t n, int d, hipStream_t st) : X_(X), n_(n), d_(d), st_(st) {}

    struct Clu {
        std::vector<int32_t> idx;   // members, ascending
        DevBuf didx;
        std::vector<double> center;
        double sse = 0.0;
    };

    // Left fold of v[0..n) (device) into the host.
    double fold1(const double *v, int64_t n) {
        double *out = ws_.get<double>("fold_out", 1);
        left_fold(v, n, n, 1, nullptr, out, nullptr, ws_, st_);
        double h = 0.0;
        FL_HIP(hipMemcpyAsync(&h, out, 8, hipMemcpyDeviceToHost, st_));
        FL_HIP(hipStreamSynchronize(st_));
        return h;
    }

    // Centre of the selected members sel[0..*cnt) (count m on the host):
    // per feature the left fold of their values in order, / m.
    void centre_of(const int32_t *sel, const int32_t *dcnt, int64_t m, double *dc) {
        if (m <= 0) return;
        double *buf = ws_.get<double>("gather", (int64_t)d_ * m);
        double *sums = ws_.get<double>("sums", d_);
        const unsigned gy = (unsigned)std::min(d_, 64);
        hipLaunchKernelGGL(bk2_gather_kernel<TX>, dim3(grid1(m, 256), gy), dim3(256), 0, st_, X_,
                           d_, sel, dcnt, m, buf);
        left_fold(buf, m, m, d_, nullptr, sums, nullptr, ws_, st_);
        hipLaunchKernelGGL(bk2_centre_kernel, dim3(grid1(d_, 256)), dim3(256), 0, st_, sums,
                           dcnt, d_, dc);
        check_launch();
    }

    // Stable selection of the members whose flag is set; returns the count.
    int64_t select(const int32_t *didx, const uint8_t *flag, int64_t n, int32_t *out,
                   int32_t *dcnt) {
        size_t tb = 0;
        FL_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, didx, flag, out, dcnt, (int)n, st_));
        void *tmp = ws_.get<unsigned char>("sel_tmp", (int64_t)tb);
        FL_HIP(hipcub::DeviceSelect::Flagged(tmp, tb, didx, flag, out, dcnt, (int)n, st_));
        int32_t h = 0;
        FL_HIP(hipMemcpyAsync(&h, dcnt, 4, hipMemcpyDeviceToHost, st_));
        FL_HIP(hipStreamSynchronize(st_));
        return h;
    }

    // run_2means (bisecting_kmeans.rs:225-322) over cluster c; labels in lab
    // (device), centres in c0 / c1 (device), returns the total SSE.
    double two_means(const Clu &c, int64_t max_iter, StdRng09 &g, uint8_t *lab, double *c0,
                     double *c1) {
        const int64_t n = (int64_t)c.idx.size();
        const int64_t i0 = g.range_from(0, n);
        int64_t i1 = g.range_from(0, n);
        if (n > 1)
            while (i1 == i0) i1 = g.range_from(0, n);
        hipLaunchKernelGGL(bk2_row_kernel<TX>, dim3(grid1(d_, 256)), dim3(256), 0, st_, X_,
                           (int64_t)c.idx[(size_t)i0], d_, c0);
        hipLaunchKernelGGL(bk2_row_kernel<TX>, dim3(grid1(d_, 256)), dim3(256), 0, st_, X_,
                           (int64_t)c.idx[(size_t)i1], d_, c1);
        check_launch();
        FL_HIP(hipMemsetAsync(lab, 0, (size_t)n, st_));
        const int32_t *didx = (const int32_t *)c.didx.ptr();
        uint8_t *flag0 = ws_.get<uint8_t>("flag0", n);
        int32_t *changed = ws_.get<int32_t>("changed", 1);
        int32_t *sel = ws_.get<int32_t>("sel", n);
        int32_t *cnt = ws_.get<int32_t>("cnt", 2);
        for (int64_t it = 0; it < max_iter; ++it) {
            FL_HIP(hipMemsetAsync(changed, 0, 4, st_));
            hipLaunchKernelGGL(bk2_assign_kernel<TX>, dim3(grid1(n, 256)), dim3(256), 0, st_, X_,
                               d_, didx, n, c0, c1, lab, flag0, changed);
            check_launch();
            int32_t ch = 0;
            FL_HIP(hipMemcpyAsync(&ch, changed, 4, hipMemcpyDeviceToHost, st_));
            FL_HIP(hipStreamSynchronize(st_));
            if (!ch) break;
            const int64_t n0 = select(didx, flag0, n, sel, cnt);
            centre_of(sel, cnt, n0, c0);
            const int64_t n1 = select(didx, lab, n, sel, cnt + 1);
            centre_of(sel, cnt + 1, n1, c1);
        }
        double *dd = ws_.get<double>("dd", n);
        hipLaunchKernelGGL(bk2_dist_kernel<TX>, dim3(grid1(n, 256)), dim3(256), 0, st_, X_, d_,
                           didx, (const int32_t *)nullptr, n, lab, c0, c1, dd);
        check_launch();
        return fold1(dd, n);
    }

    // SSE of the selected members sel[0..m) against centre c (device).
    double sse_of(const int32_t *sel, int64_t m, const double *dc) {
        if (m <= 0) return 0.0;
        double *dd = ws_.get<double>("dd", m);
        hipLaunchKernelGGL(bk2_dist_kernel<TX>, dim3(grid1(m, 256)), dim3(256), 0, st_, X_, d_,
                           sel, (const int32_t *)nullptr, m, (const uint8_t *)nullptr, dc, dc, dd);
        check_launch();
        return fold1(dd, m);
    }

    fl_status fit(const fl_bisecting_params &p, std::vector<double> &centers,
                  std::vector<int64_t> &labels, double &inertia) {
        const size_t db = (size_t)d_ * 8;
        std::vector<std::unique_ptr<Clu>> cl;
        {
            // compute_cluster_stats of every sample (:192-223)
            auto c = std::make_unique<Clu>();
            c->idx.resize((size_t)n_);
            for (int64_t i = 0; i < n_; ++i) c->idx[(size_t)i] = (int32_t)i;
            int32_t *didx = (int32_t *)c->didx.get((size_t)n_ * 4);
            FL_HIP(hipMemcpyAsync(didx, c->idx.data(), (size_t)n_ * 4, hipMemcpyHostToDevice, st_));
            int32_t *cnt = ws_.get<int32_t>("cnt", 2);
            const int32_t nn = (int32_t)n_;
            FL_HIP(hipMemcpyAsync(cnt, &nn, 4, hipMemcpyHostToDevice, st_));
            double *dc = ws_.get<double>("cen", d_);
            centre_of(didx, cnt, n_, dc);
            c->center.resize((size_t)d_);
            FL_HIP(hipMemcpyAsync(c->center.data(), dc, db, hipMemcpyDeviceToHost, st_));
            c->sse = sse_of(didx, n_, dc);
            cl.push_back(std::move(c));
        }
        const uint64_t base = p.has_random_state ? p.random_state : 0;
        uint64_t split = 0;
        int64_t stall = 0;
        double *c0 = ws_.get<double>("c0", d_), *c1 = ws_.get<double>("c1", d_);
        double *b0 = ws_.get<double>("b0", d_), *b1 = ws_.get<double>("b1", d_);
        while ((int64_t)cl.size() < p.n_clusters) {
            int64_t t = -1;
            for (int64_t i = 0; i < (int64_t)cl.size(); ++i) {
                const Clu &c = *cl[(size_t)i];
                if (c.idx.size() < 2) continue;
                if (t < 0) {
                    t = i;
                } else if (p.strategy == FL_BISECT_LARGEST_CLUSTER) {
                    if (!(cl[(size_t)t]->idx.size() > c.idx.size())) t = i;
                } else if (!(cl[(size_t)t]->sse > c.sse)) {
                    t = i;
                }
            }
            if (t < 0) break;   // no cluster of two or more samples
            const Clu &tc = *cl[(size_t)t];
            const int64_t tn = (int64_t)tc.idx.size();
            uint8_t *lab = ws_.get<uint8_t>("lab", tn);
            uint8_t *blab = ws_.get<uint8_t>("blab", tn);
            double best = 1.7976931348623157e308;
            bool have = false;
            for (int64_t run = 0; run < p.n_init; ++run) {
                StdRng09 g(base + split * 1000u + (uint64_t)run);
                const double s = two_means(tc, p.max_iter, g, lab, c0, c1);
                if (s < best) {
                    best = s;
                    have = true;
                    FL_HIP(hipMemcpyAsync(blab, lab, (size_t)tn, hipMemcpyDeviceToDevice, st_));
                    FL_HIP(hipMemcpyAsync(b0, c0, db, hipMemcpyDeviceToDevice, st_));
                    FL_HIP(hipMemcpyAsync(b1, c1, db, hipMemcpyDeviceToDevice, st_));
                }
            }
            ++split;
            if (!have)
                return fail(FL_NUMERICAL_INSTABILITY,
                            "NumericalInstability: BisectingKMeans: no 2-means run of the "
                            "bisection gave an SSE below f64::MAX");
            // children in member order (:430-470)
            const int32_t *didx = (const int32_t *)tc.didx.ptr();
            uint8_t *flag0 = ws_.get<uint8_t>("flag0", tn);
            hipLaunchKernelGGL(bk2_flip_kernel, dim3(grid1(tn, 256)), dim3(256), 0, st_, blab, tn,
                               flag0);
            check_launch();
            int32_t *cnt = ws_.get<int32_t>("cnt", 2);
            auto k0 = std::make_unique<Clu>(), k1 = std::make_unique<Clu>();
            int32_t *s0 = (int32_t *)k0->didx.get((size_t)tn * 4);
            int32_t *s1 = (int32_t *)k1->didx.get((size_t)tn * 4);
            const int64_t n0 = select(didx, flag0, tn, s0, cnt);
            const int64_t n1 = select(didx, blab, tn, s1, cnt + 1);
            k0->idx.resize((size_t)n0);
            k1->idx.resize((size_t)n1);
            if (n0) FL_HIP(hipMemcpyAsync(k0->idx.data(), s0, (size_t)n0 * 4, hipMemcpyDeviceToHost, st_));
            if (n1) FL_HIP(hipMemcpyAsync(k1->idx.data(), s1, (size_t)n1 * 4, hipMemcpyDeviceToHost, st_));
            k0->center.resize((size_t)d_);
            k1->center.resize((size_t)d_);
            FL_HIP(hipMemcpyAsync(k0->center.data(), b0, db, hipMemcpyDeviceToHost, st_));
            FL_HIP(hipMemcpyAsync(k1->center.data(), b1, db, hipMemcpyDeviceToHost, st_));
            k0->sse = sse_of(s0, n0, b0);
            k1->sse = sse_of(s1, n1, b1);
            FL_HIP(hipStreamSynchronize(st_));
            cl.erase(cl.begin() + t);
            if (n0) cl.push_back(std::move(k0));
            if (n1) cl.push_back(std::move(k1));
            if (n0 == 0 || n1 == 0) {
                // the reference keeps drawing new seeds for the same cluster
                if (++stall >= kMaxStall)
                    return fail(FL_NUMERICAL_INSTABILITY,
                                "NumericalInstability: BisectingKMeans: " +
                                    std::to_string(kMaxStall) +
                                    " bisections in a row left one side empty (identical "
                                    "samples?); the reference's loop would not end");
            } else {
                stall = 0;
            }
        }
        const int64_t k = (int64_t)cl.size();
        centers.assign((size_t)(k * d_), 0.0);
        labels.assign((size_t)n_, 0);
        inertia = 0.0;
        for (int64_t c = 0; c < k; ++c) {
            std::copy(cl[(size_t)c]->center.begin(), cl[(size_t)c]->center.end(),
                      centers.begin() + c * d_);
            for (int32_t i : cl[(size_t)c]->idx) labels[(size_t)i] = c;
            inertia = inertia + cl[(size_t)c]->sse;
        }
        return FL_OK;
    }

    static constexpr int64_t kMaxStall = 1000;

  private:
    const TX *X_;
    int64_t n_;
    int d_;
    hipStream_t st_;
    Workspace ws_;
};

// :326-366, in the reference's order.
fl_status bisect_validate(const fl_bisecting_params *p, int64_t n, int64_t d, int32_t dtype) {
    if (!p) return fail(FL_INVALID_PARAMETER, "InvalidParameter: params: must not be NULL");
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    if (p->n_clusters <= 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_clusters: must be at least 1");
    if (p->n_init <= 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_init: must be at least 1");
    if (p->max_iter < 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: max_iter: must not be negative");
    if (p->strategy != FL_BISECT_LARGEST_CLUSTER && p->strategy != FL_BISECT_LARGEST_SSE)
        return fail(FL_INVALID_PARAMETER,
                    "InvalidParameter: strategy: must be FL_BISECT_LARGEST_CLUSTER or "
                    "FL_BISECT_LARGEST_SSE");
    if (n < p->n_clusters)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(p->n_clusters) +
                        ", actual " + std::to_string(n < 0 ? 0 : n) +
                        ": BisectingKMeans requires at least n_clusters samples");
    if (d <= 0 || d > (1 << 20))
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_features: must be in 1..2^20");
    if (n > INT32_MAX)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_samples: at most 2^31 - 1");
    return FL_OK;
}

fl_status bisect_run(const void *dX, int32_t dtype, int64_t n, int64_t d,
                     const fl_bisecting_params *p, std::vector<double> &centers,
                     std::vector<int64_t> &labels, double &inertia) {
    StreamGuard sg;
    fl_status s;
    if (dtype == FL_F32) {
        Bisect<float> b((const float *)dX, n, (int)d, sg.s);
        s = b.fit(*p, centers, labels, inertia);
    } else {
        Bisect<double> b((const double *)dX, n, (int)d, sg.s);
        s = b.fit(*p, centers, labels, inertia);
    }
    FL_HIP(hipStreamSynchronize(sg.s));
    return s;
}

}  // namespace

extern "C" {

void fl_bisecting_params_default(fl_bisecting_params *p, int64_t n_clusters) {
    std::memset(p, 0, sizeof(*p));
    p->n_clusters = n_clusters;   // bisecting_kmeans.rs:93-102
    p->max_iter = 300;
    p->n_init = 10;
    p->has_random_state = 0;
    p->strategy = FL_BISECT_LARGEST_CLUSTER;
}

fl_status fl_bisecting_kmeans_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                                  const fl_bisecting_params *params, double *centers_out,
                                  int64_t *labels_out, double *inertia_out,
                                  int64_t *n_clusters_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = bisect_validate(params, n, d, dtype)) return s;
        if (!X || !centers_out || !labels_out || !inertia_out || !n_clusters_out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        DevBuf bx;
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        std::vector<double> centers;
        std::vector<int64_t> labels;
        double inertia = 0.0;
        if (fl_status s = bisect_run(dX, dtype, n, d, params, centers, labels, inertia)) return s;
        std::memcpy(centers_out, centers.data(), centers.size() * 8);
        std::memcpy(labels_out, labels.data(), labels.size() * 8);
        *inertia_out = inertia;
        *n_clusters_out = (int64_t)(centers.size() / (size_t)d);
        return FL_OK;
    });
}

fl_status fl_bisecting_kmeans_fit_device(const void *dX, int32_t dtype, int64_t n, int64_t d,
                                         const fl_bisecting_params *params,
                                         double *d_centers_out, int64_t *d_labels_out,
                                         double *inertia_out, int64_t *n_clusters_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = bisect_validate(params, n, d, dtype)) return s;
        if (!dX || !d_centers_out || !d_labels_out || !inertia_out || !n_clusters_out)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        std::vector<double> centers;
        std::vector<int64_t> labels;
        double inertia = 0.0;
        if (fl_status s = bisect_run(dX, dtype, n, d, params, centers, labels, inertia)) return s;
        FL_HIP(hipMemcpy(d_centers_out, centers.data(), centers.size() * 8, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(d_labels_out, labels.data(), labels.size() * 8, hipMemcpyHostToDevice));
        *inertia_out = inertia;
        *n_clusters_out = (int64_t)(centers.size() / (size_t)d);
        return FL_OK;
    });
}

}  // extern "C"
