// kmeans.hip -- KMeans fit / predict / transform on gfx950.
//
// Reference: ferrolearn-cluster/src/kmeans.rs.  Fit (:403-513) runs n_init
// Lloyd runs of assign (:267-352) -> recompute centroids (:358-401) -> swap
// -> n_iter = iter + 1 -> stop when max_shift < tol, and keeps the run with
// the strictly smallest inertia.  Labels and inertia come from the last
// assignment, the centres are the post-update ones (:468-495).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <type_traits>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "host.hpp"
#include "kmeans_kernels.hpp"

using namespace fl;

// Fixed device scratch of fl_kmeans_predict_device_noalloc.  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in, never the device memory: kernels and
// copies only ever get the address hipGetSymbolAddress returns for the
// current device (scratch_address below).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Largest cluster count (labels are int32; the filters' tiles index
// centres below 2^31 - 64).
constexpr int64_t kMaxClusters = ((int64_t)1 << 31) - 65;
// Up to this many clusters the centroid update counts members in LDS.
constexpr int64_t kLdsClusters = 8192;
constexpr int kKmM = 4;   // candidates kept per row by the filter

struct KmAcc {
    double ms_dist = 0, ms_fin = 0, ms_fb = 0, ms_upd = 0;
    int64_t n_rows = 0, n_cand = 0, n_fb = 0, n_launch = 0;
    int64_t n_second = 0;   // rows the first filter (FP8) left to the BF16x3 one
};

template <typename TX>
__global__ void copy_row_f64_kernel(const TX *__restrict__ X, int64_t row, int d,
                                    double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d) out[j] = (double)X[row * d + j];
}

__global__ void labels_to_i64_kernel(const int32_t *__restrict__ a, int64_t n,
                                     int64_t *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// One KMeans problem on the device: X stays resident, scratch is reused.
template <typename TX>
class Lloyd {
  public:
    // scratch / scratch_bytes: a fixed buffer that every device buffer of
    // this object is carved from (nothing is allocated; phase timing is off).
    // C_range: centres to include in the range check (predict, where the
    // centres are not means of the samples).
    Lloyd(const TX *X, int64_t n, int d, int64_t k, int32_t precision, hipStream_t st,
          void *scratch = nullptr, size_t scratch_bytes = 0, const double *C_range = nullptr)
        : X_(X), n_(n), d_(d), k_(k), st_(st), ws_(scratch, scratch_bytes) {
        timing_ = scratch == nullptr;
        // The direct exact scan is cheaper than filter + re-score when there
        // are few centres.
        direct_ = precision == FL_PREC_EXACT || k <= 32;
        if (!direct_) {
            xn_ = ws_.get<double>("xn", n);
            xs_ = ws_.get<double>("xs", n);
            hipLaunchKernelGGL(row_norms_kernel<TX>, dim3(grid1(n, 256)), dim3(256), 0, st, X, n,
                               d, xn_, xs_);
            check_launch();
        }
        // Range of the samples (the centres, means of samples, stay inside
        // it): the BF16 bound needs finite values below 2^56 in magnitude,
        // the f64 filter below 2^250; other data uses the f64 GEMM form.
        if (!direct_) {
            unsigned int *flag = ws_.get<unsigned int>("range", 1);
            FL_HIP(hipMemsetAsync(flag, 0, 4, st));
            hipLaunchKernelGGL(range_flags_kernel<TX>, dim3(1024), dim3(256), 0, st, X, n * d,
                               flag);
            if (C_range)
                hipLaunchKernelGGL(range_flags_kernel<double>, dim3(64), dim3(256), 0, st, C_range,
                                   k * d, flag);
            check_launch();
            unsigned int hflag = 0;
            FL_HIP(hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, st));
            FL_HIP(hipStreamSynchronize(st));
            const bool fits = k < ((int64_t)1 << 31) - 64;
            // BF16 filter: samples are split into hi / lo bf16 once per fit.
            // The FP8 form (d <= 128) filters with E4M3 first and sends the
            // rows it cannot certify through the BF16x3 filter.
            if ((precision == FL_PREC_BF16 || precision == FL_PREC_BF16X3 ||
                 precision == FL_PREC_FP8) &&
                bf16_nkc(d) > 0 && fits && (hflag & 1u) == 0) {
                passes_ = precision == FL_PREC_BF16 ? 1 : 3;
                nkc_ = bf16_nkc(d);
                const int dpad = nkc_ * 32;
                xh_ = ws_.get<unsigned short>("xh", n * dpad);
                xl_ = ws_.get<unsigned short>("xl", n * dpad);
                hipLaunchKernelGGL(bf16_split_kernel<TX>, dim3(grid1(n * dpad, 256)), dim3(256), 0,
                                   st, X, n, d, dpad, xh_, xl_);
                check_launch();
                if (precision == FL_PREC_FP8 && d <= 128) {
                    fp8_ = true;
                    x8_ = ws_.get<uint8_t>("x8", n * 128);
                    sx_ = ws_.get<int32_t>("sx", n);
                    hipLaunchKernelGGL(fp8_split_rows_kernel<TX>, dim3((unsigned)ceil_div(n, 4)),
                                       dim3(256), 0, st, X, n, d, x8_, sx_);
                    check_launch();
                }
            } else if (f64_d4(d) > 0 && fits && (hflag & 2u) == 0) {
                d4_ = f64_d4(d);   // f64 MFMA argmin filter
            }
        }
    }

    bool bf16() const { return passes_ > 0; }

    // labels/best for every row against centres C (k*d f64, device).
    // Phase timers are read after the caller's next synchronisation
    // (flush), so that timing adds no host round trip to an iteration.
    int tbegin() {
        if (!timing_) return -1;
        if (!tm_) tm_.reset(new PhaseTimer(st_));
        return tm_->begin();
    }
    void tend(int id, double *slot) {
        if (id < 0) return;
        tm_->end(id);
        pend_.push_back({id, slot});
    }
    // Add the pending phase times to their slots and, when given, the
    // fallback count of the last assignment (device) to *n_fb.
    // With n_second, also the rows the FP8 filter sent to the BF16x3 one.
    void flush(int64_t *n_fb = nullptr, int64_t *n_second = nullptr) {
        if (n_fb) {
            int32_t c[2] = {0, 0};
            copy_counts(c);
            FL_HIP(hipStreamSynchronize(st_));
            *n_fb += c[0];
            if (n_second) *n_second += c[1];
        }
        flush_upto(pend_.size());
    }
    // Enqueue copies of the last assignment's counts into h[0] (rows of the
    // exact scan) and h[1] (rows the FP8 filter left; untouched otherwise).
    // h must be pinned host memory for the copies to stay asynchronous.
    void copy_counts(int32_t *h) {
        FL_HIP(hipMemcpyAsync(&h[0], ws_.get<int32_t>("fcnt", 1), 4, hipMemcpyDeviceToHost, st_));
        if (fp8_)
            FL_HIP(hipMemcpyAsync(&h[1], ws_.get<int32_t>("fcnt8", 1), 4, hipMemcpyDeviceToHost,
                                  st_));
    }
    // Pending phase timers: mark() names the position after the phases
    // enqueued so far; flush_upto(m) adds the first m to their slots (waiting
    // only for those), drop_from(m) forgets those after m.
    size_t mark() const { return pend_.size(); }
    void flush_upto(size_t m) {
        for (size_t i = 0; i < m; ++i) {
            *pend_[i].second += tm_->ms(pend_[i].first);
            tm_->release(pend_[i].first);
        }
        pend_.erase(pend_.begin(), pend_.begin() + (std::ptrdiff_t)m);
    }
    void drop_from(size_t m) {
        for (size_t i = m; i < pend_.size(); ++i) tm_->release(pend_[i].first);
        pend_.resize(m);
    }

    void assign(const double *C, int32_t *labels, double *best, KmAcc &acc) {
        acc.n_rows += n_;
        if (direct_) {
            const int t = tbegin();
            hipLaunchKernelGGL(km_assign_exact_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                               X_, C, n_, d_, k_, labels, best);
            check_launch();
            tend(t, &acc.ms_dist);
            FL_HIP(hipMemsetAsync(ws_.get<int32_t>("fcnt", 1), 0, sizeof(int32_t), st_));
            acc.n_launch += 1;
            return;
        }
        double *cn = ws_.get<double>("cn", k_);
        double *cs = ws_.get<double>("cs", k_);
        hipLaunchKernelGGL(row_norms_kernel<double>, dim3(grid1(k_, 256)), dim3(256), 0, st_, C,
                           k_, d_, cn, cs);
        check_launch();
        Cand cand;
        cand.key = ws_.get<double>("ckey", n_ * kKmM);
        cand.idx = ws_.get<int32_t>("cidx", n_ * kKmM);
        cand.cnt = ws_.get<int32_t>("ccnt", n_);
        cand.theta = ws_.get<double>("cth", n_);
        cand.theta_idx = ws_.get<int32_t>("cthi", n_);
        int t0;
        bool fused = false;   // the filter kernel also took the exact decision
        if (passes_ > 0) {
            // Centre split, centre terms and the largest centre norm change
            // every iteration; the samples' split was made once.
            const int dpad = nkc_ * 32;
            const Bound bd = bound_bf16(d_, passes_);
            unsigned short *ch = ws_.get<unsigned short>("ch", k_ * dpad);
            unsigned short *cl = ws_.get<unsigned short>("cl", k_ * dpad);
            float *cxnp = ws_.get<float>("cxnp", k_);
            unsigned long long *smax = ws_.get<unsigned long long>("smax", 1);
            double *qoff = ws_.get<double>("qoff", n_);
            hipLaunchKernelGGL(bf16_split_kernel<double>, dim3(grid1(k_ * dpad, 256)), dim3(256), 0,
                               st_, C, k_, d_, dpad, ch, cl);
            hipLaunchKernelGGL(bf16_stream_terms_kernel, dim3(grid1(k_, 256)), dim3(256), 0, st_,
                               cn, k_, bd.beta, cxnp);
            FL_HIP(hipMemsetAsync(smax, 0, 8, st_));
            hipLaunchKernelGGL(max_nonneg_kernel, dim3(16), dim3(256), 0, st_, cs, k_, smax);
            hipLaunchKernelGGL(bf16_resident_terms_kernel, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                               xn_, xs_, n_, d_, bd.alpha, bd.beta, (const double *)smax, qoff);
            check_launch();
            int32_t *flags8 = nullptr, *fcnt8 = nullptr;
            if (fp8_) {
                // FP8 filter with the decision fused over every row; the rows
                // it cannot certify are listed in flags8 for the BF16x3 pass.
                const Bound b8 = bound_fp8(d_);
                uint8_t *c8 = ws_.get<uint8_t>("c8", k_ * 128);
                int32_t *sc = ws_.get<int32_t>("sc", 1);
                unsigned long long *cmax = ws_.get<unsigned long long>("cmax", 1);
                float *cxnp8 = ws_.get<float>("cxnp8", k_);
                float *afac = ws_.get<float>("afac", n_);
                double *qoff8 = ws_.get<double>("qoff8", n_);
                flags8 = ws_.get<int32_t>("flags8", n_);
                fcnt8 = ws_.get<int32_t>("fcnt8", 1);
                FL_HIP(hipMemsetAsync(cmax, 0, 8, st_));
                hipLaunchKernelGGL(max_abs_bits_kernel, dim3(64), dim3(256), 0, st_, C, k_ * d_,
                                   cmax);
                hipLaunchKernelGGL(fp8_common_exp_kernel, dim3(1), dim3(64), 0, st_,
                                   (const unsigned long long *)cmax, sc);
                hipLaunchKernelGGL(fp8_split_common_kernel, dim3(grid1(k_ * 128, 256)), dim3(256),
                                   0, st_, C, k_, d_, (const int32_t *)sc, c8);
                hipLaunchKernelGGL(bf16_stream_terms_kernel, dim3(grid1(k_, 256)), dim3(256), 0,
                                   st_, cn, k_, b8.beta, cxnp8);
                hipLaunchKernelGGL(fp8_resident_factor_kernel, dim3(grid1(n_, 256)), dim3(256), 0,
                                   st_, (const int32_t *)sx_, (const int32_t *)sc, n_, afac);
                hipLaunchKernelGGL(fp8_resident_terms_kernel, dim3(grid1(n_, 256)), dim3(256), 0,
                                   st_, xn_, xs_, n_, d_, b8.alpha, b8.beta, (const double *)smax,
                                   qoff8);
                check_launch();
                Bf16Args a8;
                a8.nA = n_;
                a8.nB = k_;
                a8.b_per_split = k_;
                a8.ah = a8.al = (const unsigned short *)x8_;
                a8.bh = a8.bl = (const unsigned short *)c8;
                a8.qoff = qoff8;
                a8.xnp = cxnp8;
                a8.afac = afac;
                a8.out = cand;
                a8.dec_X = X_;
                a8.dec_x_f32 = sizeof(TX) == 4;
                a8.d = d_;
                a8.dec_C = C;
                a8.labels = labels;
                a8.best = best;
                a8.flags = flags8;
                a8.fcount = fcnt8;
                FL_HIP(hipMemsetAsync(fcnt8, 0, sizeof(int32_t), st_));
                const int t8 = tbegin();
                if (k_ <= FR_MAXK)
                    launch_fp8_res(a8, st_);   // every centre resident in LDS
                else
                    launch_bf16(0, 0, 2, a8, 1, st_);
                tend(t8, &acc.ms_dist);
                acc.n_launch += 1;
            }
            Bf16Args a;
            a.nA = n_;
            a.rows = flags8;          // FP8 form: only the rows FP8 left
            a.rows_count = fcnt8;
            a.nB = k_;
            a.b_per_split = k_;
            a.ah = xh_;
            a.al = xl_;
            a.bh = ch;
            a.bl = cl;
            a.qoff = qoff;
            a.xnp = cxnp;
            a.out = cand;
            a.dec_X = X_;
            a.dec_x_f32 = sizeof(TX) == 4;
            a.d = d_;
            a.dec_C = C;
            a.labels = labels;
            a.best = best;
            a.flags = ws_.get<int32_t>("flags", n_);
            a.fcount = ws_.get<int32_t>("fcnt", 1);
            FL_HIP(hipMemsetAsync(a.fcount, 0, sizeof(int32_t), st_));
            t0 = tbegin();
            if (!a.rows && passes_ == 3 && nkc_ <= 4) {
                // every row: the centres resident in LDS
                BfResArgs ra;
                ra.a = a;
                ra.pk = ws_.get<float>("res_pk", n_);
                ra.psec = ws_.get<float>("res_ps", n_);
                ra.pidx = ws_.get<int32_t>("res_pi", n_);
                launch_bf16_res(nkc_, ra, st_);
            } else {
                launch_bf16(passes_, 0, nkc_, a, 1, st_);   // argmin form, decision fused
            }
            tend(t0, fp8_ ? &acc.ms_fin : &acc.ms_dist);
            fused = true;
        } else if (d4_ > 0) {
            // f64 filter: centre terms, padding and the largest centre norm
            // change every iteration.
            const int dp = 4 * d4_;
            const Bound bd = bound_fp64(d_);
            const double *cB = C;
            if (dp != d_) {
                double *cp = ws_.get<double>("cpad", k_ * dp);
                hipLaunchKernelGGL(pad_rows_kernel<double>, dim3(grid1(k_ * dp, 256)), dim3(256),
                                   0, st_, C, k_, d_, dp, cp);
                cB = cp;
            }
            double *cxnp = ws_.get<double>("cxnp64", k_);
            unsigned long long *smax = ws_.get<unsigned long long>("smax", 1);
            double *qoff = ws_.get<double>("qoff", n_);
            hipLaunchKernelGGL(f64_stream_terms_kernel, dim3(grid1(k_, 256)), dim3(256), 0, st_,
                               cn, k_, bd.beta, cxnp);
            FL_HIP(hipMemsetAsync(smax, 0, 8, st_));
            hipLaunchKernelGGL(max_nonneg_kernel, dim3(16), dim3(256), 0, st_, cs, k_, smax);
            hipLaunchKernelGGL(f64_resident_terms_kernel, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                               xn_, xs_, n_, d_, bd.alpha, bd.beta, (const double *)smax, qoff);
            check_launch();
            F64Args a;
            a.nA = n_;
            a.nB = k_;
            a.d = d_;
            a.b_per_split = k_;
            a.A = X_;
            a.B = cB;
            a.qoff = qoff;
            a.xnp = cxnp;
            a.out = cand;
            t0 = tbegin();
            launch_f64_km<TX>(d4_, a, st_);   // argmin form: 4 candidates per row
            tend(t0, &acc.ms_dist);
        } else {
            TopmArgs a;
            a.nA = n_;
            a.nB = k_;
            a.d = d_;
            a.b_per_split = k_;
            a.an = xn_;
            a.as = xs_;
            a.bn = cn;
            a.bs = cs;
            const Bound bd = bound_fp64(d_);
            a.alpha = bd.alpha;
            a.beta = bd.beta;
            a.sqrt_key = 0;
            a.out = cand;
            t0 = tbegin();
            launch_topm<TX, double>(kKmM, MODE_MFMA64, X_, C, a, 1, st_);
            tend(t0, &acc.ms_dist);
        }
        acc.n_launch += 1;
        int32_t *flags = ws_.get<int32_t>("flags", n_);
        int32_t *fcnt = ws_.get<int32_t>("fcnt", 1);
        if (!fused) {
            FL_HIP(hipMemsetAsync(fcnt, 0, sizeof(int32_t), st_));
            const int t1 = tbegin();
            hipLaunchKernelGGL(km_finalize_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_, X_,
                               C, n_, d_, k_, kKmM, cand, labels, best, flags, fcnt);
            check_launch();
            tend(t1, &acc.ms_fin);
        }
        acc.n_cand += n_ * (int64_t)(fused ? 1 : (k_ < kKmM ? k_ : kKmM));
        // exact scan of the uncertified rows; the grid strides over the
        // device-side count (zero rows: every wave exits at once)
        const int t2 = tbegin();
        if (d_ <= 512) {
            // eight rows per block share each centre read
            constexpr int R = 8;
            double *Ct = ws_.get<double>("ctrans", k_ * d_);
            hipLaunchKernelGGL(km_transpose_kernel, dim3(grid1(k_ * d_, 256)), dim3(256), 0, st_,
                               C, k_, d_, Ct);
            const size_t sm = (size_t)R * d_ * 8;
            FL_HIP(hipFuncSetAttribute((const void *)km_fallback_rows_kernel<TX, R>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
            hipLaunchKernelGGL((km_fallback_rows_kernel<TX, R>),
                               dim3((unsigned)std::min<int64_t>(ceil_div(n_, R), 8192)),
                               dim3(256), sm, st_, X_, Ct, d_, k_, flags, fcnt, labels, best);
        } else if (d_ <= 8192) {
            double *Ct = ws_.get<double>("ctrans", k_ * d_);
            hipLaunchKernelGGL(km_transpose_kernel, dim3(grid1(k_ * d_, 256)), dim3(256), 0, st_,
                               C, k_, d_, Ct);
            const size_t sm = (size_t)d_ * 8;
            FL_HIP(hipFuncSetAttribute((const void *)km_fallback_lds_kernel<TX>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
            hipLaunchKernelGGL(km_fallback_lds_kernel<TX>,
                               dim3((unsigned)std::min<int64_t>(n_, 8192)), dim3(256), sm, st_, X_,
                               Ct, d_, k_, flags, fcnt, labels, best);
        } else {
            hipLaunchKernelGGL(km_fallback_kernel<TX>,
                               dim3((unsigned)std::min<int64_t>(grid1(n_, 4), 2048)), dim3(256), 0,
                               st_, X_, C, d_, k_, flags, fcnt, labels, best);
        }
        check_launch();
        tend(t2, &acc.ms_fb);
    }

    // Sum of best[0..n) into *out (device): on the reference's serial path
    // its row-order left fold, bit for bit; otherwise 1024-row blocks, a
    // fixed order (the reference's rayon order is unspecified there).
    void inertia(const double *best, double *out) {
        if (n_ * d_ < KM_SERIAL_WORK) {
            left_fold(best, n_, n_, 1, nullptr, out, nullptr, ws_, st_);
            return;
        }
        const int64_t nb = ceil_div(n_, SUM_BLK);
        double *part = ws_.get<double>("part", nb);
        hipLaunchKernelGGL(block_sum_kernel, dim3((unsigned)nb), dim3(256), 0, st_, best, n_, part);
        hipLaunchKernelGGL(runs_sum_kernel, dim3(1), dim3(64), 0, st_, part, nb, out);
        check_launch();
    }

    // Reference-order centroid update; shift (device) = max centre movement.
    // The member order (rows of each cluster in ascending index) comes from
    // per-block counts in LDS up to kLdsClusters clusters, and from a stable
    // radix sort of the row indices by label above that.
    void update(const int32_t *labels, const double *oldc, double *newc, double *shift,
                KmAcc &acc) {
        const int t = tbegin();
        int32_t *count = ws_.get<int32_t>("count", k_);
        int64_t *offset = ws_.get<int64_t>("offset", k_);
        int32_t *order = ws_.get<int32_t>("order", n_);
        if (k_ > kLdsClusters) {
            int32_t *ids = ws_.get<int32_t>("row_ids", n_);
            uint32_t *keys = ws_.get<uint32_t>("labels_sorted", n_);
            hipLaunchKernelGGL(km_fill_i32_kernel, dim3(grid1(k_, 256)), dim3(256), 0, st_, count,
                               k_, 0);
            hipLaunchKernelGGL(km_count_ids_kernel, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                               labels, n_, count, ids);
            hipLaunchKernelGGL(km_offsets_kernel, dim3(1), dim3(256), 0, st_, count, k_, offset);
            size_t tb = 0;
            const uint32_t *lk = reinterpret_cast<const uint32_t *>(labels);
            FL_HIP(rocprim::radix_sort_pairs(nullptr, tb, lk, keys, ids, order, (size_t)n_, 0,
                                             (unsigned)label_bits(k_), st_));
            void *tmp = ws_.get<unsigned char>("sort_tmp", (int64_t)tb);
            FL_HIP(rocprim::radix_sort_pairs(tmp, tb, lk, keys, ids, order, (size_t)n_, 0,
                                             (unsigned)label_bits(k_), st_));
        } else {
            const int64_t nblk = ceil_div(n_, UPD_RB);
            int32_t *blockcnt = ws_.get<int32_t>("blockcnt", nblk * k_);
            int32_t *blockbase = ws_.get<int32_t>("blockbase", nblk * k_);
            const size_t hist_b = (size_t)k_ * 4, ctr_b = (size_t)k_ * 8;
            FL_HIP(hipFuncSetAttribute((const void *)km_block_hist_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)hist_b));
            FL_HIP(hipFuncSetAttribute((const void *)km_scatter_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctr_b));
            hipLaunchKernelGGL(km_block_hist_kernel, dim3((unsigned)nblk), dim3(256), hist_b, st_,
                               labels, n_, k_, blockcnt);
            hipLaunchKernelGGL(km_col_scan_kernel, dim3((unsigned)k_), dim3(256), 0, st_,
                               blockcnt, nblk, k_, blockbase, count);
            hipLaunchKernelGGL(km_offsets_kernel, dim3(1), dim3(256), 0, st_, count, k_, offset);
            hipLaunchKernelGGL(km_scatter_kernel, dim3((unsigned)nblk), dim3(64), ctr_b, st_,
                               labels, n_, k_, label_bits(k_), blockbase, offset, order);
        }
        hipLaunchKernelGGL(km_sum_kernel<TX>, dim3((unsigned)k_, (unsigned)ceil_div(d_, 64)),
                           dim3(64), 0, st_, X_, d_, order, offset, count, oldc, newc);
        unsigned long long *mx = ws_.get<unsigned long long>("shift_max", 1);
        FL_HIP(hipMemsetAsync(mx, 0, 8, st_));
        hipLaunchKernelGGL(km_shift_kernel, dim3(grid1(k_, 64)), dim3(64), 0, st_, newc, oldc, k_,
                           d_, mx);
        hipLaunchKernelGGL(km_shift_finish, dim3(1), dim3(64), 0, st_, mx, shift);
        check_launch();
        tend(t, &acc.ms_upd);
    }

    // The library's k-means++ into out (k*d f64, device).
    void kmeanspp(uint64_t seed, double *out) {
        Rng rng(seed);
        kmeanspp(rng, out);
    }
    // k-means++ with every trial of a centre in one pass over X (the same
    // draws, sums, searches and choices as the per-trial passes below):
    // block-interleaved distance arrays, one array per trial plus md.
    void kmeanspp_batched(Rng &rng, double *out, int64_t nt) {
        const int64_t nblk = ceil_div(n_, KPP_BLK);
        const int64_t stride = nblk * KPP_BLK;
        double *pool = ws_.get<double>("kpp_pool", (nt + 1) * stride);
        double *bs = ws_.get<double>("kpp_bs", (nt + 1) * nblk);
        double *tot = ws_.get<double>("kpp_tot", nt + 1);
        double *thr = ws_.get<double>("kpp_thr", nt);
        int64_t *cand = ws_.get<int64_t>("kpp_cand", nt);
        double *ndarr = ws_.get<double>("kpp_nd", nt * stride);
        double *ndbs = ws_.get<double>("kpp_ndbs", nt * nblk);
        double *md = pool;        // current min distances, interleaved
        double *mdbs = bs;        // their block sums
        hipLaunchKernelGGL(kpp_first_il_kernel<TX>, dim3(grid1(stride, 256)), dim3(256), 0, st_,
                           X_, n_, d_, out, nblk, md);
        hipLaunchKernelGGL(kpp_bsums_il_kernel, dim3(grid1(nblk, 256)), dim3(256), 0, st_, md,
                           nblk, 1, mdbs);
        hipLaunchKernelGGL(kpp_totals_kernel, dim3(1), dim3(64), 0, st_, mdbs, nblk, tot);
        check_launch();
        double total = 0.0;
        FL_HIP(hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, st_));
        FL_HIP(hipStreamSynchronize(st_));
        std::vector<double> hthr((size_t)nt), hpot((size_t)nt);
        std::vector<int64_t> hcand((size_t)nt);
        // rows per distance block in one tile: LDS for the tile and the
        // candidates within 80 KB
        int I = 4;
        while (I > 1 && (size_t)64 * I * (d_ + 1) * sizeof(TX) + (size_t)nt * d_ * 8 > 81920)
            I /= 2;
        const size_t sm = (size_t)nt * d_ * 8 + (size_t)64 * I * (d_ + 1) * sizeof(TX);
        const dim3 tgrid((unsigned)ceil_div(nblk, 64), (unsigned)(KPP_BLK / I));
        auto trials = [&](auto ntc) {
            constexpr int NT = decltype(ntc)::value;
            FL_HIP(hipFuncSetAttribute((const void *)kpp_trials_il_kernel<TX, NT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
            hipLaunchKernelGGL((kpp_trials_il_kernel<TX, NT>), tgrid, dim3(64 * I), sm, st_, X_, n_,
                               d_, cand, nblk, I, md, ndarr);
        };
        auto launch_trials = [&]() {
            switch (nt) {
                case 1: return trials(std::integral_constant<int, 1>());
                case 2: return trials(std::integral_constant<int, 2>());
                case 3: return trials(std::integral_constant<int, 3>());
                case 4: return trials(std::integral_constant<int, 4>());
                case 5: return trials(std::integral_constant<int, 5>());
                case 6: return trials(std::integral_constant<int, 6>());
                case 7: return trials(std::integral_constant<int, 7>());
                case 8: return trials(std::integral_constant<int, 8>());
                case 9: return trials(std::integral_constant<int, 9>());
                case 10: return trials(std::integral_constant<int, 10>());
                case 11: return trials(std::integral_constant<int, 11>());
            }
            throw DeviceError("internal: unsupported k-means++ trial count");
        };
        for (int64_t c = 1; c < k_; ++c) {
            if (total <= 0.0) {
                const int64_t idx = rng.range(n_);
                hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_,
                                   X_, idx, d_, out + c * d_);
                check_launch();
                continue;
            }
            for (int64_t t = 0; t < nt; ++t) hthr[(size_t)t] = rng.f64() * total;
            FL_HIP(hipMemcpyAsync(thr, hthr.data(), (size_t)nt * 8, hipMemcpyHostToDevice, st_));
            hipLaunchKernelGGL(kpp_search_il_kernel, dim3((unsigned)nt), dim3(64), 0, st_, md, n_,
                               mdbs, nblk, thr, cand);
            launch_trials();
            hipLaunchKernelGGL(kpp_bsums_il_kernel, dim3(grid1(nt * nblk, 256)), dim3(256), 0,
                               st_, ndarr, nblk, (int)nt, ndbs);
            hipLaunchKernelGGL(kpp_totals_kernel, dim3((unsigned)nt), dim3(64), 0, st_, ndbs,
                               nblk, tot);
            check_launch();
            FL_HIP(hipMemcpyAsync(hcand.data(), cand, (size_t)nt * 8, hipMemcpyDeviceToHost, st_));
            FL_HIP(hipMemcpyAsync(hpot.data(), tot, (size_t)nt * 8, hipMemcpyDeviceToHost, st_));
            FL_HIP(hipStreamSynchronize(st_));
            // kmeans.rs:169-253: the first trial, then any strictly smaller potential
            int64_t bt = 0;
            for (int64_t t = 1; t < nt; ++t)
                if (hpot[(size_t)t] < hpot[(size_t)bt]) bt = t;
            FL_HIP(hipMemcpyAsync(md, ndarr + bt * stride, (size_t)stride * 8,
                                  hipMemcpyDeviceToDevice, st_));
            FL_HIP(hipMemcpyAsync(mdbs, ndbs + bt * nblk, (size_t)nblk * 8,
                                  hipMemcpyDeviceToDevice, st_));
            hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_, X_,
                               hcand[(size_t)bt], d_, out + c * d_);
            check_launch();
            total = hpot[(size_t)bt];
        }
        FL_HIP(hipStreamSynchronize(st_));
    }

    // k-means++ exactly as kmeans.rs:169-253 draws and sums it: the
    // generator's range / f64 draws in the reference's order, the total,
    // the threshold search and every trial's potential as plain left folds
    // (one device thread each), the strictly smallest potential kept.  The
    // caller's generator is left where the reference's would be.
    template <typename G>
    void kmeanspp_seq(G &rng, double *out) {
        const int64_t first = rng.range(n_);
        hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_, X_,
                           first, d_, out);
        check_launch();
        if (k_ == 1) {
            FL_HIP(hipStreamSynchronize(st_));
            return;
        }
        int64_t n_trials = 2 + (int64_t)std::floor(std::log((double)k_));
        if (n_trials < 1) n_trials = 1;
        if (n_trials > 64) n_trials = 64;
        const int nt = (int)n_trials;
        double *md = ws_.get<double>("kps_md", n_);
        double *pre = ws_.get<double>("kps_pre", n_);
        double *nd = ws_.get<double>("kps_nd", n_trials * n_);
        double *pre_t = ws_.get<double>("kps_pret", n_trials * n_);
        double *pot = ws_.get<double>("kps_pot", n_trials);
        // candidates of odd and even centres (a centre's are set to n - 1
        // while the previous centre's are still read)
        unsigned long long *cand = ws_.get<unsigned long long>("kps_cand", 2 * n_trials);
        KppSeqCtl *ctl = ws_.get<KppSeqCtl>("kps_ctl", 1);
        double *r = ws_.get<double>("kps_r", (k_ - 1) * n_trials);
        // Every centre's draws, made now on the assumption that the
        // reference's total stays > 0; the generator's state before each
        // centre is kept to replay the tail from if it does not.
        std::vector<G> snap;
        snap.reserve((size_t)(k_ - 1));
        std::vector<double> hr((size_t)((k_ - 1) * n_trials));
        for (int64_t c = 1; c < k_; ++c) {
            snap.push_back(rng);
            for (int t = 0; t < nt; ++t) hr[(size_t)((c - 1) * nt + t)] = rng.f64();
        }
        FL_HIP(hipMemcpyAsync(r, hr.data(), hr.size() * 8, hipMemcpyHostToDevice, st_));
        FL_HIP(hipMemsetAsync(ctl, 0, sizeof(KppSeqCtl), st_));
        hipLaunchKernelGGL(kpp_first_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_, X_, n_,
                           d_, out, md);
        left_fold(md, n_, n_, 1, nullptr, &ctl->total, pre, ws_, st_);
        const dim3 gn((unsigned)grid1(n_, 256), (unsigned)nt);
        hipLaunchKernelGGL(kpp_fill_u64_kernel, dim3(1), dim3(64), 0, st_, cand + nt, (int64_t)nt,
                           (unsigned long long)(n_ - 1));
        for (int64_t c = 1; c < k_; ++c) {
            unsigned long long *cc = cand + (c & 1) * nt, *cn = cand + ((c + 1) & 1) * nt;
            hipLaunchKernelGGL(kpp_seq_threshold_kernel, gn, dim3(256), 0, st_, ctl,
                               r + (c - 1) * nt, (int)c, pre, n_, cc);
            hipLaunchKernelGGL(kpp_seq_trials_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                               ctl, X_, n_, d_, nt, cc, md, nd);
            left_fold(nd, n_, n_, nt, &ctl->stop, pot, pre_t, ws_, st_);
            hipLaunchKernelGGL(kpp_seq_select_adopt_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0,
                               st_, ctl, pot, nt, cc, X_, d_, out + c * d_, n_, nd, pre_t, md, pre,
                               cn);
            check_launch();
        }
        int32_t stop = 0;
        FL_HIP(hipMemcpyAsync(&stop, &ctl->stop, 4, hipMemcpyDeviceToHost, st_));
        FL_HIP(hipStreamSynchronize(st_));
        if (stop > 0) {
            // total <= 0 at centre `stop`: the minimum distances never change
            // again, so every later centre is a uniformly drawn row (:201-205)
            rng = snap[(size_t)(stop - 1)];
            for (int64_t c = stop; c < k_; ++c) {
                const int64_t idx = rng.range(n_);
                hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_,
                                   X_, idx, d_, out + c * d_);
                check_launch();
            }
            FL_HIP(hipStreamSynchronize(st_));
        }
    }

    // k-means++ from a generator the caller keeps using: the library's
    // stream with its two-level sums, or the reference's with plain folds.
    void kmeanspp_from(Rng &rng, double *out) { kmeanspp(rng, out); }
    void kmeanspp_from(StdRng09 &rng, double *out) { kmeanspp_seq(rng, out); }

    // Same, drawing from a generator the caller continues to use afterwards
    // (MiniBatchKMeans samples its batches from the same stream).
    void kmeanspp(Rng &rng, double *out) {
        const int64_t first = rng.range(n_);
        hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_, X_,
                           first, d_, out);
        check_launch();
        if (k_ == 1) {
            FL_HIP(hipStreamSynchronize(st_));
            return;
        }
        double *md = ws_.get<double>("kpp_md", n_);
        double *nd = ws_.get<double>("kpp_nd", n_);
        double *bd = ws_.get<double>("kpp_bd", n_);
        const int64_t nblk = ceil_div(n_, KPP_BLK);
        double *bsum = ws_.get<double>("kpp_bsum", nblk);
        double *tot = ws_.get<double>("kpp_tot", 1);
        double *thr = ws_.get<double>("kpp_thr", 64);
        int64_t *cand = ws_.get<int64_t>("kpp_cand", 64);
        hipLaunchKernelGGL(kpp_first_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_, X_, n_,
                           d_, out, md);
        check_launch();
        int64_t n_trials = 2 + (int64_t)std::floor(std::log((double)k_));
        if (n_trials < 1) n_trials = 1;
        if (n_trials > 64) n_trials = 64;
        if (n_trials <= KPP_MAXT &&
            (size_t)n_trials * d_ * 8 + (size_t)64 * (d_ + 1) * sizeof(TX) <= 81920) {
            kmeanspp_batched(rng, out, n_trials);
            return;
        }
        std::vector<double> hthr((size_t)n_trials);
        std::vector<int64_t> hcand((size_t)n_trials);
        auto total_of = [&](const double *v) {
            hipLaunchKernelGGL(kpp_block_sums_kernel, dim3(grid1(nblk, 64)), dim3(64), 0, st_, v,
                               n_, bsum);
            hipLaunchKernelGGL(seq_sum_kernel, dim3(1), dim3(64), 0, st_, bsum, nblk, tot);
            check_launch();
            double h = 0.0;
            FL_HIP(hipMemcpyAsync(&h, tot, 8, hipMemcpyDeviceToHost, st_));
            FL_HIP(hipStreamSynchronize(st_));
            return h;
        };
        for (int64_t c = 1; c < k_; ++c) {
            const double total = total_of(md);   // leaves md's block sums in bsum
            if (total <= 0.0) {
                const int64_t idx = rng.range(n_);
                hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_,
                                   X_, idx, d_, out + c * d_);
                check_launch();
                continue;
            }
            for (int64_t t = 0; t < n_trials; ++t) hthr[(size_t)t] = rng.f64() * total;
            FL_HIP(hipMemcpyAsync(thr, hthr.data(), (size_t)n_trials * 8, hipMemcpyHostToDevice,
                                  st_));
            hipLaunchKernelGGL(kpp_search_kernel, dim3(1), dim3(64), 0, st_, md, n_, bsum, nblk,
                               thr, (int)n_trials, cand);
            check_launch();
            FL_HIP(hipMemcpyAsync(hcand.data(), cand, (size_t)n_trials * 8,
                                  hipMemcpyDeviceToHost, st_));
            FL_HIP(hipStreamSynchronize(st_));
            int64_t best = 0;
            double best_pot = 0.0;
            bool have = false;
            for (int64_t t = 0; t < n_trials; ++t) {
                hipLaunchKernelGGL(kpp_trial_kernel<TX>, dim3(grid1(n_, 256)), dim3(256), 0, st_,
                                   X_, n_, d_, hcand[(size_t)t], md, nd);
                check_launch();
                const double pot = total_of(nd);
                if (!have || pot < best_pot) {
                    have = true;
                    best_pot = pot;
                    best = hcand[(size_t)t];
                    std::swap(nd, bd);
                }
            }
            hipLaunchKernelGGL(copy_row_f64_kernel<TX>, dim3(grid1(d_, 64)), dim3(64), 0, st_, X_,
                               best, d_, out + c * d_);
            check_launch();
            std::swap(md, bd);
        }
        FL_HIP(hipStreamSynchronize(st_));
    }

    Workspace &ws() { return ws_; }

  private:
    const TX *X_;
    int64_t n_;
    int d_;
    int64_t k_;
    hipStream_t st_;
    bool direct_ = true;
    int passes_ = 0, nkc_ = 0;   // BF16 filter (passes_ > 0)
    bool fp8_ = false;           // FP8 filter first (BF16x3 for what it leaves)
    uint8_t *x8_ = nullptr;      // samples as E4M3, 128 bytes per row
    int32_t *sx_ = nullptr;      // their per-row scale exponents
    int d4_ = 0;                 // f64 MFMA filter (d4_ > 0)
    double *xn_ = nullptr, *xs_ = nullptr;
    unsigned short *xh_ = nullptr, *xl_ = nullptr;
    bool timing_ = true;
    std::unique_ptr<PhaseTimer> tm_;
    std::vector<std::pair<int, double *>> pend_;
    Workspace ws_;
};

// kmeans.rs:421-449, same order and conditions.
fl_status validate_fit(int64_t n, int64_t d, const fl_kmeans_params *p, int32_t dtype) {
    if (!p) return fail(FL_INVALID_PARAMETER, "InvalidParameter: params: must not be NULL");
    if (p->n_clusters == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_clusters: must be at least 1");
    if (n == 0)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(p->n_clusters) +
                        ", actual 0: KMeans requires at least n_clusters samples");
    if (n < p->n_clusters)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(p->n_clusters) +
                        ", actual " + std::to_string(n) +
                        ": KMeans requires at least n_clusters samples");
    if (p->n_init == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_init: must be at least 1");
    // Limits of this implementation (no reference counterpart).
    if (p->n_clusters < 0 || p->n_clusters > kMaxClusters)
        return fail(FL_INVALID_PARAMETER,
                    "InvalidParameter: n_clusters: the GPU path supports 1..2^31-65 clusters");
    if (p->n_init < 0 || p->max_iter < 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_init/max_iter: negative");
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    if (d <= 0 || d > (1 << 20))
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_features: must be in 1..2^20");
    if (n > 0x7fffffffLL)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_samples: at most 2^31-1 rows");
    if (p->rng != FL_RNG_LIBRARY && p->rng != FL_RNG_STDRNG)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: rng: unknown generator");
    if (p->precision != FL_PREC_FP64 && p->precision != FL_PREC_EXACT &&
        p->precision != FL_PREC_BF16 && p->precision != FL_PREC_BF16X3 &&
        p->precision != FL_PREC_FP8)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: precision: unknown mode");
    return FL_OK;
}

template <typename TX>
fl_status fit_device(const TX *X, int64_t n, int d, const fl_kmeans_params *p,
                     const double *d_init, double *d_centers, int64_t *d_labels,
                     double *inertia_out, int64_t *n_iter_out, fl_report *rep) {
    WallClock wc;
    StreamGuard sg;
    hipStream_t st = sg.s;
    const int64_t k = p->n_clusters;
    Lloyd<TX> L(X, n, d, k, p->precision, st);
    Workspace &ws = L.ws();
    // Iteration i reads centres C[i&1] and writes C[(i+1)&1], labels lab[i&1]
    // and scalars [inertia, shift] at sc + 2 (i&1).  Iteration i+1 is
    // enqueued before the host waits for iteration i's shift (through pinned
    // memory), so the device never idles on the convergence test; an
    // iteration enqueued after convergence touches none of the buffers the
    // result is read from, and its counts and phase times are discarded.
    double *C[2] = {ws.get<double>("cur", k * d), ws.get<double>("nxt", k * d)};
    int32_t *lab[2] = {ws.get<int32_t>("labels", n), ws.get<int32_t>("labels2", n)};
    double *best = ws.get<double>("best", n);
    double *sc = ws.get<double>("scalars", 4);
    FL_HIP(hipMemsetAsync(lab[0], 0, (size_t)n * 4, st));
    struct IterOut {
        double inertia, shift;
        int32_t counts[2];
    };
    PinnedHost<IterOut> hout(2);   // zeroed: counts[1] stays 0 without the FP8 filter
    EventPair ev;
    const uint64_t base = p->has_random_state ? p->random_state : 0;
    KmAcc acc;
    double best_inertia = 0.0;
    int64_t best_iter = 0;
    for (int64_t run = 0; run < p->n_init; ++run) {
        if (d_init)
            FL_HIP(hipMemcpyAsync(C[0], d_init + run * k * d, (size_t)(k * d) * 8,
                                  hipMemcpyDeviceToDevice, st));
        else if (p->rng == FL_RNG_STDRNG) {
            StdRng09 g(base + (uint64_t)run);   // kmeans.rs:460
            L.kmeanspp_seq(g, C[0]);
        } else
            L.kmeanspp(base + (uint64_t)run, C[0]);
        double inertia = 1.7976931348623157e308;
        int64_t n_iter = 0;
        for (int64_t it = 0;; ++it) {
            const bool enq = it < p->max_iter;
            const KmAcc before = acc;
            const size_t mk = L.mark();
            if (enq) {
                const int s = (int)(it & 1);
                L.assign(C[s], lab[s], best, acc);
                L.inertia(best, sc + 2 * s);
                L.update(lab[s], C[s], C[s ^ 1], sc + 2 * s + 1, acc);
                L.copy_counts(hout[s].counts);
                FL_HIP(hipMemcpyAsync(&hout[s].inertia, sc + 2 * s, 16, hipMemcpyDeviceToHost, st));
                FL_HIP(hipEventRecord(ev.e[s], st));
            }
            if (it == 0) {
                if (!enq) break;   // max_iter == 0
                continue;
            }
            // iteration it-1 is complete once its event is
            const int ps = (int)((it - 1) & 1);
            FL_HIP(hipEventSynchronize(ev.e[ps]));
            L.flush_upto(mk);
            acc.n_fb += hout[ps].counts[0];
            acc.n_second += hout[ps].counts[1];
            inertia = hout[ps].inertia;
            n_iter = it;
            if (hout[ps].shift < p->tol || !enq) {
                if (enq) {   // discard iteration it
                    L.drop_from(0);
                    acc.n_rows = before.n_rows;
                    acc.n_cand = before.n_cand;
                    acc.n_launch = before.n_launch;
                }
                break;
            }
        }
        if (run == 0 || inertia < best_inertia) {
            best_inertia = inertia;
            best_iter = n_iter;
            FL_HIP(hipMemcpyAsync(d_centers, C[n_iter & 1], (size_t)(k * d) * 8,
                                  hipMemcpyDeviceToDevice, st));
            const int32_t *fl = n_iter > 0 ? lab[(n_iter - 1) & 1] : lab[0];
            hipLaunchKernelGGL(labels_to_i64_kernel, dim3(grid1(n, 256)), dim3(256), 0, st, fl,
                               n, d_labels);
            check_launch();
        }
    }
    FL_HIP(hipStreamSynchronize(st));
    *inertia_out = best_inertia;
    *n_iter_out = best_iter;
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->ms_total = wc.ms();
        rep->ms_distance = acc.ms_dist;
        rep->ms_finalize = acc.ms_fin;
        rep->ms_fallback = acc.ms_fb;
        rep->ms_update = acc.ms_upd;
        rep->n_rows = acc.n_rows;
        rep->n_candidates = acc.n_cand;
        rep->n_fallback = acc.n_fb;
        rep->n_iter = best_iter;
        rep->n_dist_launches = acc.n_launch;
        rep->n_second = acc.n_second;
    }
    return FL_OK;
}

fl_status fit_dispatch(const void *dX, int32_t dtype, int64_t n, int64_t d,
                       const fl_kmeans_params *p, const double *d_init, double *d_centers,
                       int64_t *d_labels, double *inertia, int64_t *n_iter, fl_report *rep) {
    if (dtype == FL_F32)
        return fit_device<float>((const float *)dX, n, (int)d, p, d_init, d_centers, d_labels,
                                 inertia, n_iter, rep);
    return fit_device<double>((const double *)dX, n, (int)d, p, d_init, d_centers, d_labels,
                              inertia, n_iter, rep);
}

// ------------------------------------------------------ allocation-free --
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (on its first call there) and kept.
void *scratch_address() {
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        throw DeviceError("device index beyond the scratch address table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Upper bound of the scratch bytes Lloyd::assign takes for a chunk of m rows
// against k centres of d features (every slot any filter form can take, each
// rounded up to 256 bytes), plus the chunk's int32 labels and best folds.
size_t noalloc_bytes(int64_t m, int64_t k, int d) {
    const int64_t dpad = (int64_t)bf16_nkc(d) * 32;   // BF16 forms (0: not used)
    const int64_t dp = (int64_t)f64_d4(d) * 4;        // f64 filter padding (0: not used)
    const int64_t rows = m * (8 + 8 + 4 * dpad + 32 + 16 + 4 + 8 + 4 + 8 + 4 + 4 + 8);
    const int64_t cent = k * (8 + 8 + 4 * dpad + 8 + 8 * dp + 8 * (int64_t)d);
    return (size_t)(rows + cent + 32 * 256);
}

fl_status check_predict(int64_t k, int64_t d_centers, int64_t n, int64_t d, int32_t dtype) {
    if (d != d_centers)
        return fail(FL_SHAPE_MISMATCH,
                    "ShapeMismatch: expected [" + std::to_string(d_centers) + "], actual [" +
                        std::to_string(d) +
                        "]: number of features must match fitted KMeans model");
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    if (k <= 0 || k > kMaxClusters)
        return fail(FL_INVALID_PARAMETER,
                    "InvalidParameter: n_clusters: the GPU path supports 1..2^31-65 clusters");
    if (n < 0 || n > 0x7fffffffLL)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_samples: out of range");
    return FL_OK;
}

// ------------------------------------------------------------- mini-batch --
// mini_batch_kmeans.rs:451-472 (sample_batch_indices) with the library's
// generator: a full Fisher-Yates shuffle when batch >= n, else the first
// `batch` steps of one over 0..n (the swapped entries kept in a map, so a
// step costs O(1) instead of rebuilding the pool).
template <typename G>
std::vector<int32_t> sample_batch(G &rng, int64_t n, int64_t batch) {
    std::vector<int32_t> out;
    if (batch >= n) {
        out.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) out[(size_t)i] = (int32_t)i;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t j = rng.range_from(i, n);
            std::swap(out[(size_t)i], out[(size_t)j]);
        }
        return out;
    }
    std::unordered_map<int64_t, int64_t> pool;
    pool.reserve((size_t)batch * 2);
    auto at = [&](int64_t i) {
        auto it = pool.find(i);
        return it == pool.end() ? i : it->second;
    };
    out.resize((size_t)batch);
    for (int64_t i = 0; i < batch; ++i) {
        const int64_t j = rng.range_from(i, n);
        const int64_t vi = at(i), vj = at(j);
        pool[i] = vj;
        pool[j] = vi;
        out[(size_t)i] = (int32_t)vj;
    }
    return out;
}

// Fit<Array2<F>, ()> for MiniBatchKMeans (mini_batch_kmeans.rs:478-592):
// per run (seed base + run * 1000003) initial centres (given, or the
// library's k-means++ from that run's generator), counts = 1, max_iter
// iterations of sample -> assign batch -> online update -> n_iter += 1 ->
// stop if shift < tol; then every row is assigned (the candidate filter of
// `precision` with the exact re-score) and the inertia is the left fold of
// the best folds in row order.  The run with strictly smaller inertia wins.
template <typename TX>
fl_status mb_fit_device(const TX *X, int64_t n, int d, const fl_minibatch_params *p,
                        const double *d_init, double *d_centers, int64_t *d_labels,
                        double *inertia_out, int64_t *n_iter_out, fl_report *rep) {
    WallClock wc;
    StreamGuard sg;
    hipStream_t st = sg.s;
    const int64_t k = p->n_clusters;
    Lloyd<TX> L(X, n, d, k, p->precision, st);
    Workspace &ws = L.ws();
    double *cur = ws.get<double>("mb_cur", k * d);
    double *old = ws.get<double>("mb_old", k * d);
    int64_t *counts = ws.get<int64_t>("mb_counts", k);
    int32_t *labels = ws.get<int32_t>("labels", n);
    double *best = ws.get<double>("best", n);
    double *dsc = ws.get<double>("scalars", 2);   // [shift, inertia]
    const int64_t bmax = p->batch_size < n ? p->batch_size : n;
    int32_t *bidx = ws.get<int32_t>("mb_bidx", bmax);
    int32_t *blab = ws.get<int32_t>("mb_blab", bmax);
    const uint64_t base = p->has_random_state ? p->random_state : 0;
    KmAcc acc;
    double best_inertia = 0.0;
    int64_t best_iter = 0;
    PhaseTimer tm(st);
    double ms_batch = 0.0;
    // one run from its generator (the library's or the reference's stream)
    auto one_run = [&](int64_t run, auto &rng) {
        if (d_init)
            FL_HIP(hipMemcpyAsync(cur, d_init + run * k * d, (size_t)(k * d) * 8,
                                  hipMemcpyDeviceToDevice, st));
        else
            L.kmeanspp_from(rng, cur);
        hipLaunchKernelGGL(fill_i64_kernel, dim3(grid1(k, 256)), dim3(256), 0, st, counts, k,
                           (int64_t)1);
        check_launch();
        int64_t n_iter = 0;
        for (int64_t it = 0; it < p->max_iter; ++it) {
            const std::vector<int32_t> bi = sample_batch(rng, n, p->batch_size);
            const int64_t b = (int64_t)bi.size();
            const int t0 = tm.begin();
            FL_HIP(hipMemcpyAsync(bidx, bi.data(), (size_t)b * 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(mb_assign_batch_kernel<TX>, dim3(grid1(b, 4)), dim3(256), 0, st, X,
                               cur, d, k, bidx, b, blab);
            FL_HIP(hipMemcpyAsync(old, cur, (size_t)(k * d) * 8, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(mb_update_kernel<TX>, dim3((unsigned)k), dim3(64), 0, st, X, d,
                               bidx, blab, b, cur, counts);
            unsigned long long *mx = ws.get<unsigned long long>("shift_max", 1);
            FL_HIP(hipMemsetAsync(mx, 0, 8, st));
            hipLaunchKernelGGL(km_shift_kernel, dim3(grid1(k, 64)), dim3(64), 0, st, cur, old, k,
                               d, mx);
            hipLaunchKernelGGL(km_shift_finish, dim3(1), dim3(64), 0, st, mx, dsc);
            check_launch();
            tm.end(t0);
            double shift = 0.0;
            FL_HIP(hipMemcpyAsync(&shift, dsc, 8, hipMemcpyDeviceToHost, st));
            FL_HIP(hipStreamSynchronize(st));
            ms_batch += tm.ms(t0);
            n_iter = it + 1;
            if (shift < p->tol) break;
        }
        L.assign(cur, labels, best, acc);
        left_fold(best, n, n, 1, nullptr, dsc + 1, nullptr, ws, st);
        double inertia = 0.0;
        FL_HIP(hipMemcpyAsync(&inertia, dsc + 1, 8, hipMemcpyDeviceToHost, st));
        L.flush(&acc.n_fb);
        if (run == 0 || inertia < best_inertia) {
            best_inertia = inertia;
            best_iter = n_iter;
            FL_HIP(hipMemcpyAsync(d_centers, cur, (size_t)(k * d) * 8, hipMemcpyDeviceToDevice,
                                  st));
            hipLaunchKernelGGL(labels_to_i64_kernel, dim3(grid1(n, 256)), dim3(256), 0, st, labels,
                               n, d_labels);
            check_launch();
        }
    };
    for (int64_t run = 0; run < p->n_init; ++run) {
        const uint64_t seed = base + (uint64_t)run * 1000003ULL;   // mini_batch_kmeans.rs:532
        if (p->rng == FL_RNG_STDRNG) {
            StdRng09 g(seed);
            one_run(run, g);
        } else {
            Rng g(seed);
            one_run(run, g);
        }
    }
    FL_HIP(hipStreamSynchronize(st));
    *inertia_out = best_inertia;
    *n_iter_out = best_iter;
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->ms_total = wc.ms();
        rep->ms_distance = acc.ms_dist;
        rep->ms_finalize = acc.ms_fin;
        rep->ms_fallback = acc.ms_fb;
        rep->ms_update = ms_batch;
        rep->n_rows = acc.n_rows;
        rep->n_candidates = acc.n_cand;
        rep->n_fallback = acc.n_fb;
        rep->n_iter = best_iter;
        rep->n_dist_launches = acc.n_launch;
    }
    return FL_OK;
}

// mini_batch_kmeans.rs:495-526: n_clusters, batch_size, n_init, then the
// sample count, in that order.
fl_status validate_mb(int64_t n, int64_t d, const fl_minibatch_params *p, int32_t dtype) {
    if (!p) return fail(FL_INVALID_PARAMETER, "InvalidParameter: params: must not be NULL");
    if (p->n_clusters == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_clusters: must be at least 1");
    if (p->batch_size == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: batch_size: must be at least 1");
    if (p->n_init == 0)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_init: must be at least 1");
    if (n == 0)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(p->n_clusters) +
                        ", actual 0: MiniBatchKMeans requires at least n_clusters samples");
    if (n < p->n_clusters)
        return fail(FL_INSUFFICIENT_SAMPLES,
                    "InsufficientSamples: required " + std::to_string(p->n_clusters) +
                        ", actual " + std::to_string(n) +
                        ": MiniBatchKMeans requires at least n_clusters samples");
    // Limits of this implementation (no reference counterpart).
    if (p->n_clusters < 0 || p->n_clusters > kMaxClusters)
        return fail(FL_INVALID_PARAMETER,
                    "InvalidParameter: n_clusters: the GPU path supports 1..2^31-65 clusters");
    if (p->batch_size < 0 || p->n_init < 0 || p->max_iter < 0)
        return fail(FL_INVALID_PARAMETER,
                    "InvalidParameter: batch_size/n_init/max_iter: negative");
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    if (d <= 0 || d > (1 << 20))
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_features: must be in 1..2^20");
    if (n > 0x7fffffffLL)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_samples: at most 2^31-1 rows");
    if (p->rng != FL_RNG_LIBRARY && p->rng != FL_RNG_STDRNG)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: rng: unknown generator");
    if (p->precision != FL_PREC_FP64 && p->precision != FL_PREC_EXACT &&
        p->precision != FL_PREC_BF16 && p->precision != FL_PREC_BF16X3 &&
        p->precision != FL_PREC_FP8)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: precision: unknown mode");
    return FL_OK;
}

fl_status mb_dispatch(const void *dX, int32_t dtype, int64_t n, int64_t d,
                      const fl_minibatch_params *p, const double *d_init, double *d_centers,
                      int64_t *d_labels, double *inertia, int64_t *n_iter, fl_report *rep) {
    if (dtype == FL_F32)
        return mb_fit_device<float>((const float *)dX, n, (int)d, p, d_init, d_centers, d_labels,
                                    inertia, n_iter, rep);
    return mb_fit_device<double>((const double *)dX, n, (int)d, p, d_init, d_centers, d_labels,
                                 inertia, n_iter, rep);
}

}  // namespace

namespace fl {
// kmeans.rs:421-449 validation, then this build's limits (shared with the
// batched entry point).
fl_status validate_kmeans_fit(int64_t n, int64_t d, const fl_kmeans_params *p, int32_t dtype) {
    return validate_fit(n, d, p, dtype);
}

// Exact nearest centre (lowest index on ties) and its fold for n device rows
// of dtype: the Lloyd assignment step on its own (BF16x3 filter with the
// exact decision, fallbacks as in a fit).  Used by the cell router.
void nearest_centers(const void *X, int32_t dtype, int64_t n, int d, const double *C, int64_t k,
                     int32_t *labels, double *best, hipStream_t st) {
    KmAcc acc;
    if (dtype == FL_F32) {
        Lloyd<float> L((const float *)X, n, d, k, FL_PREC_BF16X3, st);
        L.assign(C, labels, best, acc);
        L.flush();
    } else {
        Lloyd<double> L((const double *)X, n, d, k, FL_PREC_BF16X3, st);
        L.assign(C, labels, best, acc);
        L.flush();
    }
}
}  // namespace fl

extern "C" {

void fl_minibatch_params_default(fl_minibatch_params *p, int64_t n_clusters) {
    std::memset(p, 0, sizeof(*p));
    p->n_clusters = n_clusters;
    p->batch_size = 1024;
    p->max_iter = 100;
    p->tol = 0.0;
    p->n_init = 3;
    p->has_random_state = 0;
    p->precision = FL_PREC_FP64;
}

fl_status fl_minibatch_kmeans_fit_device(const void *dX, int32_t dtype, int64_t n, int64_t d,
                                         const fl_minibatch_params *params, const double *d_init,
                                         double *d_centers_out, int64_t *d_labels_out,
                                         double *inertia_out, int64_t *n_iter_out,
                                         fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = validate_mb(n, d, params, dtype)) return s;
        return mb_dispatch(dX, dtype, n, d, params, d_init, d_centers_out, d_labels_out,
                           inertia_out, n_iter_out, report);
    });
}

fl_status fl_minibatch_kmeans_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                                  const fl_minibatch_params *params, const double *init_centers,
                                  double *centers_out, int64_t *labels_out, double *inertia_out,
                                  int64_t *n_iter_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = validate_mb(n, d, params, dtype)) return s;
        const int64_t k = params->n_clusters;
        DevBuf bx, bi, bc, bl;
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        double *dI = nullptr;
        if (init_centers) {
            const size_t ib = (size_t)(params->n_init * k * d) * 8;
            dI = (double *)bi.get(ib);
            FL_HIP(hipMemcpy(dI, init_centers, ib, hipMemcpyHostToDevice));
        }
        double *dC = (double *)bc.get((size_t)(k * d) * 8);
        int64_t *dL = (int64_t *)bl.get((size_t)n * 8);
        fl_status s = mb_dispatch(dX, dtype, n, d, params, dI, dC, dL, inertia_out, n_iter_out,
                                  report);
        if (s) return s;
        FL_HIP(hipMemcpy(centers_out, dC, (size_t)(k * d) * 8, hipMemcpyDeviceToHost));
        FL_HIP(hipMemcpy(labels_out, dL, (size_t)n * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}


void fl_kmeans_params_default(fl_kmeans_params *p, int64_t n_clusters) {
    std::memset(p, 0, sizeof(*p));
    p->n_clusters = n_clusters;
    p->max_iter = 300;
    p->tol = 1e-4;
    p->n_init = 10;
    p->has_random_state = 0;
    p->precision = FL_PREC_FP64;
}

fl_status fl_kmeans_fit_device(const void *dX, int32_t dtype, int64_t n, int64_t d,
                               const fl_kmeans_params *params, const double *d_init,
                               double *d_centers_out, int64_t *d_labels_out, double *inertia_out,
                               int64_t *n_iter_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = validate_fit(n, d, params, dtype)) return s;
        return fit_dispatch(dX, dtype, n, d, params, d_init, d_centers_out, d_labels_out,
                            inertia_out, n_iter_out, report);
    });
}

fl_status fl_kmeans_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                        const fl_kmeans_params *params, const double *init_centers,
                        double *centers_out, int64_t *labels_out, double *inertia_out,
                        int64_t *n_iter_out, fl_report *report) {
    return guarded([&]() -> fl_status {
        if (fl_status s = validate_fit(n, d, params, dtype)) return s;
        const int64_t k = params->n_clusters;
        DevBuf bx, bi, bc, bl;
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        double *dI = nullptr;
        if (init_centers) {
            const size_t ib = (size_t)(params->n_init * k * d) * 8;
            dI = (double *)bi.get(ib);
            FL_HIP(hipMemcpy(dI, init_centers, ib, hipMemcpyHostToDevice));
        }
        double *dC = (double *)bc.get((size_t)(k * d) * 8);
        int64_t *dL = (int64_t *)bl.get((size_t)n * 8);
        fl_status s = fit_dispatch(dX, dtype, n, d, params, dI, dC, dL, inertia_out, n_iter_out,
                                   report);
        if (s) return s;
        FL_HIP(hipMemcpy(centers_out, dC, (size_t)(k * d) * 8, hipMemcpyDeviceToHost));
        FL_HIP(hipMemcpy(labels_out, dL, (size_t)n * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

fl_status fl_kmeans_predict(const double *centers, int64_t k, int64_t d_centers, const void *X,
                            int32_t dtype, int64_t n, int64_t d, int64_t *labels_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_predict(k, d_centers, n, d, dtype)) return s;
        if (n == 0) return FL_OK;
        StreamGuard sg;
        DevBuf bx, bc, bl, bb, b64;
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        double *dC = (double *)bc.get((size_t)(k * d) * 8);
        FL_HIP(hipMemcpy(dC, centers, (size_t)(k * d) * 8, hipMemcpyHostToDevice));
        int32_t *dl = (int32_t *)bl.get((size_t)n * 4);
        double *db = (double *)bb.get((size_t)n * 8);
        int64_t *d64 = (int64_t *)b64.get((size_t)n * 8);
        KmAcc acc;
        if (dtype == FL_F32) {
            Lloyd<float> L((const float *)dX, n, (int)d, k, FL_PREC_FP64, sg.s);
            L.assign(dC, dl, db, acc);
            L.flush();
        } else {
            Lloyd<double> L((const double *)dX, n, (int)d, k, FL_PREC_FP64, sg.s);
            L.assign(dC, dl, db, acc);
            L.flush();
        }
        hipLaunchKernelGGL(labels_to_i64_kernel, dim3(grid1(n, 256)), dim3(256), 0, sg.s, dl, n,
                           d64);
        check_launch();
        FL_HIP(hipStreamSynchronize(sg.s));
        FL_HIP(hipMemcpy(labels_out, d64, (size_t)n * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

fl_status fl_kmeans_predict_device_noalloc(const double *dC, int64_t k, int64_t d_centers,
                                          const void *dX, int32_t dtype, int64_t n, int64_t d,
                                          int32_t precision, int64_t *d_labels_out,
                                          void *stream) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_predict(k, d_centers, n, d, dtype)) return s;
        if (precision != FL_PREC_FP64 && precision != FL_PREC_EXACT &&
            precision != FL_PREC_BF16 && precision != FL_PREC_BF16X3 && precision != FL_PREC_FP8)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: precision: unknown mode");
        if (n == 0) return FL_OK;
        if (!dC || !dX || !d_labels_out)
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: centers, X and labels_out must not be NULL");
        // Rows go through in chunks that fit the fixed scratch; the centres'
        // share of it does not shrink with the chunk.
        const int64_t m_min = std::min<int64_t>(n, 256);
        if (noalloc_bytes(m_min, k, (int)d) > kNoAllocScratchBytes)
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: n_clusters x n_features: k=" + std::to_string(k) +
                            ", d=" + std::to_string(d) + " needs " +
                            std::to_string(noalloc_bytes(m_min, k, (int)d) >> 20) +
                            " MiB of scratch; the allocation-free predict has " +
                            std::to_string(kNoAllocScratchBytes >> 20) +
                            " MiB (about 16*k*d bytes plus ~400 bytes per 256-row chunk row)");
        int64_t m = m_min;
        while (m < n && noalloc_bytes(std::min<int64_t>(n, 2 * m), k, (int)d) <= kNoAllocScratchBytes)
            m = std::min<int64_t>(n, 2 * m);
        void *base = scratch_address();
        const hipStream_t st = (hipStream_t)stream;
        auto run = [&](auto tag) {
            using TX = decltype(tag);
            const TX *X = (const TX *)dX;
            for (int64_t r0 = 0; r0 < n; r0 += m) {
                const int64_t mm = std::min<int64_t>(m, n - r0);
                Lloyd<TX> L(X + r0 * d, mm, (int)d, k, precision, st, base, kNoAllocScratchBytes,
                            dC);
                int32_t *lab = L.ws().template get<int32_t>("labels", mm);
                double *best = L.ws().template get<double>("best", mm);
                KmAcc acc;
                L.assign(dC, lab, best, acc);
                hipLaunchKernelGGL(labels_to_i64_kernel, dim3(grid1(mm, 256)), dim3(256), 0, st,
                                   lab, mm, d_labels_out + r0);
                check_launch();
            }
        };
        if (dtype == FL_F32)
            run(float());
        else
            run(double());
        FL_HIP(hipStreamSynchronize(st));
        return FL_OK;
    });
}

fl_status fl_kmeans_transform(const double *centers, int64_t k, int64_t d_centers, const void *X,
                              int32_t dtype, int64_t n, int64_t d, double *dist_out) {
    return guarded([&]() -> fl_status {
        if (fl_status s = check_predict(k, d_centers, n, d, dtype)) return s;
        if (n == 0) return FL_OK;
        DevBuf bx, bc, bo;
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        double *dC = (double *)bc.get((size_t)(k * d) * 8);
        FL_HIP(hipMemcpy(dC, centers, (size_t)(k * d) * 8, hipMemcpyHostToDevice));
        double *dO = (double *)bo.get((size_t)(n * k) * 8);
        if (dtype == FL_F32)
            hipLaunchKernelGGL(km_transform_kernel<float>, dim3(grid1(n * k, 256)), dim3(256), 0,
                               0, (const float *)dX, dC, n, (int)d, k, dO);
        else
            hipLaunchKernelGGL(km_transform_kernel<double>, dim3(grid1(n * k, 256)), dim3(256), 0,
                               0, (const double *)dX, dC, n, (int)d, k, dO);
        check_launch();
        FL_HIP(hipMemcpy(dist_out, dO, (size_t)(n * k) * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

fl_status fl_kmeans_plusplus(const void *X, int32_t dtype, int64_t n, int64_t d, int64_t k,
                             uint64_t seed, double *centers_out) {
    return guarded([&]() -> fl_status {
        fl_kmeans_params p;
        fl_kmeans_params_default(&p, k);
        if (fl_status s = validate_fit(n, d, &p, dtype)) return s;
        StreamGuard sg;
        DevBuf bx, bc;
        const size_t xb = (size_t)(n * d) * dtype_size(dtype);
        void *dX = bx.get(xb);
        FL_HIP(hipMemcpy(dX, X, xb, hipMemcpyHostToDevice));
        double *dC = (double *)bc.get((size_t)(k * d) * 8);
        if (dtype == FL_F32) {
            Lloyd<float> L((const float *)dX, n, (int)d, k, FL_PREC_EXACT, sg.s);
            L.kmeanspp(seed, dC);
        } else {
            Lloyd<double> L((const double *)dX, n, (int)d, k, FL_PREC_EXACT, sg.s);
            L.kmeanspp(seed, dC);
        }
        FL_HIP(hipMemcpy(centers_out, dC, (size_t)(k * d) * 8, hipMemcpyDeviceToHost));
        return FL_OK;
    });
}

fl_status fl_left_fold_device(const double *d_v, int64_t n, int64_t count, int64_t ld,
                              double *d_sums, double *d_prefix) {
    return guarded([&]() -> fl_status {
        if (n < 0 || count < 0 || (count > 1 && ld < n) || count > 65535)
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: left fold needs n >= 0, ld >= n, count <= 65535");
        if (count == 0) return FL_OK;
        if (!d_v || !d_sums)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: left fold: null pointer");
        StreamGuard sg;
        Workspace ws;
        left_fold(d_v, n, count > 1 ? ld : n, count, nullptr, d_sums, d_prefix, ws, sg.s);
        FL_HIP(hipStreamSynchronize(sg.s));
        return FL_OK;
    });
}

}  // extern "C"
