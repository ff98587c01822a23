// router.hip -- learned cell router for exact kNN on gfx950.
//
// No reference counterpart: the reference answers kNN by brute force
// (kdtree.rs:364-391) or trees (knn.rs:86-150).  This index keeps that
// answer exactly: cells from a coarse KMeans, a BF16 perceptron that
// chooses which cells a query scans (layers, loss and Adam as in
// ferrolearn-neural/src/mlp.rs), an exact scan of those cells, a
// triangle-inequality certificate, and the brute force for every query the
// certificate does not cover.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "host.hpp"
#include "kmeans_kernels.hpp"
#include "mlp_bf16.hpp"
#include "router_kernels.hpp"

using namespace fl;

namespace fl {
void knn_describe(const fl_knn *h, const void **X, int32_t *dtype, int64_t *n, int64_t *d);
int64_t knn_n_neighbors(const fl_knn *h);
void nearest_centers(const void *X, int32_t dtype, int64_t n, int d, const double *C, int64_t k,
                     int32_t *labels, double *best, hipStream_t st);
}

namespace {

constexpr int64_t kMaxCells = 8192;
constexpr int64_t kMaxHidden = 1024;
constexpr int kMaxProbe = 31;
constexpr int kMaxRoutedK = 32;
constexpr int64_t kInferRows = 32768;   // network inference chunk
constexpr int64_t kMaxTargets = 32;     // certification targets per training query
constexpr int kRouteAll = 3;            // internal route: every cell (exact by construction)
constexpr int64_t kFullScanMax = 2048;  // queries a last all-cells scan takes
constexpr int64_t kPairListBytes = (int64_t)1 << 31;   // per-pair top-k lists of one pass

// Router network: d -> H -> H -> C with padded shapes (dp: 32, cp: 64).
struct Mlp {
    int d = 0, dp = 0, H = 0, C = 0, cp = 0;
    int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, np = 0;
    DevBuf P, G, Mo, V, W1b, W2b, W3b;
    int64_t t = 0;   // Adam steps taken

    void shape(int d_, int H_, int C_) {
        d = d_;
        dp = (int)ceil_div(d_, 32) * 32;
        H = H_;
        C = C_;
        cp = (int)ceil_div(C_, 64) * 64;
        oW1 = 0;
        ob1 = oW1 + (int64_t)dp * H;
        oW2 = ob1 + H;
        ob2 = oW2 + (int64_t)H * H;
        oW3 = ob2 + H;
        ob3 = oW3 + (int64_t)H * cp;
        np = ob3 + cp;
        P.get((size_t)np * 4);
        G.get((size_t)np * 4);
        Mo.get((size_t)np * 4);
        V.get((size_t)np * 4);
        W1b.get((size_t)dp * H * 2);
        W2b.get((size_t)H * H * 2);
        W3b.get((size_t)H * cp * 2);
    }
    float *p() const { return (float *)P.ptr(); }
    float *g() const { return (float *)G.ptr(); }
    uint16_t *w1b() const { return (uint16_t *)W1b.ptr(); }
    uint16_t *w2b() const { return (uint16_t *)W2b.ptr(); }
    uint16_t *w3b() const { return (uint16_t *)W3b.ptr(); }
    int64_t n_plain() const {
        return (int64_t)d * H + H + (int64_t)H * H + H + (int64_t)H * C + C;
    }
    // plain (mlp.rs order, unpadded) <-> padded layout, on the host
    void pad(const float *plain, std::vector<float> &out) const {
        out.assign((size_t)np, 0.f);
        const float *s = plain;
        for (int i = 0; i < d; ++i)
            for (int o = 0; o < H; ++o) out[(size_t)(oW1 + (int64_t)i * H + o)] = *s++;
        for (int o = 0; o < H; ++o) out[(size_t)(ob1 + o)] = *s++;
        for (int i = 0; i < H; ++i)
            for (int o = 0; o < H; ++o) out[(size_t)(oW2 + (int64_t)i * H + o)] = *s++;
        for (int o = 0; o < H; ++o) out[(size_t)(ob2 + o)] = *s++;
        for (int i = 0; i < H; ++i)
            for (int o = 0; o < C; ++o) out[(size_t)(oW3 + (int64_t)i * cp + o)] = *s++;
        for (int o = 0; o < C; ++o) out[(size_t)(ob3 + o)] = *s++;
    }
    void unpad(const std::vector<float> &in, float *plain) const {
        float *s = plain;
        for (int i = 0; i < d; ++i)
            for (int o = 0; o < H; ++o) *s++ = in[(size_t)(oW1 + (int64_t)i * H + o)];
        for (int o = 0; o < H; ++o) *s++ = in[(size_t)(ob1 + o)];
        for (int i = 0; i < H; ++i)
            for (int o = 0; o < H; ++o) *s++ = in[(size_t)(oW2 + (int64_t)i * H + o)];
        for (int o = 0; o < H; ++o) *s++ = in[(size_t)(ob2 + o)];
        for (int i = 0; i < H; ++i)
            for (int o = 0; o < C; ++o) *s++ = in[(size_t)(oW3 + (int64_t)i * cp + o)];
        for (int o = 0; o < C; ++o) *s++ = in[(size_t)(ob3 + o)];
    }
    void upload(const std::vector<float> &padded, hipStream_t st) {
        FL_HIP(hipMemcpyAsync(p(), padded.data(), (size_t)np * 4, hipMemcpyHostToDevice, st));
        FL_HIP(hipMemsetAsync(Mo.ptr(), 0, (size_t)np * 4, st));
        FL_HIP(hipMemsetAsync(V.ptr(), 0, (size_t)np * 4, st));
        t = 0;
        refresh_bf16(st);
    }
    void refresh_bf16(hipStream_t st) {
        hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(grid1((int64_t)dp * H, 256)), dim3(256), 0, st,
                           p() + oW1, (int64_t)dp * H, w1b());
        hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(grid1((int64_t)H * H, 256)), dim3(256), 0, st,
                           p() + oW2, (int64_t)H * H, w2b());
        hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(grid1((int64_t)H * cp, 256)), dim3(256), 0, st,
                           p() + oW3, (int64_t)H * cp, w3b());
        check_launch();
    }
    // mlp.rs:181-192 (Xavier uniform, limit sqrt(6 / (fan_in + fan_out)))
    // with the library's generator; biases zero (mlp.rs:1200-1203).
    void xavier(uint64_t seed, hipStream_t st) {
        Rng rng(seed);
        std::vector<float> plain((size_t)n_plain(), 0.f);
        float *s = plain.data();
        auto fill = [&](int fi, int fo) {
            const double lim = std::sqrt(6.0 / (double)(fi + fo));
            for (int64_t i = 0; i < (int64_t)fi * fo; ++i) *s++ = (float)(-lim + 2.0 * lim * rng.f64());
        };
        fill(d, H);
        s += H;
        fill(H, H);
        s += H;
        fill(H, C);
        std::vector<float> padded;
        pad(plain.data(), padded);
        upload(padded, st);
    }
};

MlpGemmArgs gemm_args() {
    MlpGemmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.scale = 1.f;
    return a;
}

template <bool AK, bool BK>
void gemm(const MlpGemmArgs &a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0) return;
    hipLaunchKernelGGL((mlp_gemm_kernel<AK, BK>), dim3(grid1(a.N, MG_BN), grid1(a.M, MG_BM)),
                       dim3(256), 0, st, a);
    check_launch();
}

// Buffers of one batch of nbp (padded) rows.
struct Batch {
    uint16_t *xb, *a1, *a2, *g3, *d2, *d1, *ones;
    float *z;
};

Batch batch_buffers(Workspace &ws, const Mlp &net, int64_t nbp, bool backward, hipStream_t st) {
    Batch b;
    b.ones = nullptr;
    if (backward) {
        // bf16 1.0 per row: bias gradients are G^T 1 on the GEMM kernel
        b.ones = ws.get<uint16_t>("mlp_ones", nbp);
        FL_HIP(hipMemsetD16Async((hipDeviceptr_t)b.ones, 0x3F80, (size_t)nbp, st));
    }
    b.xb = ws.get<uint16_t>("mlp_xb", nbp * net.dp);
    b.a1 = ws.get<uint16_t>("mlp_a1", nbp * net.H);
    b.a2 = ws.get<uint16_t>("mlp_a2", nbp * net.H);
    b.z = ws.get<float>("mlp_z", nbp * net.cp);
    b.g3 = backward ? ws.get<uint16_t>("mlp_g3", nbp * net.cp) : nullptr;
    b.d2 = backward ? ws.get<uint16_t>("mlp_d2", nbp * net.H) : nullptr;
    b.d1 = backward ? ws.get<uint16_t>("mlp_d1", nbp * net.H) : nullptr;
    return b;
}

// z = relu(relu(x W1 + b1) W2 + b2) W3 + b3 (mlp.rs:217-261).
void mlp_forward(const Mlp &net, const Batch &b, int nbp, hipStream_t st) {
    MlpGemmArgs a = gemm_args();
    a.A = b.xb, a.lda = net.dp, a.B = net.w1b(), a.ldb = net.H;
    a.M = nbp, a.N = net.H, a.K = net.dp;
    a.bias = net.p() + net.ob1, a.relu = 1, a.Cb = b.a1, a.ldcb = net.H;
    gemm<true, false>(a, st);
    a = gemm_args();
    a.A = b.a1, a.lda = net.H, a.B = net.w2b(), a.ldb = net.H;
    a.M = nbp, a.N = net.H, a.K = net.H;
    a.bias = net.p() + net.ob2, a.relu = 1, a.Cb = b.a2, a.ldcb = net.H;
    gemm<true, false>(a, st);
    a = gemm_args();
    a.A = b.a2, a.lda = net.H, a.B = net.w3b(), a.ldb = net.cp;
    a.M = nbp, a.N = net.cp, a.K = net.H;
    a.bias = net.p() + net.ob3, a.Cf = b.z, a.ldcf = net.cp;
    gemm<true, false>(a, st);
}

// grad_b[n] = scale * sum over rows of G[row][n] (rows past the batch are
// zero), as the GEMM G^T 1.
void bias_grad(const uint16_t *G, int ldg, int N, int nbp, float scale, float *out,
               const uint16_t *ones, hipStream_t st) {
    MlpGemmArgs a = gemm_args();
    a.A = G, a.lda = ldg, a.B = ones, a.ldb = 1;
    a.M = N, a.N = 1, a.K = nbp, a.scale = scale;
    a.Cf = out, a.ldcf = 1;
    gemm<false, false>(a, st);
}

// Softmax cross-entropy, then the gradients of mlp.rs:302-345 into net.G.
void mlp_backward(Mlp &net, const Batch &b, const int32_t *cells, int k, int nb, int nbp,
                  float alpha, double *row_loss, hipStream_t st) {
    const size_t sm = (size_t)net.cp * 4;
    FL_HIP(hipFuncSetAttribute((const void *)mlp_softmax_xent_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(mlp_softmax_xent_kernel, dim3((unsigned)nbp), dim3(256), sm, st, b.z,
                       (int64_t)net.cp, net.C, net.cp, cells, k, nb, b.g3, (int64_t)net.cp,
                       row_loss);
    check_launch();
    const float inv = 1.f / (float)nb;
    float *gp = net.g();
    // layer 3: dW3 = a2^T g3 / nb + alpha W3, db3 = mean(g3)
    MlpGemmArgs a = gemm_args();
    a.A = b.a2, a.lda = net.H, a.B = b.g3, a.ldb = net.cp;
    a.M = net.H, a.N = net.cp, a.K = nbp, a.scale = inv;
    if (alpha != 0.f) a.addf = net.p() + net.oW3, a.ldadd = net.cp, a.addscale = alpha;
    a.Cf = gp + net.oW3, a.ldcf = net.cp;
    gemm<false, false>(a, st);
    bias_grad(b.g3, net.cp, net.cp, nbp, inv, gp + net.ob3, b.ones, st);
    // delta2 = (g3 W3^T) * (a2 > 0)
    a = gemm_args();
    a.A = b.g3, a.lda = net.cp, a.B = net.w3b(), a.ldb = net.cp;
    a.M = nbp, a.N = net.H, a.K = net.cp;
    a.mask = b.a2, a.ldmask = net.H, a.Cb = b.d2, a.ldcb = net.H;
    gemm<true, true>(a, st);
    // layer 2
    a = gemm_args();
    a.A = b.a1, a.lda = net.H, a.B = b.d2, a.ldb = net.H;
    a.M = net.H, a.N = net.H, a.K = nbp, a.scale = inv;
    if (alpha != 0.f) a.addf = net.p() + net.oW2, a.ldadd = net.H, a.addscale = alpha;
    a.Cf = gp + net.oW2, a.ldcf = net.H;
    gemm<false, false>(a, st);
    bias_grad(b.d2, net.H, net.H, nbp, inv, gp + net.ob2, b.ones, st);
    a = gemm_args();
    a.A = b.d2, a.lda = net.H, a.B = net.w2b(), a.ldb = net.H;
    a.M = nbp, a.N = net.H, a.K = net.H;
    a.mask = b.a1, a.ldmask = net.H, a.Cb = b.d1, a.ldcb = net.H;
    gemm<true, true>(a, st);
    // layer 1
    a = gemm_args();
    a.A = b.xb, a.lda = net.dp, a.B = b.d1, a.ldb = net.H;
    a.M = net.dp, a.N = net.H, a.K = nbp, a.scale = inv;
    if (alpha != 0.f) a.addf = net.p() + net.oW1, a.ldadd = net.H, a.addscale = alpha;
    a.Cf = gp + net.oW1, a.ldcf = net.H;
    gemm<false, false>(a, st);
    bias_grad(b.d1, net.H, net.H, nbp, inv, gp + net.ob1, b.ones, st);
    check_launch();
}

// mlp.rs:372-409 with beta1 0.9, beta2 0.999, eps 1e-8.
void mlp_adam(Mlp &net, float lr, hipStream_t st) {
    net.t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)net.t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)net.t));
    struct Seg {
        int64_t off, n;
        uint16_t *bf;
    } segs[6] = {{net.oW1, (int64_t)net.dp * net.H, net.w1b()},
                 {net.ob1, net.H, nullptr},
                 {net.oW2, (int64_t)net.H * net.H, net.w2b()},
                 {net.ob2, net.H, nullptr},
                 {net.oW3, (int64_t)net.H * net.cp, net.w3b()},
                 {net.ob3, net.cp, nullptr}};
    for (const Seg &s : segs)
        hipLaunchKernelGGL(mlp_adam_kernel, dim3(grid1(s.n, 256)), dim3(256), 0, st,
                           net.p() + s.off, net.g() + s.off, (float *)net.Mo.ptr() + s.off,
                           (float *)net.V.ptr() + s.off, s.n, lr, b1, b2, eps, bc1, bc2, s.bf);
    check_launch();
}

// Stable counting sort of n labels in [0, k): order, first row and count
// per label (the centroid update's sort, kmeans_kernels.hpp).
void counting_sort(const int32_t *labels, int64_t n, int64_t k, int32_t *order, int64_t *first,
                   int32_t *count, Workspace &ws, hipStream_t st) {
    const int64_t nblk = std::max<int64_t>(1, ceil_div(n, UPD_RB));
    int32_t *blockcnt = ws.get<int32_t>("cs_blockcnt", nblk * k);
    int32_t *blockbase = ws.get<int32_t>("cs_blockbase", nblk * k);
    const size_t hist_b = (size_t)k * 4, ctr_b = (size_t)k * 8;
    FL_HIP(hipFuncSetAttribute((const void *)km_block_hist_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)hist_b));
    FL_HIP(hipFuncSetAttribute((const void *)km_scatter_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctr_b));
    hipLaunchKernelGGL(km_block_hist_kernel, dim3((unsigned)nblk), dim3(256), hist_b, st, labels,
                       n, k, blockcnt);
    hipLaunchKernelGGL(km_col_scan_kernel, dim3((unsigned)k), dim3(256), 0, st, blockcnt, nblk, k,
                       blockbase, count);
    hipLaunchKernelGGL(km_offsets_kernel, dim3(1), dim3(256), 0, st, count, k, first);
    hipLaunchKernelGGL(km_scatter_kernel, dim3((unsigned)nblk), dim3(64), ctr_b, st, labels, n, k,
                       label_bits(k), blockbase, first, order);
    check_launch();
}

fl_status check_dtype(int32_t dtype) {
    if (dtype != FL_F32 && dtype != FL_F64)
        return fail(FL_INVALID_PARAMETER, "InvalidParameter: dtype: must be FL_F32 or FL_F64");
    return FL_OK;
}

}  // namespace

struct fl_router {
    fl_knn *knn = nullptr;
    const void *X = nullptr;
    int32_t dtype = FL_F64;
    int64_t n = 0, d = 0;
    int C = 0;
    DevBuf centers, radius, cdist, cell_first, cell_count, row_id, xs, label32, row_cell;
    DevBuf cell_by_size;   // cells by row count, largest first (scan work order)
    // BF16x3 MFMA bounds: the centres' bf16 hi / lo rows (padded to mfma_dp
    // features) and norms; mfma: -1 not built yet, 0 not usable (d or the
    // centres' range), 1 ready
    DevBuf mh, ml, mn2, mn1;
    int mfma = -1, mfma_dp = 0;
    Mlp net;
    hipStream_t stream = nullptr;
    Workspace ws;
    ~fl_router() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

struct RtAcc {
    double ms_route = 0, ms_bounds = 0, ms_scan = 0, ms_merge = 0, ms_fb = 0;
    int64_t n_fb = 0, n_rescan = 0, rows = 0;
};

template <typename TX>
void build_cells(fl_router *r, uint64_t seed, int64_t iters) {
    hipStream_t st = r->stream;
    const int64_t n = r->n, d = r->d, C = r->C;
    const TX *X = (const TX *)r->X;
    // initial centres: C distinct rows drawn with the library's generator
    // (the first C steps of a Fisher-Yates shuffle of 0..n)
    Rng rng(seed);
    std::vector<int32_t> pick((size_t)C);
    {
        std::unordered_map<int64_t, int64_t> pool;
        auto at = [&](int64_t i) {
            auto it = pool.find(i);
            return it == pool.end() ? i : it->second;
        };
        for (int64_t i = 0; i < C; ++i) {
            const int64_t j = i + rng.range(n - i);
            const int64_t vi = at(i), vj = at(j);
            pool[i] = vj;
            pool[j] = vi;
            pick[(size_t)i] = (int32_t)vj;
        }
    }
    int32_t *dpick = r->ws.get<int32_t>("pick", C);
    double *init = r->ws.get<double>("init", C * d);
    FL_HIP(hipMemcpyAsync(dpick, pick.data(), (size_t)C * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(rt_rows_f64_kernel<TX>, dim3(grid1(C * d, 256)), dim3(256), 0, st, X,
                       (int)d, dpick, C, init);
    check_launch();
    FL_HIP(hipStreamSynchronize(st));
    double *cent = (double *)r->centers.get((size_t)(C * d) * 8);
    int64_t *lab64 = r->ws.get<int64_t>("lab64", n);
    fl_kmeans_params kp;
    fl_kmeans_params_default(&kp, C);
    kp.max_iter = iters;
    kp.tol = 0.0;
    kp.n_init = 1;
    kp.precision = FL_PREC_BF16X3;
    double inertia = 0.0;
    int64_t n_iter = 0;
    fl_status s = fl_kmeans_fit_device(r->X, r->dtype, n, d, &kp, init, cent, lab64, &inertia,
                                       &n_iter, nullptr);
    if (s != FL_OK) throw StatusError{s, std::string("router cells: ") + fl_last_error()};
    // One more assignment against the final centres: the cells are then the
    // exact nearest-centre cells of `cent` (the fit's labels belong to the
    // centres before its last update), which the bisector bound needs.
    kp.max_iter = 1;
    double *next = r->ws.get<double>("cent_next", C * d);
    s = fl_kmeans_fit_device(r->X, r->dtype, n, d, &kp, cent, next, lab64, &inertia, &n_iter,
                             nullptr);
    if (s != FL_OK) throw StatusError{s, std::string("router cells: ") + fl_last_error()};
    double *cd = (double *)r->cdist.get((size_t)(C * C) * 8);
    hipLaunchKernelGGL(rt_center_dist_kernel, dim3(grid1(C * C, 256)), dim3(256), 0, st, cent,
                       (int)C, (int)d, cd);
    check_launch();
    int32_t *lab = (int32_t *)r->label32.get((size_t)n * 4);
    hipLaunchKernelGGL(rt_labels32_kernel, dim3(grid1(n, 256)), dim3(256), 0, st, lab64, n, lab);
    check_launch();
    int32_t *order = (int32_t *)r->row_id.get((size_t)n * 4);
    int64_t *first = (int64_t *)r->cell_first.get((size_t)C * 8);
    int32_t *count = (int32_t *)r->cell_count.get((size_t)C * 4);
    counting_sort(lab, n, C, order, first, count, r->ws, st);
    TX *xs = (TX *)r->xs.get((size_t)(n * d) * sizeof(TX));
    int32_t *rc = (int32_t *)r->row_cell.get((size_t)n * 4);
    hipLaunchKernelGGL(rt_gather_sorted_kernel<TX>, dim3(grid1(n * d, 256)), dim3(256), 0, st, X,
                       n, (int)d, order, lab, xs, rc);
    unsigned long long *mf = r->ws.get<unsigned long long>("maxfold", C);
    FL_HIP(hipMemsetAsync(mf, 0, (size_t)C * 8, st));
    hipLaunchKernelGGL(rt_radius_kernel<TX>, dim3(grid1(n, 256)), dim3(256), 0, st, xs, n, (int)d,
                       rc, cent, mf);
    double *rad = (double *)r->radius.get((size_t)C * 8);
    hipLaunchKernelGGL(rt_radius_finish_kernel, dim3(grid1(C, 256)), dim3(256), 0, st, mf, count,
                       (int)C, rad);
    check_launch();
    // the scan's work items go out largest cell first, so that a launch does
    // not end waiting on a large cell whose items started last
    std::vector<int32_t> hc((size_t)C), perm((size_t)C);
    FL_HIP(hipMemcpyAsync(hc.data(), count, (size_t)C * 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    for (int64_t c = 0; c < C; ++c) perm[(size_t)c] = (int32_t)c;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int32_t a, int32_t b) { return hc[(size_t)a] > hc[(size_t)b]; });
    int32_t *pd = (int32_t *)r->cell_by_size.get((size_t)C * 4);
    FL_HIP(hipMemcpyAsync(pd, perm.data(), (size_t)C * 4, hipMemcpyHostToDevice, st));
    FL_HIP(hipStreamSynchronize(st));
}

// Lower-bound slack: the fold's d + 2 roundings and the sqrt (relative
// (d + 4) u on a distance), plus the roundings of rt_lower_bound itself.
double bound_slack(int64_t d) { return (double)(d + 8) * std::ldexp(1.0, -52); }

template <int K>
void launch_bounds_k(const RtBoundsArgs &a, int32_t dtype, hipStream_t st) {
    const size_t sm = rt_lds_bytes(K);
    const unsigned grid = (unsigned)ceil_div(a.nq, RT_Q);
    if (dtype == FL_F32) {
        FL_HIP(hipFuncSetAttribute((const void *)rt_bounds_kernel<float, K>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        hipLaunchKernelGGL((rt_bounds_kernel<float, K>), dim3(grid), dim3(256), sm, st,
                           a);
    } else {
        FL_HIP(hipFuncSetAttribute((const void *)rt_bounds_kernel<double, K>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        hipLaunchKernelGGL((rt_bounds_kernel<double, K>), dim3(grid), dim3(256), sm, st,
                           a);
    }
    check_launch();
}

template <int K>
void launch_bounds_mfma_k(const RtBoundsMfmaArgs &m, hipStream_t st) {
    const size_t sm = rtm_lds_bytes(K, m.dp);
    FL_HIP(hipFuncSetAttribute((const void *)rt_bounds_mfma_kernel<K>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((rt_bounds_mfma_kernel<K>), dim3((unsigned)ceil_div(m.b.nq, RT_Q)),
                       dim3(256), sm, st, m);
    check_launch();
}

// The centres' side of the MFMA bounds, once per router: usable when the
// padded width fits the kernel's LDS plan and every centre is finite with
// magnitude <= 2^56 (the BF16 split's range).
void prepare_mfma_bounds(fl_router *r) {
    if (r->mfma >= 0) return;
    hipStream_t st = r->stream;
    const int d = (int)r->d, C = r->C;
    const int dp = (int)ceil_div(d, 32) * 32;
    r->mfma = 0;
    if (dp > RTM_MAXDP) return;
    const double *cen = (const double *)r->centers.ptr();
    unsigned int *flag = r->ws.get<unsigned int>("mfma_flag", 1);
    FL_HIP(hipMemsetAsync(flag, 0, 4, st));
    hipLaunchKernelGGL(range_flags_kernel<double>, dim3(64), dim3(256), 0, st, cen,
                       (int64_t)C * d, flag);
    check_launch();
    unsigned int hf = 0;
    FL_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (hf & 1u) return;
    uint16_t *mh = (uint16_t *)r->mh.get((size_t)C * dp * 2);
    uint16_t *ml = (uint16_t *)r->ml.get((size_t)C * dp * 2);
    double *n2 = (double *)r->mn2.get((size_t)C * 8);
    double *n1 = (double *)r->mn1.get((size_t)C * 8);
    hipLaunchKernelGGL(bf16_split_kernel<double>, dim3(grid1((int64_t)C * dp, 256)), dim3(256), 0,
                       st, cen, (int64_t)C, d, dp, (unsigned short *)mh, (unsigned short *)ml);
    hipLaunchKernelGGL(row_norms_kernel<double>, dim3(grid1(C, 256)), dim3(256), 0, st, cen,
                       (int64_t)C, d, n2, n1);
    check_launch();
    FL_HIP(hipStreamSynchronize(st));
    r->mfma_dp = dp;
    r->mfma = 1;
}

// Per query, the kb smallest (lower bound, cell) over all cells: first the
// exact nearest centre, then every cell's ball and bisector bounds (from
// BF16x3 MFMA lower bounds of the folds when the data's range allows, else
// from the exact folds).  With ckey / cid, also the n_probe cells with the
// nearest centres.
template <typename TQ>
void smallest_bounds(fl_router *r, const TQ *Q, int64_t nq, int kb, double *bkey, int32_t *bid,
                     int n_probe = 0, double *ckey = nullptr, int32_t *cid = nullptr,
                     bool allow_mfma = true) {
    hipStream_t st = r->stream;
    const int32_t qdt = std::is_same<TQ, float>::value ? FL_F32 : FL_F64;
    double *nfold = r->ws.get<double>("near_fold", nq);
    int32_t *ncell = r->ws.get<int32_t>("near_cell", nq);
    RtBoundsArgs a;
    a.Q = Q, a.nq = nq, a.d = (int)r->d, a.centers = (const double *)r->centers.ptr();
    a.radius = (const double *)r->radius.ptr(), a.cdist = (const double *)r->cdist.ptr();
    a.C = r->C, a.ep = bound_slack(r->d);
    // the nearest centre and its exact fold: the KMeans assignment step
    nearest_centers(Q, qdt, nq, (int)r->d, (const double *)r->centers.ptr(), r->C, ncell, nfold,
                    st);
    a.near_fold = nfold, a.near_cell = ncell, a.kb = kb, a.centre_key = 0;
    a.out_key = bkey, a.out_id = bid;
    if (allow_mfma) prepare_mfma_bounds(r);
    bool use_mfma = allow_mfma && r->mfma == 1 && nq > 0;
    RtBoundsMfmaArgs m;
    if (use_mfma) {
        // the queries' range, then their bf16 hi / lo rows and norms
        const int d = (int)r->d, dp = r->mfma_dp;
        unsigned int *flag = r->ws.get<unsigned int>("mfma_flag", 1);
        FL_HIP(hipMemsetAsync(flag, 0, 4, st));
        hipLaunchKernelGGL(range_flags_kernel<TQ>, dim3(256), dim3(256), 0, st, Q, nq * d, flag);
        check_launch();
        unsigned int hf = 0;
        FL_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        use_mfma = (hf & 1u) == 0;
        if (use_mfma) {
            uint16_t *qh = r->ws.get<uint16_t>("mq_h", nq * dp);
            uint16_t *ql = r->ws.get<uint16_t>("mq_l", nq * dp);
            double *qn2 = r->ws.get<double>("mq_n2", nq);
            double *qn1 = r->ws.get<double>("mq_n1", nq);
            hipLaunchKernelGGL(bf16_split_kernel<TQ>, dim3(grid1(nq * dp, 256)), dim3(256), 0, st,
                               Q, nq, d, dp, (unsigned short *)qh, (unsigned short *)ql);
            hipLaunchKernelGGL(row_norms_kernel<TQ>, dim3(grid1(nq, 256)), dim3(256), 0, st, Q,
                               nq, d, qn2, qn1);
            check_launch();
            const Bound bd = bound_bf16(d, 3);
            m.b = a, m.qh = qh, m.ql = ql, m.qn2 = qn2, m.qn1 = qn1;
            m.ch = (const uint16_t *)r->mh.ptr(), m.cl = (const uint16_t *)r->ml.ptr();
            m.cn2 = (const double *)r->mn2.ptr(), m.cn1 = (const double *)r->mn1.ptr();
            m.dp = dp, m.alpha = bd.alpha, m.beta = bd.beta;
            m.eta = (double)(3 * d + 1) * std::ldexp(1.0, -100);
        }
    }
    auto run = [&](const RtBoundsArgs &ab, int kk) {
        if (use_mfma) {
            // each of the four sub-lists keeps K >= kk entries, which is all
            // the merge of the kk smallest needs
            m.b = ab;
            if (kk <= 4)
                launch_bounds_mfma_k<4>(m, st);
            else if (kk <= 8)
                launch_bounds_mfma_k<8>(m, st);
            else if (kk <= 16)
                launch_bounds_mfma_k<16>(m, st);
            else
                launch_bounds_mfma_k<32>(m, st);
        } else if (kk <= 16) {
            launch_bounds_k<16>(ab, qdt, st);
        } else {
            launch_bounds_k<32>(ab, qdt, st);
        }
    };
    run(a, kb);
    if (ckey) {
        // the same pass keyed by the fold with each centre: the n_probe
        // nearest non-empty cells (ties: lower cell)
        a.centre_key = 1, a.kb = n_probe, a.out_key = ckey, a.out_id = cid;
        run(a, n_probe);
    }
}

template <typename TQ, typename TX, int K>
void launch_scan_k(const RtScanArgs &a, int64_t items, hipStream_t st) {
    const size_t sm = rt_lds_bytes(K);
    FL_HIP(hipFuncSetAttribute((const void *)rt_scan_kernel<TQ, TX, K>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL((rt_scan_kernel<TQ, TX, K>), dim3((unsigned)items), dim3(256), sm, st, a);
    check_launch();
}

template <int M>
void launch_top_cells(const float *Z, int64_t ldz, int64_t rows, int C, int n_probe,
                      int32_t *out, hipStream_t st) {
    hipLaunchKernelGGL((rt_top_cells_kernel<M>), dim3(grid1(rows, 4)), dim3(256), 0, st, Z, ldz,
                       rows, C, n_probe, out);
    check_launch();
}

// Routed search of nq device queries (one pass; nq * n_probe < 2^31).
// Queries whose certificate fails are searched again: up to kFullScanMax of
// them over every cell whose ball bound does not exceed the k-th distance
// found so far, more over the cells with the smallest bounds (8 or twice
// n_probe, doubling up to min(31, n_cells)); what still fails (non-finite
// queries) goes to the brute force.  A rescan starts
// from the k best its parent found (prev_d, prev_i) and skips the cells its
// parent scanned (prev_routes, prev_np per query) -- its parent only: an
// all-cells pass after a widened rescan can scan a cell of the first search
// again, and the merge then drops the (distance, index) pairs it already
// holds, so the result is the same.
template <typename TQ, typename TX>
void route_search(fl_router *r, const TQ *Q, int64_t nq, int k, int n_probe, int route,
                  double *dist, int64_t *idx, RtAcc &acc, int depth = 0,
                  const double *prev_d = nullptr, const int64_t *prev_i = nullptr,
                  const int32_t *prev_routes = nullptr, int prev_np = 0) {
    hipStream_t st = r->stream;
    const int d = (int)r->d, C = r->C;
    PhaseTimer tm(st);
    int32_t *routes = r->ws.get<int32_t>("routes", nq * n_probe);
    // 1. network choice
    int t = tm.begin();
    if (route == FL_ROUTE_NETWORK) {
        Mlp &net = r->net;
        for (int64_t q0 = 0; q0 < nq; q0 += kInferRows) {
            const int nb = (int)std::min<int64_t>(kInferRows, nq - q0);
            const int nbp = (int)ceil_div(nb, 64) * 64;
            Batch b = batch_buffers(r->ws, net, nbp, false, st);
            hipLaunchKernelGGL(mlp_gather_rows_kernel<TQ>, dim3(grid1((int64_t)nbp * net.dp, 256)),
                               dim3(256), 0, st, Q, d, (const int32_t *)nullptr, q0, nb, nbp,
                               net.dp, b.xb);
            check_launch();
            mlp_forward(net, b, nbp, st);
            int32_t *out = routes + q0 * n_probe;
            if (n_probe <= 8)
                launch_top_cells<8>(b.z, net.cp, nb, C, n_probe, out, st);
            else if (n_probe <= 16)
                launch_top_cells<16>(b.z, net.cp, nb, C, n_probe, out, st);
            else
                launch_top_cells<32>(b.z, net.cp, nb, C, n_probe, out, st);
        }
    }
    tm.end(t);
    // 2. lower bounds: the n_probe + 1 smallest per query (none when every
    // cell is scanned: nothing is left to bound)
    const int kb = route == kRouteAll ? 0 : n_probe + 1;
    double *bkey = r->ws.get<double>("bkey", nq * kb);
    int32_t *bid = r->ws.get<int32_t>("bid", nq * kb);
    const int tb = tm.begin();
    if (route == kRouteAll) {
        hipLaunchKernelGGL(rt_routes_all_kernel, dim3(grid1(nq * n_probe, 256)), dim3(256), 0, st,
                           nq, n_probe, routes);
        check_launch();
    } else {
        if (route == FL_ROUTE_CENTRES) {
            double *ckey = r->ws.get<double>("ckey", nq * n_probe);
            int32_t *cid = r->ws.get<int32_t>("cid", nq * n_probe);
            smallest_bounds(r, Q, nq, kb, bkey, bid, n_probe, ckey, cid);
            hipLaunchKernelGGL(rt_routes_from_bounds_kernel, dim3(grid1(nq * n_probe, 256)),
                               dim3(256), 0, st, cid, n_probe, nq, n_probe, routes);
            check_launch();
        } else {
            smallest_bounds(r, Q, nq, kb, bkey, bid);
        }
        if (route == FL_ROUTE_BOUNDS) {
            hipLaunchKernelGGL(rt_routes_from_bounds_kernel, dim3(grid1(nq * n_probe, 256)),
                               dim3(256), 0, st, bid, kb, nq, n_probe, routes);
            check_launch();
        }
    }
    tm.end(tb);
    // 3-5. exact scan of the chosen cells and merge with the certificate.
    // A pass scans the pairs whose scan_routes entry is a cell (bucket C of
    // the counting sort gets no work items) and merges them with an earlier
    // k best (pd, pi) when given.
    const int64_t P = nq * n_probe;
    int32_t *pord = r->ws.get<int32_t>("pord", P);
    int64_t *pfirst = r->ws.get<int64_t>("pfirst", C + 1);
    int32_t *pcount = r->ws.get<int32_t>("pcount", C + 1);
    int32_t *wfirst = r->ws.get<int32_t>("wfirst", C);
    int32_t *wtotal = r->ws.get<int32_t>("wtotal", 1);
    double *pkey = r->ws.get<double>("pkey", P * k);
    int32_t *pid = r->ws.get<int32_t>("pid", P * k);
    uint8_t *flag = r->ws.get<uint8_t>("flag", nq);
    unsigned long long *rows = r->ws.get<unsigned long long>("rows", 1);
    FL_HIP(hipMemsetAsync(rows, 0, 8, st));
    std::vector<int> scan_ids, merge_ids;
    auto pass = [&](const int32_t *sroutes, const double *pd, const int64_t *pi) {
        const int ts = tm.begin();
        counting_sort(sroutes, P, C + 1, pord, pfirst, pcount, r->ws, st);
        hipLaunchKernelGGL(rt_work_scan_kernel, dim3(1), dim3(256), 0, st, pcount,
                           (const int32_t *)r->cell_by_size.ptr(), C, wfirst, wtotal);
        check_launch();
        int32_t items = 0;
        FL_HIP(hipMemcpyAsync(&items, wtotal, 4, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        int32_t *wcell = r->ws.get<int32_t>("wcell", items);
        int32_t *wstart = r->ws.get<int32_t>("wstart", items);
        hipLaunchKernelGGL(rt_work_fill_kernel, dim3(grid1(C, 256)), dim3(256), 0, st, pcount,
                           wfirst, C, wcell, wstart);
        check_launch();
        if (items > 0) {
            RtScanArgs a;
            a.Q = Q, a.Xs = r->xs.ptr(), a.row_id = (const int32_t *)r->row_id.ptr();
            a.cell_first = (const int64_t *)r->cell_first.ptr();
            a.cell_count = (const int32_t *)r->cell_count.ptr();
            a.pair_order = pord, a.pair_first = pfirst, a.pair_count = pcount;
            a.work_cell = wcell, a.work_start = wstart;
            a.d = d, a.n_probe = n_probe, a.k = k, a.out_key = pkey, a.out_id = pid;
            if (k <= 16)
                launch_scan_k<TQ, TX, 16>(a, items, st);
            else
                launch_scan_k<TQ, TX, 32>(a, items, st);
        }
        tm.end(ts);
        const int tmg = tm.begin();
        if (k <= 16)
            hipLaunchKernelGGL((rt_merge_certify_kernel<TQ, 16>), dim3(grid1(nq, 128)), dim3(128),
                               0, st, pkey, pid, nq, n_probe, k, routes, sroutes, C, pd, pi, bkey,
                               bid, kb, Q, d, (const int32_t *)r->cell_count.ptr(), dist, idx,
                               flag, rows);
        else
            hipLaunchKernelGGL((rt_merge_certify_kernel<TQ, 32>), dim3(grid1(nq, 128)), dim3(128),
                               0, st, pkey, pid, nq, n_probe, k, routes, sroutes, C, pd, pi, bkey,
                               bid, kb, Q, d, (const int32_t *)r->cell_count.ptr(), dist, idx,
                               flag, rows);
        check_launch();
        tm.end(tmg);
        scan_ids.push_back(ts);
        merge_ids.push_back(tmg);
    };
    if (prev_routes) {
        // a rescan: skip the cells the search before it scanned, and cells
        // whose bound exceeds that search's k-th distance
        int32_t *sr = r->ws.get<int32_t>("sroutes", P);
        if (route == kRouteAll)
            hipLaunchKernelGGL(rt_skip_all_kernel<TQ>, dim3(grid1(P, 256)), dim3(256), 0, st, Q,
                               d, nq, C, (const double *)r->centers.ptr(),
                               (const double *)r->radius.ptr(), bound_slack(d), prev_routes,
                               prev_np, prev_d, k, sr);
        else
            hipLaunchKernelGGL(rt_skip_routes_kernel, dim3(grid1(P, 256)), dim3(256), 0, st,
                               routes, nq, n_probe, prev_routes, prev_np, C,
                               route == FL_ROUTE_BOUNDS ? bkey : nullptr, kb, prev_d, k, sr);
        check_launch();
        pass(sr, prev_d, prev_i);
    } else if (n_probe >= 2 && route != kRouteAll) {
        // each query's first cell, then its other cells whose bound does not
        // exceed the k-th distance the first found
        int32_t *sr = r->ws.get<int32_t>("sroutes", P);
        hipLaunchKernelGGL(rt_first_slot_kernel, dim3(grid1(P, 256)), dim3(256), 0, st, routes,
                           P, n_probe, C, sr);
        check_launch();
        pass(sr, nullptr, nullptr);
        double *p1d = r->ws.get<double>("p1d", nq * k);
        int64_t *p1i = r->ws.get<int64_t>("p1i", nq * k);
        FL_HIP(hipMemcpyAsync(p1d, dist, (size_t)(nq * k) * 8, hipMemcpyDeviceToDevice, st));
        FL_HIP(hipMemcpyAsync(p1i, idx, (size_t)(nq * k) * 8, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(rt_skip_by_bound_kernel<TQ>, dim3(grid1(P, 256)), dim3(256), 0, st, Q,
                           d, nq, n_probe, routes, C, (const double *)r->centers.ptr(),
                           (const double *)r->radius.ptr(), (const double *)r->cdist.ptr(),
                           r->ws.get<double>("near_fold", nq), r->ws.get<int32_t>("near_cell", nq),
                           bound_slack(d), p1d, k, sr);
        check_launch();
        pass(sr, p1d, p1i);
    } else {
        pass(routes, nullptr, nullptr);
    }
    std::vector<uint8_t> hflag((size_t)nq);
    unsigned long long hrows = 0;
    FL_HIP(hipMemcpyAsync(hflag.data(), flag, (size_t)nq, hipMemcpyDeviceToHost, st));
    FL_HIP(hipMemcpyAsync(&hrows, rows, 8, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    acc.ms_route += tm.ms(t);
    acc.ms_bounds += tm.ms(tb);
    for (int id : scan_ids) acc.ms_scan += tm.ms(id);
    for (int id : merge_ids) acc.ms_merge += tm.ms(id);
    acc.rows += (int64_t)hrows;
    // 6. uncertified queries: a wider routed scan, then the brute force
    std::vector<int32_t> which;
    for (int64_t q = 0; q < nq; ++q)
        if (hflag[(size_t)q]) which.push_back((int32_t)q);
    const int64_t m = (int64_t)which.size();
    if (m == 0) return;
    WallClock wc;
    const std::string fp = "fb" + std::to_string(depth) + "_";
    int32_t *dw = r->ws.get<int32_t>(fp + "which", m);
    TQ *fq = r->ws.get<TQ>(fp + "q", m * d);
    double *fd = r->ws.get<double>(fp + "d", m * k);
    int64_t *fi = r->ws.get<int64_t>(fp + "i", m * k);
    FL_HIP(hipMemcpyAsync(dw, which.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(rt_gather_queries_kernel<TQ>, dim3(grid1(m * d, 256)), dim3(256), 0, st, Q,
                       d, dw, m, fq);
    check_launch();
    // what this search found and scanned: a rescan starts from it
    double *pd = r->ws.get<double>(fp + "pd", m * k);
    int64_t *pi = r->ws.get<int64_t>(fp + "pi", m * k);
    int32_t *pr = r->ws.get<int32_t>(fp + "pr", m * n_probe);
    hipLaunchKernelGGL(rt_gather_prev_kernel, dim3(grid1(m * std::max(k, n_probe), 256)),
                       dim3(256), 0, st, dist, idx, routes, n_probe, dw, m, k, pd, pi, pr);
    check_launch();
    // widening probes: max(8, 2 n_probe), doubling up to min(31, n_cells)
    const int wide = std::min(std::min(std::max(8, 2 * n_probe), kMaxProbe), C);
    if (route != kRouteAll && n_probe < C && m <= kFullScanMax) {
        // a few queries left: every cell whose ball bound does not exceed the
        // k-th distance found so far, with the same exact kernels; that
        // certifies every finite query
        if (depth == 0) acc.n_rescan += m;
        // in chunks whose per-pair lists (C pairs per query, k entries of 12
        // bytes) stay within kPairListBytes
        const int64_t mc = std::max<int64_t>(1, kPairListBytes / ((int64_t)C * k * 12));
        for (int64_t q0 = 0; q0 < m; q0 += mc) {
            const int64_t mq = std::min(mc, m - q0);
            RtAcc inner;
            route_search<TQ, TX>(r, fq + q0 * d, mq, k, C, kRouteAll, fd + q0 * k, fi + q0 * k,
                                 inner, depth + 1, pd + q0 * k, pi + q0 * k, pr + q0 * n_probe,
                                 n_probe);
            acc.rows += inner.rows;
            acc.n_fb += inner.n_fb;
        }
    } else if (route != kRouteAll && depth < 3 && (wide > n_probe || route != FL_ROUTE_BOUNDS)) {
        if (depth == 0) acc.n_rescan += m;
        RtAcc inner;
        route_search<TQ, TX>(r, fq, m, k, wide, FL_ROUTE_BOUNDS, fd, fi, inner, depth + 1, pd, pi,
                             pr, n_probe);
        acc.rows += inner.rows;
        acc.n_fb += inner.n_fb;
    } else {
        FL_HIP(hipStreamSynchronize(st));
        acc.n_fb += m;
        const int32_t qdt = std::is_same<TQ, float>::value ? FL_F32 : FL_F64;
        const fl_status s = fl_knn_kneighbors_device(r->knn, fq, qdt, m, d, k, fd, fi, nullptr);
        if (s != FL_OK) throw StatusError{s, fl_last_error()};
    }
    hipLaunchKernelGGL(rt_scatter_results_kernel, dim3(grid1(m * k, 256)), dim3(256), 0, st, fd,
                       fi, dw, m, k, dist, idx);
    check_launch();
    FL_HIP(hipStreamSynchronize(st));
    acc.ms_fb += wc.ms();
}

template <typename TQ>
void train_epochs(fl_router *r, const TQ *Q, int64_t nq, const int32_t *cells, int k,
                  const fl_router_train_params *p, double *loss_first, double *loss_last) {
    hipStream_t st = r->stream;
    Mlp &net = r->net;
    const int64_t B = p->batch_size;
    const int64_t nbp_max = ceil_div(std::min<int64_t>(B, nq), 64) * 64;
    Batch b = batch_buffers(r->ws, net, nbp_max, true, st);
    int32_t *perm_d = r->ws.get<int32_t>("perm", nq);
    int32_t *cells_b = r->ws.get<int32_t>("cells_b", nbp_max * k);
    double *row_loss = r->ws.get<double>("row_loss", nq + nbp_max);
    std::vector<int32_t> perm((size_t)nq);
    for (int64_t i = 0; i < nq; ++i) perm[(size_t)i] = (int32_t)i;
    std::vector<double> hl((size_t)nq);
    Rng rng(p->seed);
    for (int64_t ep = 0; ep < p->epochs; ++ep) {
        // mlp.rs:1306: shuffle the sample order every epoch
        for (int64_t i = nq - 1; i > 0; --i) {
            const int64_t j = rng.range(i + 1);
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
        FL_HIP(hipMemcpyAsync(perm_d, perm.data(), (size_t)nq * 4, hipMemcpyHostToDevice, st));
        for (int64_t s = 0; s < nq; s += B) {
            const int nb = (int)std::min<int64_t>(B, nq - s);
            const int nbp = (int)ceil_div(nb, 64) * 64;
            hipLaunchKernelGGL(mlp_gather_rows_kernel<TQ>, dim3(grid1((int64_t)nbp * net.dp, 256)),
                               dim3(256), 0, st, Q, (int)r->d, perm_d + s, (int64_t)0, nb, nbp,
                               net.dp, b.xb);
            hipLaunchKernelGGL(mlp_gather_cells_kernel, dim3(grid1((int64_t)nb * k, 256)),
                               dim3(256), 0, st, cells, k, perm_d + s, nb, cells_b);
            check_launch();
            mlp_forward(net, b, nbp, st);
            mlp_backward(net, b, cells_b, k, nb, nbp, (float)p->alpha, row_loss + s, st);
            mlp_adam(net, (float)p->learning_rate, st);
        }
        FL_HIP(hipMemcpyAsync(hl.data(), row_loss, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        // mlp.rs:1366,1403: the epoch loss is the mean of the batch losses
        double sum = 0.0;
        int64_t nbatch = 0;
        for (int64_t s = 0; s < nq; s += B) {
            const int64_t e = std::min<int64_t>(nq, s + B);
            double bl = 0.0;
            for (int64_t i = s; i < e; ++i) bl += hl[(size_t)i];
            sum += bl / (double)(e - s);
            ++nbatch;
        }
        const double l = nbatch ? sum / (double)nbatch : 0.0;
        if (ep == 0) *loss_first = l;
        *loss_last = l;
    }
}

void zero_report(fl_router_report *rep) {
    if (rep) std::memset(rep, 0, sizeof(*rep));
}

}  // namespace

extern "C" {

void fl_router_params_default(fl_router_params *p, int64_t n_cells) {
    if (!p) return;
    p->n_cells = n_cells;
    p->cell_iter = 10;
    p->hidden = 512;
    p->seed = 0;
}

void fl_router_train_params_default(fl_router_train_params *p) {
    if (!p) return;
    p->n_neighbors = 10;
    p->epochs = 5;
    p->batch_size = 4096;
    p->learning_rate = 1e-3;
    p->alpha = 0.0;
    p->seed = 0;
    p->target = FL_TARGET_CERTIFY;
}

fl_status fl_router_build(fl_knn_handle knn, const fl_router_params *p, fl_router_handle *out,
                          fl_router_report *rep) {
    return guarded([&]() -> fl_status {
        if (!out) return fail(FL_INVALID_PARAMETER, "InvalidParameter: out: must not be NULL");
        *out = nullptr;
        if (!knn) return fail(FL_INVALID_PARAMETER, "InvalidParameter: knn: must not be NULL");
        if (!p) return fail(FL_INVALID_PARAMETER, "InvalidParameter: params: must not be NULL");
        const void *X;
        int32_t dtype;
        int64_t n, d;
        knn_describe(knn, &X, &dtype, &n, &d);
        if (p->n_cells < 1)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_cells: must be at least 1");
        if (n < p->n_cells)
            return fail(FL_INSUFFICIENT_SAMPLES,
                        "InsufficientSamples: required " + std::to_string(p->n_cells) +
                            ", actual " + std::to_string(n) +
                            ": the router requires at least n_cells samples");
        if (p->n_cells > kMaxCells)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_cells: at most 8192");
        if (p->cell_iter < 1)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: cell_iter: must be at least 1");
        if (p->hidden < 64 || p->hidden > kMaxHidden || p->hidden % 64)
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: hidden: a multiple of 64 in 64..1024");
        WallClock wc;
        fl_router *r = new fl_router();
        try {
            r->knn = knn;
            r->X = X, r->dtype = dtype, r->n = n, r->d = d, r->C = (int)p->n_cells;
            FL_HIP(hipStreamCreate(&r->stream));
            // finite data with |x| <= 2^250: folds cannot overflow
            unsigned int *flag = r->ws.get<unsigned int>("range", 1);
            FL_HIP(hipMemsetAsync(flag, 0, 4, r->stream));
            if (dtype == FL_F32)
                hipLaunchKernelGGL(range_flags_kernel<float>, dim3(1024), dim3(256), 0, r->stream,
                                   (const float *)X, n * d, flag);
            else
                hipLaunchKernelGGL(range_flags_kernel<double>, dim3(1024), dim3(256), 0,
                                   r->stream, (const double *)X, n * d, flag);
            check_launch();
            unsigned int hf = 0;
            FL_HIP(hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, r->stream));
            FL_HIP(hipStreamSynchronize(r->stream));
            if (hf & 2u) {
                delete r;
                return fail(FL_INVALID_PARAMETER,
                            "InvalidParameter: X: the router needs finite data with |x| <= 2^250");
            }
            if (dtype == FL_F32)
                build_cells<float>(r, p->seed, p->cell_iter);
            else
                build_cells<double>(r, p->seed, p->cell_iter);
            r->net.shape((int)d, (int)p->hidden, (int)p->n_cells);
            r->net.xavier(p->seed ^ 0x5DEECE66DULL, r->stream);
            FL_HIP(hipStreamSynchronize(r->stream));
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
        zero_report(rep);
        if (rep) rep->ms_total = rep->ms_cells = wc.ms();
        return FL_OK;
    });
}

fl_status fl_router_train_device(fl_router_handle r, const void *dQ, int32_t dtype, int64_t nq,
                                 int64_t d, const fl_router_train_params *p,
                                 fl_router_report *rep) {
    return guarded([&]() -> fl_status {
        if (!r) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
        if (!p) return fail(FL_INVALID_PARAMETER, "InvalidParameter: params: must not be NULL");
        if (fl_status s = check_dtype(dtype)) return s;
        if (d != r->d)
            return fail(FL_SHAPE_MISMATCH, "ShapeMismatch: expected " + std::to_string(r->d) +
                                               ", actual " + std::to_string(d) +
                                               ": number of features must match training data");
        if (nq < 1 || !dQ)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: queries: need at least one");
        if (p->n_neighbors < 1 || p->n_neighbors > 64 || p->n_neighbors > r->n)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: 1..min(64, n)");
        if (p->batch_size < 1 || p->batch_size > 65536)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: batch_size: 1..65536");
        if (p->target != FL_TARGET_CERTIFY && p->target != FL_TARGET_NEIGHBORS)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: target: unknown");
        if (p->epochs < 0 || !(p->learning_rate > 0.0) || !(p->alpha >= 0.0))
            return fail(FL_INVALID_PARAMETER,
                        "InvalidParameter: epochs >= 0, learning_rate > 0, alpha >= 0");
        if (nq > 0x7fffffffLL)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: queries: at most 2^31-1");
        WallClock wc;
        zero_report(rep);
        const int k = (int)p->n_neighbors;
        // exact targets: the cells of each query's k nearest rows
        WallClock wt;
        // The exact neighbours come from this index's own search (bound
        // route over the widest probe, certified or brute force): the same
        // answer as the brute force, at a fraction of its cost.
        double *td = r->ws.get<double>("t_dist", nq * k);
        int64_t *ti = r->ws.get<int64_t>("t_idx", nq * k);
        if (k <= kMaxRoutedK) {
            const fl_status s = fl_router_kneighbors_device(
                r, dQ, dtype, nq, d, k, std::min<int64_t>(8, r->C), FL_ROUTE_BOUNDS, td, ti,
                nullptr);
            if (s != FL_OK) return s;
        } else {
            const fl_status s =
                fl_knn_kneighbors_device(r->knn, dQ, dtype, nq, d, k, td, ti, nullptr);
            if (s != FL_OK) return s;
        }
        int tw = k;
        int32_t *cells;
        if (p->target == FL_TARGET_NEIGHBORS) {
            cells = r->ws.get<int32_t>("t_cells", nq * k);
            hipLaunchKernelGGL(rt_target_cells_kernel, dim3(grid1(nq * k, 256)), dim3(256), 0,
                               r->stream, ti, nq * k, (const int32_t *)r->label32.ptr(), cells);
            check_launch();
        } else {
            tw = (int)std::min<int64_t>(kMaxTargets, r->C);
            double *bkey = r->ws.get<double>("t_bkey", nq * tw);
            int32_t *bid = r->ws.get<int32_t>("t_bid", nq * tw);
            if (dtype == FL_F32)
                smallest_bounds(r, (const float *)dQ, nq, tw, bkey, bid);
            else
                smallest_bounds(r, (const double *)dQ, nq, tw, bkey, bid);
            double *kth = r->ws.get<double>("t_kth", nq);
            hipLaunchKernelGGL(rt_kth_kernel, dim3(grid1(nq, 256)), dim3(256), 0, r->stream, td,
                               nq, k, kth);
            cells = r->ws.get<int32_t>("t_cells", nq * tw);
            hipLaunchKernelGGL(rt_certify_targets_kernel, dim3(grid1(nq * tw, 256)), dim3(256), 0,
                               r->stream, kth, nq, bkey, bid, tw, cells);
            check_launch();
        }
        FL_HIP(hipStreamSynchronize(r->stream));
        const double ms_t = wt.ms();
        WallClock wtr;
        double lf = 0.0, ll = 0.0;
        if (dtype == FL_F32)
            train_epochs<float>(r, (const float *)dQ, nq, cells, tw, p, &lf, &ll);
        else
            train_epochs<double>(r, (const double *)dQ, nq, cells, tw, p, &lf, &ll);
        FL_HIP(hipStreamSynchronize(r->stream));
        if (rep) {
            rep->ms_targets = ms_t;
            rep->ms_train = wtr.ms();
            rep->ms_total = wc.ms();
            rep->n_queries = nq;
            rep->loss_first = lf;
            rep->loss_last = ll;
        }
        return FL_OK;
    });
}

fl_status fl_router_cell_bounds_device(fl_router_handle r, const void *dQ, int32_t dtype,
                                       int64_t nq, int64_t d, int64_t kb, int32_t exact,
                                       double *d_keys, int32_t *d_ids) {
    return guarded([&]() -> fl_status {
        if (!r) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
        if (fl_status s = check_dtype(dtype)) return s;
        if (d != r->d)
            return fail(FL_SHAPE_MISMATCH, "ShapeMismatch: expected " + std::to_string(r->d) +
                                               ", actual " + std::to_string(d) +
                                               ": number of features must match training data");
        if (nq < 0) return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_queries: negative");
        if (kb < 1 || kb > 32 || kb > r->C)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: kb: 1..min(32, n_cells)");
        if (nq == 0) return FL_OK;
        if (!dQ || !d_keys || !d_ids)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        if (dtype == FL_F32)
            smallest_bounds(r, (const float *)dQ, nq, (int)kb, d_keys, d_ids, 0, nullptr, nullptr,
                            exact == 0);
        else
            smallest_bounds(r, (const double *)dQ, nq, (int)kb, d_keys, d_ids, 0, nullptr,
                            nullptr, exact == 0);
        FL_HIP(hipStreamSynchronize(r->stream));
        return FL_OK;
    });
}

fl_status fl_router_kneighbors_device(fl_router_handle r, const void *dQ, int32_t dtype,
                                      int64_t nq, int64_t d, int64_t k, int64_t n_probe,
                                      int32_t route, double *d_dist, int64_t *d_idx,
                                      fl_router_report *rep) {
    return guarded([&]() -> fl_status {
        if (!r) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
        if (fl_status s = check_dtype(dtype)) return s;
        if (d != r->d)
            return fail(FL_SHAPE_MISMATCH, "ShapeMismatch: expected " + std::to_string(r->d) +
                                               ", actual " + std::to_string(d) +
                                               ": number of features must match training data");
        if (nq < 0) return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_queries: negative");
        if (k == 0) k = knn_n_neighbors(r->knn);
        if (k == 0) return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_neighbors: must be at least 1");
        if (k > r->n)
            return fail(FL_INSUFFICIENT_SAMPLES,
                        "InsufficientSamples: required " + std::to_string(k) + ", actual " +
                            std::to_string(r->n) + ": kneighbors requires k <= n_samples");
        if (k < 1) return fail(FL_INVALID_PARAMETER, "InvalidParameter: k: must be at least 1");
        if (n_probe < 1 || n_probe > kMaxProbe || n_probe > r->C)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: n_probe: 1..min(31, n_cells)");
        if (route != FL_ROUTE_NETWORK && route != FL_ROUTE_BOUNDS && route != FL_ROUTE_CENTRES)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: route: unknown");
        if (nq > 0 && (!dQ || !d_dist || !d_idx))
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: buffers: must not be NULL");
        WallClock wc;
        if (k > kMaxRoutedK) {
            // beyond the routed scan's lists: every query takes the exact
            // brute force of the underlying handle (same results)
            const fl_status s =
                fl_knn_kneighbors_device(r->knn, dQ, dtype, nq, d, k, d_dist, d_idx, nullptr);
            if (s != FL_OK) return s;
            if (rep) {
                zero_report(rep);
                rep->ms_total = wc.ms();
                rep->ms_fallback = rep->ms_total;
                rep->n_queries = nq;
                rep->n_fallback = nq;
            }
            return FL_OK;
        }
        RtAcc acc;
        // queries per pass: the per-pair lists (n_probe pairs per query, k
        // entries of 12 bytes) stay within kPairListBytes
        const int64_t chunk = std::max<int64_t>(
            1, std::min<int64_t>(1 << 20, kPairListBytes / (n_probe * k * 12)));
        for (int64_t q0 = 0; q0 < nq; q0 += chunk) {
            const int64_t m = std::min(chunk, nq - q0);
            double *dd = d_dist + q0 * k;
            int64_t *di = d_idx + q0 * k;
            if (dtype == FL_F32) {
                const float *Q = (const float *)dQ + q0 * d;
                if (r->dtype == FL_F32)
                    route_search<float, float>(r, Q, m, (int)k, (int)n_probe, route, dd, di, acc);
                else
                    route_search<float, double>(r, Q, m, (int)k, (int)n_probe, route, dd, di, acc);
            } else {
                const double *Q = (const double *)dQ + q0 * d;
                if (r->dtype == FL_F32)
                    route_search<double, float>(r, Q, m, (int)k, (int)n_probe, route, dd, di, acc);
                else
                    route_search<double, double>(r, Q, m, (int)k, (int)n_probe, route, dd, di, acc);
            }
        }
        if (rep) {
            zero_report(rep);
            rep->ms_total = wc.ms();
            rep->ms_route = acc.ms_route;
            rep->ms_bounds = acc.ms_bounds;
            rep->ms_scan = acc.ms_scan;
            rep->ms_merge = acc.ms_merge;
            rep->ms_fallback = acc.ms_fb;
            rep->n_queries = nq;
            rep->n_fallback = acc.n_fb;
            rep->n_rescan = acc.n_rescan;
            rep->rows_scanned = acc.rows;
        }
        return FL_OK;
    });
}

int64_t fl_router_n_params(fl_router_handle r) { return r ? r->net.n_plain() : 0; }

fl_status fl_router_get_params(fl_router_handle r, float *params) {
    return guarded([&]() -> fl_status {
        if (!r || !params) return fail(FL_INVALID_PARAMETER, "InvalidParameter: NULL argument");
        std::vector<float> pad((size_t)r->net.np);
        FL_HIP(hipMemcpyAsync(pad.data(), r->net.p(), pad.size() * 4, hipMemcpyDeviceToHost,
                              r->stream));
        FL_HIP(hipStreamSynchronize(r->stream));
        r->net.unpad(pad, params);
        return FL_OK;
    });
}

fl_status fl_router_set_params(fl_router_handle r, const float *params) {
    return guarded([&]() -> fl_status {
        if (!r || !params) return fail(FL_INVALID_PARAMETER, "InvalidParameter: NULL argument");
        std::vector<float> pad;
        r->net.pad(params, pad);
        r->net.upload(pad, r->stream);
        FL_HIP(hipStreamSynchronize(r->stream));
        return FL_OK;
    });
}

fl_status fl_router_cells(fl_router_handle r, double *centers, double *radius, int64_t *counts,
                          int64_t *row_cell) {
    return guarded([&]() -> fl_status {
        if (!r) return fail(FL_INVALID_PARAMETER, "InvalidParameter: handle: must not be NULL");
        const int64_t C = r->C, n = r->n, d = r->d;
        if (centers)
            FL_HIP(hipMemcpy(centers, r->centers.ptr(), (size_t)(C * d) * 8, hipMemcpyDeviceToHost));
        if (radius) FL_HIP(hipMemcpy(radius, r->radius.ptr(), (size_t)C * 8, hipMemcpyDeviceToHost));
        if (counts) {
            std::vector<int32_t> c((size_t)C);
            FL_HIP(hipMemcpy(c.data(), r->cell_count.ptr(), (size_t)C * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < C; ++i) counts[i] = c[(size_t)i];
        }
        if (row_cell) {
            std::vector<int32_t> l((size_t)n);
            FL_HIP(hipMemcpy(l.data(), r->label32.ptr(), (size_t)n * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) row_cell[i] = l[(size_t)i];
        }
        return FL_OK;
    });
}

fl_status fl_router_mlp_gradients(int64_t d, int64_t hidden, int64_t n_cells, const float *params,
                                  const float *X, int64_t nb, const int32_t *cells, int64_t k,
                                  double alpha, double *loss, float *grads) {
    return guarded([&]() -> fl_status {
        if (!params || !X || !cells || !loss || !grads)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: NULL argument");
        if (d < 1 || d > 4096 || n_cells < 1 || n_cells > kMaxCells || hidden < 64 ||
            hidden > kMaxHidden || hidden % 64 || nb < 1 || nb > 65536 || k < 1 || k > 64)
            return fail(FL_INVALID_PARAMETER, "InvalidParameter: shape out of range");
        hipStream_t st;
        FL_HIP(hipStreamCreate(&st));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{st};
        Mlp net;
        net.shape((int)d, (int)hidden, (int)n_cells);
        std::vector<float> pad;
        net.pad(params, pad);
        net.upload(pad, st);
        Workspace ws;
        const int nbp = (int)ceil_div(nb, 64) * 64;
        Batch b = batch_buffers(ws, net, nbp, true, st);
        float *xd = ws.get<float>("x", nb * d);
        int32_t *cd = ws.get<int32_t>("c", nb * k);
        double *rl = ws.get<double>("rl", nbp);
        FL_HIP(hipMemcpyAsync(xd, X, (size_t)(nb * d) * 4, hipMemcpyHostToDevice, st));
        FL_HIP(hipMemcpyAsync(cd, cells, (size_t)(nb * k) * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mlp_gather_rows_kernel<float>, dim3(grid1((int64_t)nbp * net.dp, 256)),
                           dim3(256), 0, st, xd, (int)d, (const int32_t *)nullptr, (int64_t)0,
                           (int)nb, nbp, net.dp, b.xb);
        check_launch();
        mlp_forward(net, b, nbp, st);
        mlp_backward(net, b, cd, (int)k, (int)nb, nbp, (float)alpha, rl, st);
        std::vector<float> g((size_t)net.np);
        std::vector<double> hl((size_t)nb);
        FL_HIP(hipMemcpyAsync(g.data(), net.g(), g.size() * 4, hipMemcpyDeviceToHost, st));
        FL_HIP(hipMemcpyAsync(hl.data(), rl, (size_t)nb * 8, hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        net.unpad(g, grads);
        double s = 0.0;
        for (double v : hl) s += v;
        *loss = s / (double)nb;
        return FL_OK;
    });
}

void fl_router_free(fl_router_handle r) { delete r; }

}  // extern "C"
