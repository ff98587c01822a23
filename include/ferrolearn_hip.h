/*
 * ferrolearn_hip.h -- C ABI of the MI355X (gfx950) squared-Euclidean path.
 *
 * Drop-in boundary for ferrolearn's KMeans / MiniBatchKMeans Lloyd step and
 * brute-force KNeighborsClassifier / KNeighborsRegressor predict.  Every
 * entry point cites the reference interface it replaces.  Conventions:
 *
 *   - matrices are row-major, C-contiguous; labels and indices are int64_t
 *     (Rust usize); output buffers are caller-allocated;
 *   - host-pointer entry points copy inputs to the GPU and results back;
 *     the *_device entry points take device pointers already in HBM;
 *   - arguments are validated in the same order and with the same
 *     conditions as the Rust code before anything is launched; on error no
 *     output is written and fl_last_error() holds the context string;
 *   - labels, iteration counts and neighbour indices equal the reference's
 *     f64 CPU path bit for bit, whatever the precision mode: the reduced
 *     precision forms only filter candidates, and every decision is taken on
 *     exact f64 folds (kmeans.rs:159-163) with the reference's tie rules;
 *   - a handle is not shared between threads without external locking;
 *     distinct handles may run concurrently (one HIP stream each).
 */
#ifndef FERROLEARN_HIP_H
#define FERROLEARN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ferrolearn-core/src/error.rs:30-91 (FerroError), one code per variant used
 * on this path, plus FL_DEVICE_ERROR for HIP failures. */
typedef enum fl_status {
    FL_OK = 0,
    FL_SHAPE_MISMATCH = 1,        /* FerroError::ShapeMismatch          */
    FL_INSUFFICIENT_SAMPLES = 2,  /* FerroError::InsufficientSamples    */
    FL_INVALID_PARAMETER = 3,     /* FerroError::InvalidParameter       */
    FL_NUMERICAL_INSTABILITY = 4, /* FerroError::NumericalInstability   */
    FL_DEVICE_ERROR = 5           /* HIP runtime error (no Rust analogue) */
} fl_status;

/* Storage type of an input matrix.  Arithmetic is f64 either way: f32
 * values are widened exactly, as the reference's f64 path sees them. */
typedef enum fl_dtype { FL_F32 = 0, FL_F64 = 1 } fl_dtype;

/* How candidate distances are filtered before the exact f64 decision.
 *   FL_PREC_FP64  : f64 MFMA GEMM form ||a||^2 + ||b||^2 - 2 a.b (plain form)
 *   FL_PREC_BF16  : a.b on BF16 MFMA, FP32 accumulation
 *   FL_PREC_BF16X3: a.b as hi*hi + hi*lo + lo*hi BF16 MFMA products
 *   FL_PREC_EXACT : direct f64 fold for every pair (no filter)
 *   FL_PREC_FP8   : KMeans assignment: a.b on FP8 E4M3 MFMA (per-row
 *                   power-of-two scales for the samples, one scale for the
 *                   centres, FP32 accumulation), then the BF16x3 filter for
 *                   the rows its certificate leaves (d <= 128; otherwise,
 *                   and for kNN, it behaves as FL_PREC_BF16X3)
 * Results are identical in every mode; only speed and the share of rows
 * that need an exact re-score differ. */
typedef enum fl_precision {
    FL_PREC_FP64 = 0,
    FL_PREC_BF16 = 1,
    FL_PREC_BF16X3 = 2,
    FL_PREC_EXACT = 3,
    FL_PREC_FP8 = 4
} fl_precision;

/* knn.rs:72-79 (Weights) */
typedef enum fl_weights { FL_UNIFORM = 0, FL_DISTANCE = 1 } fl_weights;

/* GraphMode (ferrolearn-neighbors/src/graph.rs:26-32): the value stored for
 * every neighbour edge of a kneighbors graph. */
typedef enum fl_graph_mode { FL_GRAPH_CONNECTIVITY = 0, FL_GRAPH_DISTANCE = 1 } fl_graph_mode;

/* Per-call measurements, filled when a non-NULL pointer is passed.  Kernel
 * times come from HIP events recorded on the stream the kernels run on. */
typedef struct fl_report {
    double ms_total;        /* whole call, host wall clock                  */
    double ms_distance;     /* candidate-filter (distance) kernels          */
    double ms_finalize;     /* exact re-score, selection and certificates   */
    double ms_fallback;     /* exact full scans for uncertified rows        */
    double ms_update;       /* centroid update / vote / mean epilogues      */
    int64_t n_rows;         /* rows (queries or samples x iterations)       */
    int64_t n_candidates;   /* candidates re-scored exactly in f64          */
    int64_t n_fallback;     /* rows whose certificate failed (full scan)    */
    int64_t n_iter;         /* Lloyd iterations run (KMeans)                */
    int64_t n_dist_launches;/* launches of the distance kernel              */
    int64_t n_second;       /* KMeans FP8 form: rows the FP8 certificate
                               left to the BF16x3 filter                    */
} fl_report;

/* Last error context string of the calling thread ("" after success). */
const char *fl_last_error(void);

/* Library version: major*10000 + minor*100 + patch. */
int fl_version(void);

/* Number of visible HIP devices (0 when none; never fails). */
int fl_device_count(void);

/* Device memory helpers for callers that hold no GPU allocator of their own
 * (C / Rust hosts, tests): plain hipMalloc / hipMemcpy / hipFree on the
 * current device.  No reference counterpart. */
fl_status fl_device_malloc(void **ptr, uint64_t bytes);
fl_status fl_device_free(void *ptr);
fl_status fl_memcpy_to_device(void *dst, const void *src, uint64_t bytes);
fl_status fl_memcpy_to_host(void *dst, const void *src, uint64_t bytes);
fl_status fl_memcpy_device(void *dst, const void *src, uint64_t bytes);
fl_status fl_device_synchronize(void);

/* ------------------------------------------------------------------ */
/* KMeans -- ferrolearn-cluster/src/kmeans.rs                          */
/* ------------------------------------------------------------------ */

/* kmeans.rs:55-113 (KMeans builder).  Defaults from KMeans::new
 * (kmeans.rs:76-84): max_iter 300, tol 1e-4, n_init 10, no seed. */
/* The random stream behind k-means++ and MiniBatchKMeans' batches.
 * FL_RNG_LIBRARY: the library's own generator, with k-means++'s sums over
 * the distances taken in a fixed two-level order (blocks of 4096, then the
 * block sums) so that they run in parallel; seeds follow the reference
 * (random_state + run, random_state + run*1000003).
 * FL_RNG_STDRNG: the reference's stream itself -- StdRng::seed_from_u64 of
 * rand 0.9.4, i.e. ChaCha12 with rand_core's PCG32 key expansion, and its
 * random_range / random::<f64> draws -- with the reference's plain left
 * folds for the sums (kmeans.rs:200-228), so that a seeded fit draws what
 * the reference draws.  The sums are then sequential on the device, which
 * costs a few milliseconds per centre at n = 10^6. */
typedef enum fl_rng { FL_RNG_LIBRARY = 0, FL_RNG_STDRNG = 1 } fl_rng;

typedef struct fl_kmeans_params {
    int64_t n_clusters;
    int64_t max_iter;
    double tol;
    int64_t n_init;
    uint64_t random_state;  /* used when has_random_state != 0 */
    int32_t has_random_state;
    int32_t precision;      /* fl_precision for the assignment filter */
    int32_t rng;            /* fl_rng: generator of the k-means++ draws */
    int32_t reserved;       /* zero */
} fl_kmeans_params;

/* kmeans.rs:76-84 (KMeans::new) */
void fl_kmeans_params_default(fl_kmeans_params *p, int64_t n_clusters);

/* Fit<Array2<F>, ()> for KMeans (kmeans.rs:403-513).
 * init_centers: NULL runs the library's k-means++ per restart (seed
 * random_state + run, kmeans.rs:451,460, with the library's own generator);
 * otherwise n_init*k*d f64 initial centres, one block per restart, which
 * makes the run comparable with any other implementation started from the
 * same centres.  Outputs: centers_out k*d (post-update centres of the kept
 * run), labels_out n, inertia_out, n_iter_out (kmeans.rs:468-503). */
fl_status fl_kmeans_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                        const fl_kmeans_params *params,
                        const double *init_centers, double *centers_out,
                        int64_t *labels_out, double *inertia_out,
                        int64_t *n_iter_out, fl_report *report);

/* Same as fl_kmeans_fit with X, init_centers and every output in device
 * memory (inertia_out and n_iter_out stay host pointers). */
fl_status fl_kmeans_fit_device(const void *dX, int32_t dtype, int64_t n,
                               int64_t d, const fl_kmeans_params *params,
                               const double *d_init_centers,
                               double *d_centers_out, int64_t *d_labels_out,
                               double *inertia_out, int64_t *n_iter_out,
                               fl_report *report);

/* Predict<Array2<F>> for FittedKMeans (kmeans.rs:528-553): labels of X
 * against centres (k*d_centers f64).  ShapeMismatch if d != d_centers. */
fl_status fl_kmeans_predict(const double *centers, int64_t k,
                            int64_t d_centers, const void *X, int32_t dtype,
                            int64_t n, int64_t d, int64_t *labels_out);

/* Allocation-free form of fl_kmeans_predict.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  centers (k*d_centers f64), X (n*d
 * of dtype) and labels_out (n int64) are device pointers; `precision` picks
 * the candidate filter (fl_precision; labels are exact in every mode);
 * `stream` is a hipStream_t (NULL: the null stream).  The call allocates
 * nothing: every scratch buffer lives in one fixed 64 MiB device array of
 * the library (one per GPU, held until the process ends), and rows are
 * processed in chunks that fit it.  k*d beyond what that array holds (about
 * 16*k*d bytes plus the centre splits) gives InvalidParameter.  The call
 * synchronises its stream before it returns, so it cannot be captured in a
 * graph.  Calls that overlap on one GPU share the array and race: serialise
 * them. */
fl_status fl_kmeans_predict_device_noalloc(const double *centers, int64_t k,
                                           int64_t d_centers, const void *X,
                                           int32_t dtype, int64_t n, int64_t d,
                                           int32_t precision, int64_t *labels_out,
                                           void *stream);

/* Transform<Array2<F>> for FittedKMeans (kmeans.rs:555-618):
 * dist_out[i*k + c] = sqrt(fold(x_i, c_c)). */
fl_status fl_kmeans_transform(const double *centers, int64_t k,
                              int64_t d_centers, const void *X, int32_t dtype,
                              int64_t n, int64_t d, double *dist_out);

/* The library's k-means++ (kmeans.rs:169-253 restated with the library's
 * generator; see fl_kmeans_fit).  centers_out: k*d f64. */
fl_status fl_kmeans_plusplus(const void *X, int32_t dtype, int64_t n,
                             int64_t d, int64_t k, uint64_t seed,
                             double *centers_out);

/* Many independent KMeans fits in one call (no reference counterpart;
 * each problem is KMeans::fit, kmeans.rs:403-513).  Every problem's outputs
 * are those of fl_kmeans_fit_device on the same arguments, bit for bit,
 * whatever its precision field.  Problems with d <= 64 and k <= 1024 run
 * together on the device: k-means++ of all their restarts in lockstep, then
 * Lloyd iterations of every (problem, restart) run as one HIP graph of a
 * persistent exact-assignment kernel and a per-run update kernel, replayed
 * until every run has converged.  Other problems run as whole fits on
 * n_workers host threads.  X, init_centers (or NULL), centers_out and
 * labels_out are device pointers. */
typedef struct fl_kmeans_problem {
    const void *X;
    int32_t dtype;
    int64_t n;
    int64_t d;
    fl_kmeans_params params;
    const double *init_centers;
    double *centers_out;
    int64_t *labels_out;
    double inertia;   /* out */
    int64_t n_iter;   /* out */
    int32_t status;   /* out: this problem's fl_status */
} fl_kmeans_problem;

/* Returns FL_OK when every problem succeeded, else the first failing
 * problem's status (message names the problem).  n_workers <= 0 picks 8.
 * report: ms_total (wall), n_rows = rows assigned over every restart and
 * iteration of every problem, n_iter = sum of the kept runs' n_iter,
 * ms_distance = Lloyd iterations (device), ms_update = k-means++ (device),
 * n_dist_launches = graph replays. */
fl_status fl_kmeans_fit_batched(fl_kmeans_problem *problems, int64_t n_problems,
                                int32_t n_workers, fl_report *report);

/* Tile shapes of the batched assignment kernel (no reference counterpart).
 * Variant v folds U centres side by side for R rows per lane held in D
 * feature registers; table[4*b_d + b_k] is the variant used for problems
 * with d in bucket b_d (d <= 8, 16, 32, 64) and k in bucket b_k (k <= 16,
 * 64, 256, 1024), and must have that bucket's D.  The project's autotuner
 * (ferrolearn_amd.tuning) times the variants per bucket and loads its table
 * with fl_kmeans_batched_set_tuning.  Results never depend on the table. */
int32_t fl_kmeans_batched_n_variants(void);
fl_status fl_kmeans_batched_variant(int32_t v, int32_t *D, int32_t *U, int32_t *R);
fl_status fl_kmeans_batched_set_tuning(const int32_t *table);
fl_status fl_kmeans_batched_get_tuning(int32_t *table);

/* ------------------------------------------------------------------ */
/* MiniBatchKMeans -- ferrolearn-cluster/src/mini_batch_kmeans.rs      */
/* ------------------------------------------------------------------ */

/* mini_batch_kmeans.rs:77-93 (MiniBatchKMeans builder).  Defaults from
 * MiniBatchKMeans::new (mini_batch_kmeans.rs:108-118): batch_size 1024,
 * max_iter 100, tol 0, n_init 3, no seed. */
typedef struct fl_minibatch_params {
    int64_t n_clusters;
    int64_t batch_size;
    int64_t max_iter;
    double tol;
    int64_t n_init;
    uint64_t random_state;  /* used when has_random_state != 0 */
    int32_t has_random_state;
    int32_t precision;      /* fl_precision for the final full assignment */
    int32_t rng;            /* fl_rng: generator of k-means++ and the batches */
    int32_t reserved;       /* zero */
} fl_minibatch_params;

/* mini_batch_kmeans.rs:108-118 (MiniBatchKMeans::new) */
void fl_minibatch_params_default(fl_minibatch_params *p, int64_t n_clusters);

/* Fit<Array2<F>, ()> for MiniBatchKMeans (mini_batch_kmeans.rs:478-592).
 * Validation order: n_clusters == 0, batch_size == 0, n_init == 0 ->
 * InvalidParameter; n == 0 or n < n_clusters -> InsufficientSamples.  Run r
 * draws from the library's generator seeded with random_state + r*1000003:
 * k-means++ when init_centers is NULL (else n_init*k*d f64 initial centres),
 * then the batches.  Outputs as fl_kmeans_fit; the inertia is the left fold
 * of the final assignment's folds in row order (:366-373).  Predict and
 * transform of the fitted model are fl_kmeans_predict / fl_kmeans_transform
 * (mini_batch_kmeans.rs:607-682 is the same nearest-centre rule). */
fl_status fl_minibatch_kmeans_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                                  const fl_minibatch_params *params,
                                  const double *init_centers, double *centers_out,
                                  int64_t *labels_out, double *inertia_out,
                                  int64_t *n_iter_out, fl_report *report);

/* Same with X, init_centers and the centre / label outputs in device memory
 * (inertia_out and n_iter_out stay host pointers). */
fl_status fl_minibatch_kmeans_fit_device(const void *dX, int32_t dtype, int64_t n,
                                         int64_t d, const fl_minibatch_params *params,
                                         const double *d_init_centers,
                                         double *d_centers_out, int64_t *d_labels_out,
                                         double *inertia_out, int64_t *n_iter_out,
                                         fl_report *report);

/* ------------------------------------------------------------------ */
/* k-nearest neighbours -- ferrolearn-neighbors/src/knn.rs             */
/* ------------------------------------------------------------------ */

typedef struct fl_knn *fl_knn_handle;

/* Fit<Array2<F>, Array1<usize>> for KNeighborsClassifier (knn.rs:256-320)
 * with Algorithm::BruteForce (knn.rs:148).  Validation order: y_len != n
 * -> ShapeMismatch; n_neighbors == 0 -> InvalidParameter; n < n_neighbors
 * -> InsufficientSamples.  The handle owns a device copy of X and y, as the
 * fitted Rust model owns clones (knn.rs:311-318).  Any 1 <= n_neighbors
 * <= n: k <= 64 runs the BF16 / f64 candidate filters, larger k an exact
 * sort of every distance per query (the same (distance, index) order). */
fl_status fl_knn_fit_classifier(const void *X, int32_t dtype, int64_t n,
                                int64_t d, const int64_t *y, int64_t y_len,
                                int64_t n_neighbors, int32_t weights,
                                int32_t precision, fl_knn_handle *out);

/* Fit<Array2<F>, Array1<F>> for KNeighborsRegressor (knn.rs:660-718). */
fl_status fl_knn_fit_regressor(const void *X, int32_t dtype, int64_t n,
                               int64_t d, const double *y, int64_t y_len,
                               int64_t n_neighbors, int32_t weights,
                               int32_t precision, fl_knn_handle *out);

/* Predict for FittedKNeighborsClassifier (knn.rs:322-383): class labels. */
fl_status fl_knn_predict_classify(fl_knn_handle h, const void *Q,
                                  int32_t dtype, int64_t nq, int64_t d,
                                  int64_t *labels_out, fl_report *report);

/* Predict for FittedKNeighborsRegressor (knn.rs:720-781). */
fl_status fl_knn_predict_regress(fl_knn_handle h, const void *Q,
                                 int32_t dtype, int64_t nq, int64_t d,
                                 double *y_out, fl_report *report);

/* FittedKNeighborsClassifier::predict_proba (knn.rs:452-486):
 * proba_out is nq*n_classes, classes in fl_knn_classes order. */
fl_status fl_knn_predict_proba(fl_knn_handle h, const void *Q, int32_t dtype,
                               int64_t nq, int64_t d, double *proba_out,
                               fl_report *report);

/* kneighbors (knn.rs:823-834, 875-917): k == 0 means the fitted
 * n_neighbors.  Errors: ShapeMismatch on feature count, InvalidParameter
 * for k == 0 after defaulting, InsufficientSamples for k > n_fit.
 * dist_out / idx_out are nq*k, sorted by (distance, index). */
fl_status fl_knn_kneighbors(fl_knn_handle h, const void *Q, int32_t dtype,
                            int64_t nq, int64_t d, int64_t k,
                            double *dist_out, int64_t *idx_out,
                            fl_report *report);

/* Device-memory forms of the three predict calls: Q and the outputs are
 * device pointers. */
fl_status fl_knn_kneighbors_device(fl_knn_handle h, const void *dQ,
                                   int32_t dtype, int64_t nq, int64_t d,
                                   int64_t k, double *d_dist_out,
                                   int64_t *d_idx_out, fl_report *report);
fl_status fl_knn_predict_classify_device(fl_knn_handle h, const void *dQ,
                                         int32_t dtype, int64_t nq, int64_t d,
                                         int64_t *d_labels_out,
                                         fl_report *report);
fl_status fl_knn_predict_regress_device(fl_knn_handle h, const void *dQ,
                                        int32_t dtype, int64_t nq, int64_t d,
                                        double *d_y_out, fl_report *report);

/* Fit<Array2<F>, ()> for NearestNeighbors (nearest_neighbors.rs:198-228):
 * n_neighbors == 0 -> InvalidParameter; n < n_neighbors ->
 * InsufficientSamples.  Every Algorithm runs the GPU brute force (the tree
 * indexes return the same neighbours, nearest_neighbors.rs:657-700).  The
 * handle serves fl_knn_kneighbors(_device) and the graph calls below, where
 * k > n is InvalidParameter as in nearest_neighbors.rs:282-287 (the
 * supervised handles report InsufficientSamples, knn.rs:826-834).  A
 * negative k is a caller's explicit zero: InvalidParameter. */
fl_status fl_nn_fit(const void *X, int32_t dtype, int64_t n, int64_t d,
                    int64_t n_neighbors, int32_t precision, fl_knn_handle *out);

/* kneighbors_graph (graph.rs:91-123 via nearest_neighbors.rs / graph.rs
 * :225-304, any handle): the CSR matrix of nq rows by n_samples_fit columns
 * whose row i holds query i's k neighbours with ascending columns, valued
 * 1.0 (FL_GRAPH_CONNECTIVITY) or the distance (FL_GRAPH_DISTANCE).
 * indptr_out has nq+1 entries (i*k), indices_out and data_out nq*k.
 * Validation as fl_knn_kneighbors, then the mode. */
fl_status fl_knn_kneighbors_graph(fl_knn_handle h, const void *Q, int32_t dtype,
                                  int64_t nq, int64_t d, int64_t k, int32_t mode,
                                  int64_t *indptr_out, int64_t *indices_out,
                                  double *data_out, fl_report *report);
/* Same with Q and the three outputs in device memory. */
fl_status fl_knn_kneighbors_graph_device(fl_knn_handle h, const void *dQ,
                                         int32_t dtype, int64_t nq, int64_t d,
                                         int64_t k, int32_t mode,
                                         int64_t *d_indptr_out,
                                         int64_t *d_indices_out,
                                         double *d_data_out, fl_report *report);

/* Switch the candidate filter of a fitted handle (fl_precision). */
fl_status fl_knn_set_precision(fl_knn_handle h, int32_t precision);

/* HasClasses (ferrolearn-core/src/introspection.rs:44-50) and
 * n_samples_fit (knn.rs:838-842): sorted unique classes. */
int64_t fl_knn_n_classes(fl_knn_handle h);
fl_status fl_knn_classes(fl_knn_handle h, int64_t *classes_out);
int64_t fl_knn_n_samples_fit(fl_knn_handle h);

void fl_knn_free(fl_knn_handle h);

/* ------------------------------------------------------------------ */
/* BisectingKMeans -- ferrolearn-cluster/src/bisecting_kmeans.rs       */
/* ------------------------------------------------------------------ */

/* BisectingStrategy (bisecting_kmeans.rs:54-61). */
typedef enum fl_bisect_strategy {
    FL_BISECT_LARGEST_CLUSTER = 0,
    FL_BISECT_LARGEST_SSE = 1
} fl_bisect_strategy;

typedef struct fl_bisecting_params {
    int64_t n_clusters;
    int64_t max_iter;        /* per 2-means bisection */
    int64_t n_init;          /* 2-means runs per bisection, best SSE kept */
    uint64_t random_state;   /* used when has_random_state != 0; else 0 */
    int32_t has_random_state;
    int32_t strategy;        /* fl_bisect_strategy */
} fl_bisecting_params;

/* BisectingKMeans::new (bisecting_kmeans.rs:93-102): max_iter 300,
 * n_init 10, no seed, LargestCluster. */
void fl_bisecting_params_default(fl_bisecting_params *p, int64_t n_clusters);

/* Fit<Array2<F>, ()> for BisectingKMeans (bisecting_kmeans.rs:326-505):
 * centers_out (n_clusters * d f64; the first *n_clusters_out rows are
 * written), labels_out (n), inertia_out, n_clusters_out (fewer than
 * n_clusters when no cluster of two or more samples is left).  Labels,
 * centres and inertia equal the reference's bit for bit (f32 data widened
 * to f64).  Errors in the reference's order: n_clusters or n_init zero ->
 * InvalidParameter; n < n_clusters -> InsufficientSamples.  Where the
 * reference panics (no 2-means run below f64::MAX) or would loop without
 * end (1000 bisections in a row with an empty side, identical samples):
 * NumericalInstability.  Predict and transform of the fitted model are
 * fl_kmeans_predict / fl_kmeans_transform (bisecting_kmeans.rs:531-605
 * compute the same). */
fl_status fl_bisecting_kmeans_fit(const void *X, int32_t dtype, int64_t n,
                                  int64_t d, const fl_bisecting_params *params,
                                  double *centers_out, int64_t *labels_out,
                                  double *inertia_out, int64_t *n_clusters_out);
/* Same with X, centers_out and labels_out in device memory. */
fl_status fl_bisecting_kmeans_fit_device(const void *dX, int32_t dtype,
                                         int64_t n, int64_t d,
                                         const fl_bisecting_params *params,
                                         double *d_centers_out,
                                         int64_t *d_labels_out,
                                         double *inertia_out,
                                         int64_t *n_clusters_out);

/* ------------------------------------------------------------------ */
/* pairwise_distances_argmin(_min) -- ferrolearn-metrics/src/pairwise.rs */
/* ------------------------------------------------------------------ */

/* euclidean_distances (pairwise.rs:157-189): out (n x m, row-major f64)
 * holds sqrt(max((|x_i|^2 + |y_j|^2) - 2 x_i.y_j, 0)) for every pair, each
 * term a left fold, max returning the non-NaN operand; bit-exact.  Same
 * dtype rule and errors, in the same order, as below. */
fl_status fl_euclidean_distances(const void *X, int32_t dtype, int64_t n,
                                 int64_t dx, const void *Y, int64_t m,
                                 int64_t dy, double *out);
/* Same with X, Y and out in device memory. */
fl_status fl_euclidean_distances_device(const void *dX, int32_t dtype,
                                        int64_t n, int64_t dx, const void *dY,
                                        int64_t m, int64_t dy, double *d_out);

/* pairwise_distances_argmin_min (pairwise.rs:596-629) with
 * Metric::Euclidean (also pairwise_distances_argmin, :564-588, with
 * min_out == NULL): for each of the n rows of X the index of the closest of
 * the m rows of Y and that distance, where distances are
 * euclidean_distances' expansion (pairwise.rs:157-185)
 * sqrt(max((|x|^2 + |y|^2) - 2 x.y, 0)) with left folds, and ties go to the
 * lowest index.  X and Y share dtype (f32 data widened to f64, exactly).
 * Errors in the reference's order: dx != dy -> ShapeMismatch; n == 0 or
 * m == 0, then dx == 0 -> InsufficientSamples.  idx_out and min_out hold
 * n entries. */
fl_status fl_pairwise_distances_argmin_min(const void *X, int32_t dtype,
                                           int64_t n, int64_t dx,
                                           const void *Y, int64_t m,
                                           int64_t dy, int64_t *idx_out,
                                           double *min_out);
/* Same with X, Y and the outputs in device memory. */
fl_status fl_pairwise_distances_argmin_min_device(const void *dX, int32_t dtype,
                                                  int64_t n, int64_t dx,
                                                  const void *dY, int64_t m,
                                                  int64_t dy, int64_t *d_idx_out,
                                                  double *d_min_out);

/* ------------------------------------------------------------------ */
/* Cell router for exact kNN (no reference counterpart)                */
/* ------------------------------------------------------------------ */

/* A learned index over a fitted brute-force handle.  The database rows are
 * partitioned into cells by a coarse KMeans (kmeans.rs:403-513 with
 * initial centres drawn from the rows); each cell keeps its centre and a
 * radius bounding its rows' distance to the centre.  A perceptron
 * d -> hidden -> hidden -> n_cells (the layer, activation, loss and Adam
 * rules of ferrolearn-neural/src/mlp.rs:217-409, 1326-1340) is trained on
 * the cells that hold each training query's exact k nearest neighbours
 * (brute_force_knn, kdtree.rs:364-391).  A routed search scans the rows of
 * n_probe cells with the exact reference distance (each query's first cell,
 * then its other cells whose lower bound does not exceed the k-th distance
 * found) and certifies each query with the triangle inequality: every
 * unscanned cell's lower bound must exceed the k-th distance found.  Up to
 * 2048 queries that fail are searched again over every cell whose ball bound
 * does not exceed their k-th distance; more go first over the cells with the
 * smallest bounds (8 or twice n_probe, doubling up to min(31, n_cells)).  A
 * query that still fails (non-finite) runs the brute force, so results equal
 * fl_knn_kneighbors on the same handle, bit for bit. */
typedef struct fl_router *fl_router_handle;

typedef enum fl_route {
    FL_ROUTE_NETWORK = 0, /* cells with the largest router outputs          */
    FL_ROUTE_BOUNDS = 1,  /* cells with the smallest lower bounds (no network) */
    FL_ROUTE_CENTRES = 2  /* cells with the nearest centres (no network): the
                             classic inverted-file probe order               */
} fl_route;

typedef struct fl_router_params {
    int64_t n_cells;    /* coarse cells, 1..min(n, 8192)                     */
    int64_t cell_iter;  /* Lloyd iterations of the coarse KMeans (>= 1)      */
    int64_t hidden;     /* width of both hidden layers, multiple of 64, <= 1024 */
    uint64_t seed;      /* initial centres and network weights               */
} fl_router_params;

typedef struct fl_router_train_params {
    int64_t n_neighbors;  /* k of the exact targets, 1..64                  */
    int64_t epochs;
    int64_t batch_size;   /* 1..65536                                        */
    double learning_rate; /* Adam step (beta1 0.9, beta2 0.999, eps 1e-8)    */
    double alpha;         /* L2 penalty, grad_w += alpha * W                 */
    uint64_t seed;        /* per-epoch shuffles                              */
    int32_t target;       /* fl_router_target                                */
} fl_router_train_params;

typedef enum fl_router_target {
    /* all cells whose lower bound does not exceed the exact k-th distance:
     * the cells a search must scan to certify the exact answer (at most 32,
     * smallest bounds first) */
    FL_TARGET_CERTIFY = 0,
    /* the cells holding the exact k nearest rows */
    FL_TARGET_NEIGHBORS = 1
} fl_router_target;

typedef struct fl_router_report {
    double ms_total;
    double ms_cells;      /* build: coarse KMeans, sort, radii               */
    double ms_targets;    /* train: exact kNN of the training queries        */
    double ms_train;      /* train: epochs                                   */
    double ms_route;      /* search: network inference and cell choice       */
    double ms_bounds;     /* search: centre distances and lower bounds       */
    double ms_scan;       /* search: exact scan of the chosen cells          */
    double ms_merge;      /* search: merge and certificate                   */
    double ms_fallback;   /* search: rescans and brute force for queries whose
                             first certificate failed                      */
    int64_t n_queries;
    int64_t n_rescan;     /* queries whose first certificate failed: searched
                             again over more cells                          */
    int64_t n_fallback;   /* queries that still failed: brute force          */
    int64_t rows_scanned; /* database rows scanned by the routed search      */
    double loss_first;    /* train: mean loss of the first and last epochs   */
    double loss_last;
} fl_router_report;

void fl_router_params_default(fl_router_params *p, int64_t n_cells);
void fl_router_train_params_default(fl_router_train_params *p);

/* Build the cells over the rows of a fitted kNN handle (which must outlive
 * the router).  The data must be finite with magnitudes <= 2^250.  The
 * network starts from Xavier-uniform weights (mlp.rs:181-192) and zero
 * biases. */
fl_status fl_router_build(fl_knn_handle knn, const fl_router_params *params,
                          fl_router_handle *out, fl_router_report *report);

/* Train the network on nq device queries (dQ: nq*d of dtype). */
fl_status fl_router_train_device(fl_router_handle r, const void *dQ,
                                 int32_t dtype, int64_t nq, int64_t d,
                                 const fl_router_train_params *params,
                                 fl_router_report *report);

/* Routed kneighbors: same contract and results as fl_knn_kneighbors_device
 * (k == 0 means the handle's n_neighbors), scanning n_probe cells (1..31,
 * <= n_cells) chosen by `route` (fl_route).  k > 32 is beyond the routed
 * scan's lists: those calls run the exact brute force of the underlying
 * handle and report every query as a fallback. */
fl_status fl_router_kneighbors_device(fl_router_handle r, const void *dQ,
                                      int32_t dtype, int64_t nq, int64_t d,
                                      int64_t k, int64_t n_probe, int32_t route,
                                      double *d_dist_out, int64_t *d_idx_out,
                                      fl_router_report *report);

/* The kb (<= min(32, n_cells)) smallest per-cell lower bounds of each query
 * (d_keys, d_ids: nq x kb, ascending (key, cell)), as the search computes
 * them: from BF16x3 MFMA lower bounds of the folds (exact == 0; searches use
 * these whenever the data lie within the BF16 range) or from the exact folds
 * (exact != 0).  Every MFMA key is at most the exact key of the same cell.
 * For inspection and tests. */
fl_status fl_router_cell_bounds_device(fl_router_handle r, const void *dQ,
                                       int32_t dtype, int64_t nq, int64_t d,
                                       int64_t kb, int32_t exact, double *d_keys,
                                       int32_t *d_ids);

/* Network parameters in mlp.rs order: W1 (d x hidden), b1, W2 (hidden x
 * hidden), b2, W3 (hidden x n_cells), b3, as f32. */
int64_t fl_router_n_params(fl_router_handle r);
fl_status fl_router_get_params(fl_router_handle r, float *params);
fl_status fl_router_set_params(fl_router_handle r, const float *params);

/* The cells: centres (n_cells x d), radii (negative: empty cell), row
 * counts, and the cell of each database row. */
fl_status fl_router_cells(fl_router_handle r, double *centers, double *radius,
                          int64_t *counts, int64_t *row_cell);

/* One forward and backward pass of the router network on a host batch
 * (X: nb x d f32, cells: nb x k target cells), as used in training: mean
 * loss and the gradients in fl_router_get_params order. */
fl_status fl_router_mlp_gradients(int64_t d, int64_t hidden, int64_t n_cells,
                                  const float *params, const float *X,
                                  int64_t nb, const int32_t *cells, int64_t k,
                                  double alpha, double *loss, float *grads);

void fl_router_free(fl_router_handle r);

/* ------------------------------------------------------------------ */
/* Left fold of f64 values in index order                              */
/* ------------------------------------------------------------------ */

/* The reference's serial sums -- k-means++'s totals, threshold searches and
 * potentials (kmeans.rs:200,207-213,219-228) and MiniBatch's inertia
 * (mini_batch_kmeans.rs:366-373) are `fold(0, |acc, x| acc + x)` in index
 * order -- bit for bit the serial result: sequences under 8192 values by one
 * workgroup each, longer ones in 2048-value blocks over many compute units
 * with an exact walk over the blocks.  count sequences of n device values,
 * sequence b at d_v + b*ld:
 * d_sums[b] = the fold, and when d_prefix is not NULL, d_prefix[b*ld + i] =
 * the running sum after element i. */
fl_status fl_left_fold_device(const double *d_v, int64_t n, int64_t count, int64_t ld,
                              double *d_sums, double *d_prefix);

/* ------------------------------------------------------------------ */
/* Backend -- ferrolearn-core/src/backend.rs:37-123                    */
/* ------------------------------------------------------------------ */

/* Backend::gemm (backend.rs:41-44; faer impl backend_faer.rs:61-78):
 * C (m*n) = A (m*k) . B (k2*n), f64, row-major host pointers, on f64 MFMA
 * (summation order of k unspecified, as for a.dot(b)).  ShapeMismatch
 * (expected [m, k], actual [k2, n]) when k != k2; InvalidParameter for a
 * negative dimension or m above 65535 * 64. */
fl_status fl_backend_gemm_f64(const double *A, int64_t m, int64_t k,
                              const double *B, int64_t k2, int64_t n,
                              double *C);

#ifdef __cplusplus
}
#endif

#endif /* FERROLEARN_HIP_H */
