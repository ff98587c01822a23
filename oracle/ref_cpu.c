/*
 * ref_cpu.c -- CPU restatement of ferrolearn's squared-Euclidean hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check
 * and bench.py's CPU-baseline leg load it to check and to time the reference
 * algorithm.  The product (ferrolearn_amd/libferrolearn_hip.so) never links,
 * loads or calls it.
 *
 * Arithmetic contract (Rust semantics of the reference's f64 path):
 *   - every + - * / sqrt is one IEEE-754 binary64 op, round-to-nearest;
 *   - no FMA contraction, no reassociation (built with -O2 -ffp-contract=off,
 *     no -ffast-math);
 *   - folds run left to right from +0.0, exactly as Iterator::fold does.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference's behaviour is implementation-defined (rayon's reduction tree,
 * select_nth_unstable's tie order, rand's StdRng stream) this file picks a
 * definition and says so; those parts are "parity unpinned" against the real
 * reference and pinned only against this definition.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* a1: distance fold                                                         */
/* ------------------------------------------------------------------------ */

/* ferrolearn-cluster/src/kmeans.rs:159-163 (squared_euclidean) and its twin
 * mini_batch_kmeans.rs:216-220: acc = acc + (a_j - b_j) * (a_j - b_j). */
double or_sqeuclid(const double *a, const double *b, int64_t d) {
    double acc = 0.0;
    for (int64_t j = 0; j < d; ++j) {
        double t = a[j] - b[j];
        acc = acc + t * t;
    }
    return acc;
}

/* ferrolearn-neighbors/src/kdtree.rs:342-348 (euclidean_distance): the same
 * fold (map then fold from zero) followed by a correctly rounded sqrt. */
double or_euclid(const double *a, const double *b, int64_t d) {
    return sqrt(or_sqeuclid(a, b, d));
}

/* ------------------------------------------------------------------------ */
/* a2/a3: nearest centre, assignment, inertia                               */
/* ------------------------------------------------------------------------ */

/* kmeans.rs:295-309 (nearest_center): init (0, F::MAX); strict `<` over
 * centres in index order, so the lowest index wins an exact tie and a NaN
 * distance never wins. */
int64_t or_nearest_center(const double *row, const double *C, int64_t k,
                          int64_t d, double *dist_out) {
    int64_t best = 0;
    double bd = DBL_MAX;
    for (int64_t c = 0; c < k; ++c) {
        double v = or_sqeuclid(row, C + c * d, d);
        if (v < bd) {
            bd = v;
            best = c;
        }
    }
    if (dist_out) *dist_out = bd;
    return best;
}

static int or_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    return nthreads < 1 ? 1 : nthreads;
}

/* kmeans.rs:267-352 (assign_clusters_into / assign_serial / assign_parallel).
 * Serial below n*d < 100000 (PARALLEL_WORK_THRESHOLD, :261).  Otherwise rows
 * are cut into chunks of max(n/T, 64) (:333), each chunk folds its own
 * inertia sequentially, and the chunk sums are added in chunk order.  rayon
 * combines chunk sums with an unspecified tree (:351), so the last bits of
 * the reference's inertia are not reproducible; labels are. */
double or_assign(const double *X, int64_t n, int64_t d, const double *C,
                 int64_t k, int64_t *labels, int nthreads) {
    nthreads = or_threads(nthreads);
    if (n * d < 100000) {
        double inertia = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double dist;
            labels[i] = or_nearest_center(X + i * d, C, k, d, &dist);
            inertia = inertia + dist;
        }
        return inertia;
    }
    int64_t chunk = n / nthreads;
    if (chunk < 64) chunk = 64;
    int64_t nchunks = (n + chunk - 1) / chunk;
    double *part = (double *)calloc((size_t)nchunks, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        int64_t lo = ch * chunk, hi = lo + chunk < n ? lo + chunk : n;
        double local = 0.0;
        for (int64_t i = lo; i < hi; ++i) {
            double dist;
            labels[i] = or_nearest_center(X + i * d, C, k, d, &dist);
            local = local + dist;
        }
        part[ch] = local;
    }
    double inertia = 0.0;
    for (int64_t ch = 0; ch < nchunks; ++ch) inertia = inertia + part[ch];
    free(part);
    return inertia;
}

/* Per-row minimum squared distance (the value nearest_center returns), for
 * exact-sum inertia checks. */
void or_assign_rows(const double *X, int64_t n, int64_t d, const double *C,
                    int64_t k, int64_t *labels, double *dist, int nthreads) {
    nthreads = or_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i)
        labels[i] = or_nearest_center(X + i * d, C, k, d, &dist[i]);
}

/* ------------------------------------------------------------------------ */
/* a4: centroid update                                                       */
/* ------------------------------------------------------------------------ */

/* kmeans.rs:358-401 (recompute_centroids_into): zero sums and counts; for
 * rows in ascending index order count[l] += 1 and sum[l][j] += x[i][j];
 * divide by the count, or keep the old centre when the count is 0; return
 * sqrt(max_c fold(new_c - old_c)). */
double or_recompute_centroids(const double *X, int64_t n, int64_t d,
                              const int64_t *labels, const double *old_c,
                              int64_t k, double *new_c, double *counts) {
    for (int64_t i = 0; i < k * d; ++i) new_c[i] = 0.0;
    for (int64_t c = 0; c < k; ++c) counts[c] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t l = labels[i];
        counts[l] = counts[l] + 1.0;
        for (int64_t j = 0; j < d; ++j)
            new_c[l * d + j] = new_c[l * d + j] + X[i * d + j];
    }
    for (int64_t c = 0; c < k; ++c) {
        if (counts[c] > 0.0) {
            for (int64_t j = 0; j < d; ++j)
                new_c[c * d + j] = new_c[c * d + j] / counts[c];
        } else {
            for (int64_t j = 0; j < d; ++j) new_c[c * d + j] = old_c[c * d + j];
        }
    }
    double max_shift = 0.0;
    for (int64_t c = 0; c < k; ++c) {
        double s = or_sqeuclid(new_c + c * d, old_c + c * d, d);
        if (s > max_shift) max_shift = s;
    }
    return sqrt(max_shift);
}

/* ------------------------------------------------------------------------ */
/* a5: one Lloyd run and n_init restarts from given initial centres         */
/* ------------------------------------------------------------------------ */

/* kmeans.rs:459-495 (the body of one run): assign -> recompute -> swap ->
 * n_iter = iter + 1 -> break if max_shift < tol.  labels and inertia come
 * from the last assignment (made against the pre-update centres); the
 * returned centres are the post-update ones.  With max_iter == 0 the labels
 * stay 0, the inertia stays F::MAX and the centres are the initial ones. */
void or_kmeans_lloyd(const double *X, int64_t n, int64_t d, int64_t k,
                     const double *init, int64_t max_iter, double tol,
                     int nthreads, double *centers_out, int64_t *labels_out,
                     double *inertia_out, int64_t *n_iter_out) {
    double *cur = (double *)malloc((size_t)(k * d) * sizeof(double));
    double *nxt = (double *)malloc((size_t)(k * d) * sizeof(double));
    double *cnt = (double *)malloc((size_t)k * sizeof(double));
    memcpy(cur, init, (size_t)(k * d) * sizeof(double));
    for (int64_t i = 0; i < n; ++i) labels_out[i] = 0;
    double inertia = DBL_MAX;
    int64_t n_iter = 0;
    for (int64_t it = 0; it < max_iter; ++it) {
        inertia = or_assign(X, n, d, cur, k, labels_out, nthreads);
        double shift = or_recompute_centroids(X, n, d, labels_out, cur, k, nxt, cnt);
        double *t = cur;
        cur = nxt;
        nxt = t;
        n_iter = it + 1;
        if (shift < tol) break;
    }
    memcpy(centers_out, cur, (size_t)(k * d) * sizeof(double));
    *inertia_out = inertia;
    *n_iter_out = n_iter;
    free(cur);
    free(nxt);
    free(cnt);
}

/* kmeans.rs:459-503: n_init runs, each from its own initial centres
 * (init is n_init*k*d); a later run replaces the kept one only on strictly
 * smaller inertia.  Returns the index of the kept run. */
int64_t or_kmeans_fit_from(const double *X, int64_t n, int64_t d, int64_t k,
                           const double *inits, int64_t n_init,
                           int64_t max_iter, double tol, int nthreads,
                           double *centers_out, int64_t *labels_out,
                           double *inertia_out, int64_t *n_iter_out) {
    double *c = (double *)malloc((size_t)(k * d) * sizeof(double));
    int64_t *l = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t best = -1;
    for (int64_t r = 0; r < n_init; ++r) {
        double in;
        int64_t it;
        or_kmeans_lloyd(X, n, d, k, inits + r * k * d, max_iter, tol, nthreads,
                        c, l, &in, &it);
        if (best < 0 || in < *inertia_out) {
            best = r;
            *inertia_out = in;
            *n_iter_out = it;
            memcpy(centers_out, c, (size_t)(k * d) * sizeof(double));
            memcpy(labels_out, l, (size_t)n * sizeof(int64_t));
        }
    }
    free(c);
    free(l);
    return best;
}

/* kmeans.rs:555-618 (Transform): dist[i][c] = sqrt(fold(x_i, c)). */
void or_transform(const double *X, int64_t n, int64_t d, const double *C,
                  int64_t k, double *out, int nthreads) {
    nthreads = or_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t c = 0; c < k; ++c)
            out[i * k + c] = sqrt(or_sqeuclid(X + i * d, C + c * d, d));
}

/* ------------------------------------------------------------------------ */
/* a6: k-means++ with the library's own generator (parity unpinned vs rand) */
/* ------------------------------------------------------------------------ */

/* The reference seeds rand 0.9's StdRng (ChaCha12); that stream is not
 * vendored, so this restatement uses a documented generator instead:
 * SplitMix64 expands the seed into xoshiro256** state; random_f64 takes the
 * top 53 bits; random_range(n) is Lemire's multiply-shift (no rejection).
 * The library's host code uses the same definition so the two agree.  The
 * algorithms around the generator (or_kmeanspp, or_sample_batch,
 * or_minibatch_fit) are pinned by the pure-Python golden vectors of
 * tests/golden/make_golden.py; only the stream itself is unpinned against
 * the reference. */
typedef struct { uint64_t s[4]; } or_rng;

static uint64_t or_splitmix(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static void or_rng_seed(or_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = or_splitmix(&x);
}
static uint64_t or_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t or_rng_next(or_rng *r) {
    uint64_t *s = r->s;
    uint64_t out = or_rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = or_rotl(s[3], 45);
    return out;
}
static double or_rng_f64(or_rng *r) {
    return (double)(or_rng_next(r) >> 11) * 0x1.0p-53;
}
static int64_t or_rng_range(or_rng *r, int64_t n) {
    return (int64_t)(((__uint128_t)or_rng_next(r) * (uint64_t)n) >> 64);
}

/* Exported so the tests can pin the generator's first outputs. */
void or_rng_stream(uint64_t seed, int64_t count, uint64_t *out) {
    or_rng r;
    or_rng_seed(&r, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = or_rng_next(&r);
}

/* ------------------------------------------------------------------------ */
/* f1: the reference's own stream -- StdRng of rand 0.9.4 (Cargo.lock:2413),  */
/* i.e. rand_chacha 0.9.0's ChaCha12Rng seeded by rand_core 0.9.5's default   */
/* seed_from_u64, restated from the crates' published algorithms (the crates  */
/* are not vendored): PCG32 (XSH-RR) key expansion, 64-bit block counter and  */
/* zero stream id, 12 rounds plus feed-forward, u32 words in block order,     */
/* next_u64 = low word first; random::<f64>() = (next_u64 >> 11) * 2^-53;    */
/* random_range(lo..hi) over usize = UniformUsize: u32 sampling when hi - 1   */
/* fits in 32 bits, else u64, both with Canon's method.  The ChaCha block is  */
/* pinned by RFC 8439's vectors; the rest is parity unpinned against the real */
/* crates and pinned against an independent pure-Python restatement.          */
/* ------------------------------------------------------------------------ */

static uint32_t or_rotl32(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }

void or_chacha_block(const uint32_t *in, int rounds, uint32_t *out) {
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
#define OR_QR(a, b, c, d)                                   \
    x[a] += x[b]; x[d] = or_rotl32(x[d] ^ x[a], 16);        \
    x[c] += x[d]; x[b] = or_rotl32(x[b] ^ x[c], 12);        \
    x[a] += x[b]; x[d] = or_rotl32(x[d] ^ x[a], 8);         \
    x[c] += x[d]; x[b] = or_rotl32(x[b] ^ x[c], 7);
    for (int r = 0; r < rounds; r += 2) {
        OR_QR(0, 4, 8, 12) OR_QR(1, 5, 9, 13) OR_QR(2, 6, 10, 14) OR_QR(3, 7, 11, 15)
        OR_QR(0, 5, 10, 15) OR_QR(1, 6, 11, 12) OR_QR(2, 7, 8, 13) OR_QR(3, 4, 9, 14)
    }
#undef OR_QR
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

typedef struct {
    uint32_t key[8];
    uint64_t ctr;
    uint32_t buf[16];
    int idx;
} or_stdrng;

static void or_stdrng_seed(or_stdrng *g, uint64_t seed) {
    uint64_t s = seed;
    for (int i = 0; i < 8; ++i) {
        s = s * 6364136223846793005ULL + 11634580027462260723ULL;
        uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
        int rot = (int)(s >> 59);
        g->key[i] = rot ? (xs >> rot) | (xs << (32 - rot)) : xs;
    }
    g->ctr = 0;
    g->idx = 16;
}
static uint32_t or_stdrng_u32(or_stdrng *g) {
    if (g->idx == 16) {
        uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) st[4 + i] = g->key[i];
        st[12] = (uint32_t)g->ctr;
        st[13] = (uint32_t)(g->ctr >> 32);
        st[14] = st[15] = 0;
        or_chacha_block(st, 12, g->buf);
        g->ctr += 1;
        g->idx = 0;
    }
    return g->buf[g->idx++];
}
static uint64_t or_stdrng_u64(or_stdrng *g) {
    uint64_t lo = or_stdrng_u32(g);
    uint64_t hi = or_stdrng_u32(g);
    return (hi << 32) | lo;
}
static double or_stdrng_f64(or_stdrng *g) {
    return (double)(or_stdrng_u64(g) >> 11) * 0x1.0p-53;
}
static int64_t or_stdrng_range_from(or_stdrng *g, int64_t lo, int64_t hi) {
    uint64_t low = (uint64_t)lo, high = (uint64_t)hi - 1;
    if ((high >> 32) == 0) {
        uint32_t r = (uint32_t)high - (uint32_t)low + 1u;
        if (r == 0) return (int64_t)(uint32_t)((uint32_t)low + or_stdrng_u32(g));
        uint64_t m = (uint64_t)or_stdrng_u32(g) * r;
        uint32_t result = (uint32_t)(m >> 32), lo_order = (uint32_t)m;
        if (lo_order > 0u - r) {
            uint32_t nh = (uint32_t)(((uint64_t)or_stdrng_u32(g) * r) >> 32);
            if ((uint64_t)lo_order + nh > 0xffffffffULL) result += 1u;
        }
        return (int64_t)(uint32_t)((uint32_t)low + result);
    }
    uint64_t r = high - low + 1;
    if (r == 0) return (int64_t)(low + or_stdrng_u64(g));
    __uint128_t m = (__uint128_t)or_stdrng_u64(g) * r;
    uint64_t result = (uint64_t)(m >> 64), lo_order = (uint64_t)m;
    if (lo_order > 0ULL - r) {
        uint64_t nh = (uint64_t)(((__uint128_t)or_stdrng_u64(g) * r) >> 64);
        if (lo_order + nh < lo_order) result += 1;
    }
    return (int64_t)(low + result);
}

/* Exported for the tests: the first u32 words, and sequences of draws. */
void or_stdrng_stream(uint64_t seed, int64_t count, uint32_t *out) {
    or_stdrng g;
    or_stdrng_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = or_stdrng_u32(&g);
}
void or_stdrng_ranges(uint64_t seed, const int64_t *lo, const int64_t *hi, int64_t count,
                      int64_t *out) {
    or_stdrng g;
    or_stdrng_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = or_stdrng_range_from(&g, lo[i], hi[i]);
}
void or_stdrng_f64s(uint64_t seed, int64_t count, double *out) {
    or_stdrng g;
    or_stdrng_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = or_stdrng_f64(&g);
}

/* Either generator behind one interface: mode 0 is the library's stream
 * (two-level sums below), mode 1 the reference's StdRng stream with the
 * reference's plain left folds. */
typedef struct {
    int mode;
    or_rng lib;
    or_stdrng std;
} or_gen;

static void or_gen_seed(or_gen *g, int mode, uint64_t seed) {
    g->mode = mode;
    if (mode) or_stdrng_seed(&g->std, seed);
    else or_rng_seed(&g->lib, seed);
}
static double or_gen_f64(or_gen *g) {
    return g->mode ? or_stdrng_f64(&g->std) : or_rng_f64(&g->lib);
}
static int64_t or_gen_range_from(or_gen *g, int64_t lo, int64_t hi) {
    return g->mode ? or_stdrng_range_from(&g->std, lo, hi) : lo + or_rng_range(&g->lib, hi - lo);
}

/* The reference's sums over min_dists are plain left folds (kmeans.rs:200,
 * :207-213, :222-228).  This restatement uses a fixed two-level order
 * instead: blocks of OR_BLK consecutive values are folded left to right,
 * then the block sums are folded left to right.  The GPU runs the same two
 * levels, so both sides agree bit for bit. */
#define OR_BLK 4096
double or_blocked_sum(const double *v, int64_t n) {
    double tot = 0.0;
    for (int64_t b = 0; b < n; b += OR_BLK) {
        int64_t e = b + OR_BLK < n ? b + OR_BLK : n;
        double s = 0.0;
        for (int64_t i = b; i < e; ++i) s = s + v[i];
        tot = tot + s;
    }
    return tot;
}

/* First index i whose two-level running sum reaches threshold; n-1 if none
 * does (kmeans.rs:205-214 with the two-level order above). */
int64_t or_blocked_search(const double *v, int64_t n, double threshold) {
    double run = 0.0;
    for (int64_t b = 0; b < n; b += OR_BLK) {
        int64_t e = b + OR_BLK < n ? b + OR_BLK : n;
        double s = 0.0;
        for (int64_t i = b; i < e; ++i) s = s + v[i];
        if (run + s >= threshold) {
            double c = run;
            for (int64_t i = b; i < e; ++i) {
                c = c + v[i];
                if (c >= threshold) return i;
            }
        }
        run = run + s;
    }
    return n - 1;
}

/* kmeans.rs:200, :207-213, :222-228 exactly: the plain left fold, and the
 * first index whose running left fold reaches threshold (n-1 if none). */
double or_seq_sum(const double *v, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = s + v[i];
    return s;
}
int64_t or_seq_search(const double *v, int64_t n, double threshold) {
    double c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        c = c + v[i];
        if (c >= threshold) return i;
    }
    return n - 1;
}
static double or_gen_sum(const or_gen *g, const double *v, int64_t n) {
    return g->mode ? or_seq_sum(v, n) : or_blocked_sum(v, n);
}
static int64_t or_gen_search(const or_gen *g, const double *v, int64_t n, double thr) {
    return g->mode ? or_seq_search(v, n, thr) : or_blocked_search(v, n, thr);
}

/* kmeans.rs:169-253 (kmeans_plus_plus), greedy multi-trial k-means++:
 * first centre uniform; n_trials = max(2 + floor(ln k), 1); for each new
 * centre, if the total potential is <= 0 pick uniformly, else draw n_trials
 * candidates with probability proportional to D(x)^2 and keep the one with
 * the strictly smallest resulting potential (first wins ties). */
static void or_kmeanspp_gen(const double *X, int64_t n, int64_t d, int64_t k,
                            or_gen *rngp, double *centers, int nthreads);

void or_kmeanspp(const double *X, int64_t n, int64_t d, int64_t k,
                 uint64_t seed, double *centers, int nthreads) {
    or_gen rng;
    or_gen_seed(&rng, 0, seed);
    or_kmeanspp_gen(X, n, d, k, &rng, centers, nthreads);
}

/* The same with the reference's StdRng stream and plain left folds
 * (kmeans.rs:169-253 with StdRng::seed_from_u64(seed), :460). */
void or_kmeanspp_stdrng(const double *X, int64_t n, int64_t d, int64_t k,
                        uint64_t seed, double *centers, int nthreads) {
    or_gen rng;
    or_gen_seed(&rng, 1, seed);
    or_kmeanspp_gen(X, n, d, k, &rng, centers, nthreads);
}

/* The same, drawing from a generator the caller keeps using afterwards
 * (MiniBatchKMeans takes its batches from the run's generator,
 * mini_batch_kmeans.rs:528-555). */
static void or_kmeanspp_gen(const double *X, int64_t n, int64_t d, int64_t k,
                            or_gen *rngp, double *centers, int nthreads) {
    nthreads = or_threads(nthreads);
    or_gen rng = *rngp;
    memset(centers, 0, (size_t)(k * d) * sizeof(double));
    if (n == 0) return;
    int64_t first = or_gen_range_from(&rng, 0, n);
    memcpy(centers, X + first * d, (size_t)d * sizeof(double));
    if (k == 1) {
        *rngp = rng;
        return;
    }
    double *md = (double *)malloc((size_t)n * sizeof(double));
    double *nd = (double *)malloc((size_t)n * sizeof(double));
    double *bd = (double *)malloc((size_t)n * sizeof(double));
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) md[i] = or_sqeuclid(X + i * d, centers, d);
    int64_t n_trials = 2 + (int64_t)floor(log((double)k));
    if (n_trials < 1) n_trials = 1;
    for (int64_t c = 1; c < k; ++c) {
        double total = or_gen_sum(&rng, md, n);
        if (total <= 0.0) {
            int64_t idx = or_gen_range_from(&rng, 0, n);
            memcpy(centers + c * d, X + idx * d, (size_t)d * sizeof(double));
            continue;
        }
        int64_t best = 0;
        double best_pot = 0.0;
        int have = 0;
        for (int64_t t = 0; t < n_trials; ++t) {
            double thr = or_gen_f64(&rng) * total;
            int64_t cand = or_gen_search(&rng, md, n, thr);
            const double *cr = X + cand * d;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t i = 0; i < n; ++i) {
                double v = or_sqeuclid(X + i * d, cr, d);
                nd[i] = v < md[i] ? v : md[i];
            }
            double pot = or_gen_sum(&rng, nd, n);
            if (!have || pot < best_pot) {
                have = 1;
                best_pot = pot;
                best = cand;
                double *tmp = bd;
                bd = nd;
                nd = tmp;
            }
        }
        memcpy(centers + c * d, X + best * d, (size_t)d * sizeof(double));
        double *tmp = md;
        md = bd;
        bd = tmp;
    }
    free(md);
    free(nd);
    free(bd);
    *rngp = rng;
}

/* ------------------------------------------------------------------------ */
/* a8: MiniBatchKMeans pieces                                               */
/* ------------------------------------------------------------------------ */

/* mini_batch_kmeans.rs:410-447 (update_centers_mini_batch): for each batch
 * sample in order, count[l] += 1, lr = 1/count, c_l[j] += lr * (x[j] - c_l[j]);
 * returns max_c sqrt(fold(new_c - old_c)). */
double or_minibatch_update(const double *X, int64_t d, const int64_t *batch_idx,
                           const int64_t *batch_lab, int64_t b, double *C,
                           int64_t k, int64_t *counts) {
    double *old = (double *)malloc((size_t)(k * d) * sizeof(double));
    memcpy(old, C, (size_t)(k * d) * sizeof(double));
    for (int64_t s = 0; s < b; ++s) {
        int64_t l = batch_lab[s];
        const double *x = X + batch_idx[s] * d;
        counts[l] += 1;
        double lr = 1.0 / (double)counts[l];
        for (int64_t j = 0; j < d; ++j)
            C[l * d + j] = C[l * d + j] + lr * (x[j] - C[l * d + j]);
    }
    double mx = 0.0;
    for (int64_t c = 0; c < k; ++c) {
        double s = sqrt(or_sqeuclid(C + c * d, old + c * d, d));
        if (s > mx) mx = s;
    }
    free(old);
    return mx;
}

/* mini_batch_kmeans.rs:451-472 (sample_batch_indices) with the generator
 * above: a full shuffle when batch >= n, else a partial Fisher-Yates. */
static void or_sample_batch(int64_t n, int64_t batch, or_gen *rng, int64_t *pool,
                            int64_t *out, int64_t *out_len) {
    for (int64_t i = 0; i < n; ++i) pool[i] = i;
    if (batch >= n) {
        for (int64_t i = 0; i < n; ++i) {
            int64_t j = or_gen_range_from(rng, i, n);
            int64_t t = pool[i];
            pool[i] = pool[j];
            pool[j] = t;
        }
        memcpy(out, pool, (size_t)n * sizeof(int64_t));
        *out_len = n;
        return;
    }
    for (int64_t i = 0; i < batch; ++i) {
        int64_t j = or_gen_range_from(rng, i, n);
        int64_t t = pool[i];
        pool[i] = pool[j];
        pool[j] = t;
        out[i] = pool[i];
    }
    *out_len = batch;
}

/* mini_batch_kmeans.rs:478-592 for one run from given initial centres:
 * counts start at 1; per iteration sample a batch, assign it (serial,
 * :381-404), update online, n_iter += 1, stop if shift < tol; then assign
 * every row (parallel per row, sequential inertia fold, :343-376). */
void or_minibatch_run(const double *X, int64_t n, int64_t d, int64_t k,
                      const double *init, int64_t batch, int64_t max_iter,
                      double tol, uint64_t seed, int nthreads,
                      double *centers_out, int64_t *labels_out,
                      double *inertia_out, int64_t *n_iter_out) {
    or_gen rng;
    or_gen_seed(&rng, 0, seed);
    memcpy(centers_out, init, (size_t)(k * d) * sizeof(double));
    int64_t *counts = (int64_t *)malloc((size_t)k * sizeof(int64_t));
    for (int64_t c = 0; c < k; ++c) counts[c] = 1;
    int64_t *pool = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t bmax = batch < n ? batch : n;
    int64_t *bidx = (int64_t *)malloc((size_t)bmax * sizeof(int64_t));
    int64_t *blab = (int64_t *)malloc((size_t)bmax * sizeof(int64_t));
    int64_t it = 0;
    for (int64_t r = 0; r < max_iter; ++r) {
        int64_t blen;
        or_sample_batch(n, batch, &rng, pool, bidx, &blen);
        for (int64_t s = 0; s < blen; ++s)
            blab[s] = or_nearest_center(X + bidx[s] * d, centers_out, k, d, NULL);
        double shift = or_minibatch_update(X, d, bidx, blab, blen, centers_out, k, counts);
        it += 1;
        if (shift < tol) break;
    }
    double *dist = (double *)malloc((size_t)n * sizeof(double));
    or_assign_rows(X, n, d, centers_out, k, labels_out, dist, nthreads);
    double in = 0.0;
    for (int64_t i = 0; i < n; ++i) in = in + dist[i];
    *inertia_out = in;
    *n_iter_out = it;
    free(dist);
    free(counts);
    free(pool);
    free(bidx);
    free(blab);
}

/* mini_batch_kmeans.rs:478-592, the whole fit without given centres: run r
 * seeds the generator with seed + r*1000003, draws k-means++ from it and
 * then its batches; the run with the strictly smallest inertia is kept.
 * mode: 0 the library's generator, 1 the reference's StdRng stream. */
static void or_minibatch_fit_gen(int mode, const double *X, int64_t n, int64_t d, int64_t k,
                                 int64_t batch, int64_t max_iter, double tol, int64_t n_init,
                                 uint64_t seed, int nthreads, double *centers_out,
                                 int64_t *labels_out, double *inertia_out,
                                 int64_t *n_iter_out) {
    double *C = (double *)malloc((size_t)(k * d) * sizeof(double));
    int64_t *L = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *counts = (int64_t *)malloc((size_t)k * sizeof(int64_t));
    int64_t *pool = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t bmax = batch < n ? batch : n;
    int64_t *bidx = (int64_t *)malloc((size_t)bmax * sizeof(int64_t));
    int64_t *blab = (int64_t *)malloc((size_t)bmax * sizeof(int64_t));
    double *dist = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t r = 0; r < n_init; ++r) {
        or_gen rng;
        or_gen_seed(&rng, mode, seed + (uint64_t)r * 1000003ULL);
        or_kmeanspp_gen(X, n, d, k, &rng, C, nthreads);
        for (int64_t c = 0; c < k; ++c) counts[c] = 1;
        int64_t it = 0;
        for (int64_t q = 0; q < max_iter; ++q) {
            int64_t blen;
            or_sample_batch(n, batch, &rng, pool, bidx, &blen);
            for (int64_t s = 0; s < blen; ++s)
                blab[s] = or_nearest_center(X + bidx[s] * d, C, k, d, NULL);
            double shift = or_minibatch_update(X, d, bidx, blab, blen, C, k, counts);
            it += 1;
            if (shift < tol) break;
        }
        or_assign_rows(X, n, d, C, k, L, dist, nthreads);
        double in = 0.0;
        for (int64_t i = 0; i < n; ++i) in = in + dist[i];
        if (r == 0 || in < *inertia_out) {
            *inertia_out = in;
            *n_iter_out = it;
            memcpy(centers_out, C, (size_t)(k * d) * sizeof(double));
            memcpy(labels_out, L, (size_t)n * sizeof(int64_t));
        }
    }
    free(C);
    free(L);
    free(counts);
    free(pool);
    free(bidx);
    free(blab);
    free(dist);
}

void or_minibatch_fit(const double *X, int64_t n, int64_t d, int64_t k, int64_t batch,
                      int64_t max_iter, double tol, int64_t n_init, uint64_t seed,
                      int nthreads, double *centers_out, int64_t *labels_out,
                      double *inertia_out, int64_t *n_iter_out) {
    or_minibatch_fit_gen(0, X, n, d, k, batch, max_iter, tol, n_init, seed, nthreads,
                         centers_out, labels_out, inertia_out, n_iter_out);
}

void or_minibatch_fit_stdrng(const double *X, int64_t n, int64_t d, int64_t k, int64_t batch,
                             int64_t max_iter, double tol, int64_t n_init, uint64_t seed,
                             int nthreads, double *centers_out, int64_t *labels_out,
                             double *inertia_out, int64_t *n_iter_out) {
    or_minibatch_fit_gen(1, X, n, d, k, batch, max_iter, tol, n_init, seed, nthreads,
                         centers_out, labels_out, inertia_out, n_iter_out);
}

/* kmeans.rs:403-513 with StdRng: run r draws k-means++ from
 * StdRng::seed_from_u64(seed + r) (:460), then runs Lloyd; the run with the
 * strictly smallest inertia is kept. */
void or_kmeans_fit_stdrng(const double *X, int64_t n, int64_t d, int64_t k, int64_t max_iter,
                          double tol, int64_t n_init, uint64_t seed, int nthreads,
                          double *centers_out, int64_t *labels_out, double *inertia_out,
                          int64_t *n_iter_out) {
    double *init = (double *)malloc((size_t)(k * d) * sizeof(double));
    double *C = (double *)malloc((size_t)(k * d) * sizeof(double));
    int64_t *L = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t r = 0; r < n_init; ++r) {
        double in;
        int64_t it;
        or_kmeanspp_stdrng(X, n, d, k, seed + (uint64_t)r, init, nthreads);
        or_kmeans_lloyd(X, n, d, k, init, max_iter, tol, nthreads, C, L, &in, &it);
        if (r == 0 || in < *inertia_out) {
            *inertia_out = in;
            *n_iter_out = it;
            memcpy(centers_out, C, (size_t)(k * d) * sizeof(double));
            memcpy(labels_out, L, (size_t)n * sizeof(int64_t));
        }
    }
    free(init);
    free(C);
    free(L);
}

/* ------------------------------------------------------------------------ */
/* a9/a10/a15: brute-force kNN                                               */
/* ------------------------------------------------------------------------ */

/* kdtree.rs:364-391 (brute_force_knn): every distance sqrt(fold), then the k
 * smallest sorted ascending.  The reference breaks distance ties through
 * select_nth_unstable_by + a stable sort, an implementation-defined order;
 * this restatement defines ties by ascending index, i.e. it sorts by the
 * pair (distance, index).  k is clamped to N (:384). */
static int or_pair_less(double da, int64_t ia, double db, int64_t ib) {
    return da < db || (da == db && ia < ib);
}

int64_t or_knn_query(const double *data, int64_t N, int64_t d,
                     const double *q, int64_t k, int64_t *idx_out,
                     double *dist_out) {
    if (k > N) k = N;
    if (k <= 0) return 0;
    int64_t cnt = 0;
    for (int64_t i = 0; i < N; ++i) {
        double v = sqrt(or_sqeuclid(data + i * d, q, d));
        if (cnt == k && !or_pair_less(v, i, dist_out[k - 1], idx_out[k - 1]))
            continue;
        int64_t p = cnt < k ? cnt++ : k - 1;
        while (p > 0 && or_pair_less(v, i, dist_out[p - 1], idx_out[p - 1])) {
            dist_out[p] = dist_out[p - 1];
            idx_out[p] = idx_out[p - 1];
            --p;
        }
        dist_out[p] = v;
        idx_out[p] = i;
    }
    return k;
}

/* knn.rs:875-917 (kneighbors_impl), parallel over queries here. */
void or_kneighbors(const double *data, int64_t N, int64_t d, const double *Q,
                   int64_t nq, int64_t k, int64_t *idx, double *dist,
                   int nthreads) {
    nthreads = or_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i)
        or_knn_query(data, N, d, Q + i * d, k, idx + i * k, dist + i * k);
}

/* ------------------------------------------------------------------------ */
/* a12-a14: vote, probabilities, regression mean                             */
/* ------------------------------------------------------------------------ */

/* knn.rs:388-423 (class_score_vec).  ypos[i] is the position of y_train[i]
 * in the sorted unique class list.  Uniform: +1 per neighbour.  Distance: if
 * any neighbour distance is < 1e-15 only those count, +1 each; otherwise
 * +1/d, accumulated in neighbour order. */
void or_class_scores(const int64_t *nidx, const double *ndist, int64_t k,
                     const int64_t *ypos, int64_t n_classes, int weights,
                     double *scores) {
    const double eps = 1e-15;
    for (int64_t c = 0; c < n_classes; ++c) scores[c] = 0.0;
    if (weights == 0) {
        for (int64_t m = 0; m < k; ++m) {
            int64_t ci = ypos[nidx[m]];
            scores[ci] = scores[ci] + 1.0;
        }
        return;
    }
    int zero = 0;
    for (int64_t m = 0; m < k; ++m)
        if (ndist[m] < eps) zero = 1;
    for (int64_t m = 0; m < k; ++m) {
        int64_t ci = ypos[nidx[m]];
        if (zero) {
            if (ndist[m] < eps) scores[ci] = scores[ci] + 1.0;
        } else {
            scores[ci] = scores[ci] + 1.0 / ndist[m];
        }
    }
}

/* knn.rs:427-440 (weighted_vote): argmax with strict `>` over the sorted
 * classes, so the smallest class position wins a tie. */
int64_t or_vote(const double *scores, int64_t n_classes) {
    int64_t b = 0;
    double bs = scores[0];
    for (int64_t c = 1; c < n_classes; ++c)
        if (scores[c] > bs) {
            bs = scores[c];
            b = c;
        }
    return b;
}

/* knn.rs:452-486 (predict_proba): scores / fold(scores); uniform 1/C when
 * the total is not positive. */
void or_proba(const double *scores, int64_t n_classes, double *out) {
    double tot = 0.0;
    for (int64_t c = 0; c < n_classes; ++c) tot = tot + scores[c];
    if (tot > 0.0) {
        for (int64_t c = 0; c < n_classes; ++c) out[c] = scores[c] / tot;
    } else {
        double u = 1.0 / (double)n_classes;
        for (int64_t c = 0; c < n_classes; ++c) out[c] = u;
    }
}

/* knn.rs:952-988 (weighted_mean), neighbour order throughout. */
double or_weighted_mean(const int64_t *nidx, const double *ndist, int64_t k,
                        const double *y, int weights) {
    const double eps = 1e-15;
    if (weights == 0) {
        double s = 0.0;
        for (int64_t m = 0; m < k; ++m) s = s + y[nidx[m]];
        return s / (double)k;
    }
    int64_t nz = 0;
    for (int64_t m = 0; m < k; ++m)
        if (ndist[m] < eps) nz++;
    if (nz > 0) {
        double s = 0.0;
        for (int64_t m = 0; m < k; ++m)
            if (ndist[m] < eps) s = s + y[nidx[m]];
        return s / (double)nz;
    }
    double ws = 0.0, wt = 0.0;
    for (int64_t m = 0; m < k; ++m) {
        double w = 1.0 / ndist[m];
        ws = ws + w * y[nidx[m]];
        wt = wt + w;
    }
    return ws / wt;
}

/* knn.rs:322-383 (classifier predict) with BruteForce (knn.rs:148):
 * neighbours, class scores, vote; returns class positions. */
void or_knn_classify(const double *data, int64_t N, int64_t d,
                     const int64_t *ypos, int64_t n_classes, const double *Q,
                     int64_t nq, int64_t k, int weights, int64_t *pred_pos,
                     double *proba_or_null, int nthreads) {
    nthreads = or_threads(nthreads);
    int64_t kk = k < N ? k : N;
#pragma omp parallel num_threads(nthreads)
    {
        int64_t *ni = (int64_t *)malloc((size_t)kk * sizeof(int64_t));
        double *nd = (double *)malloc((size_t)kk * sizeof(double));
        double *sc = (double *)malloc((size_t)n_classes * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < nq; ++i) {
            or_knn_query(data, N, d, Q + i * d, k, ni, nd);
            or_class_scores(ni, nd, kk, ypos, n_classes, weights, sc);
            pred_pos[i] = or_vote(sc, n_classes);
            if (proba_or_null) or_proba(sc, n_classes, proba_or_null + i * n_classes);
        }
        free(ni);
        free(nd);
        free(sc);
    }
}

/* knn.rs:720-781 (regressor predict) with BruteForce. */
void or_knn_regress(const double *data, int64_t N, int64_t d, const double *y,
                    const double *Q, int64_t nq, int64_t k, int weights,
                    double *pred, int nthreads) {
    nthreads = or_threads(nthreads);
    int64_t kk = k < N ? k : N;
#pragma omp parallel num_threads(nthreads)
    {
        int64_t *ni = (int64_t *)malloc((size_t)kk * sizeof(int64_t));
        double *nd = (double *)malloc((size_t)kk * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < nq; ++i) {
            or_knn_query(data, N, d, Q + i * d, k, ni, nd);
            pred[i] = or_weighted_mean(ni, nd, kk, y, weights);
        }
        free(ni);
        free(nd);
    }
}

/* Thread count the OpenMP runtime would use by default (for reports). */
int or_max_threads(void) { return or_threads(0); }

/* ------------------------------------------------------------------------ */
/* f4: pairwise_distances_argmin(_min), Metric::Euclidean                     */
/* ------------------------------------------------------------------------ */

/* ferrolearn-metrics/src/pairwise.rs:157-185 (euclidean_distances) and
 * :564-629: x_sq / y_sq are left folds of v*v, dot the left fold of x*y,
 * the distance sqrt(max((x_sq + y_sq) - 2*dot, 0)) where max returns the
 * non-NaN operand (f64::max, as C's fmax), and the index the first j with
 * the smallest distance (best starts at j = 0, then strict <).  min_out may
 * be NULL. */
/* ferrolearn-metrics/src/pairwise.rs:157-189 (euclidean_distances): every
 * entry out[i*m + j] = sqrt(max((x_sq + y_sq) - 2*dot, 0)) with the left
 * folds above, max returning the non-NaN operand. */
void or_euclidean_distances(const double *X, int64_t n, const double *Y, int64_t m, int64_t d,
                            double *out, int nthreads) {
    nthreads = or_threads(nthreads);
    double *ysq = (double *)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    for (int64_t j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t k = 0; k < d; ++k) s = s + Y[j * d + k] * Y[j * d + k];
        ysq[j] = s;
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        const double *x = X + i * d;
        double xsq = 0.0;
        for (int64_t k = 0; k < d; ++k) xsq = xsq + x[k] * x[k];
        for (int64_t j = 0; j < m; ++j) {
            double dot = 0.0;
            for (int64_t k = 0; k < d; ++k) dot = dot + x[k] * Y[j * d + k];
            out[i * m + j] = sqrt(fmax((xsq + ysq[j]) - 2.0 * dot, 0.0));
        }
    }
    free(ysq);
}

void or_pairwise_argmin_min(const double *X, int64_t n, const double *Y, int64_t m, int64_t d,
                            int64_t *idx_out, double *min_out, int nthreads) {
    nthreads = or_threads(nthreads);
    double *ysq = (double *)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    for (int64_t j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t k = 0; k < d; ++k) s = s + Y[j * d + k] * Y[j * d + k];
        ysq[j] = s;
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        const double *x = X + i * d;
        double xsq = 0.0;
        for (int64_t k = 0; k < d; ++k) xsq = xsq + x[k] * x[k];
        int64_t bi = 0;
        double bv = 0.0;
        for (int64_t j = 0; j < m; ++j) {
            double dot = 0.0;
            for (int64_t k = 0; k < d; ++k) dot = dot + x[k] * Y[j * d + k];
            const double v = sqrt(fmax((xsq + ysq[j]) - 2.0 * dot, 0.0));
            if (j == 0 || v < bv) {
                bv = v;
                bi = j;
            }
        }
        idx_out[i] = bi;
        if (min_out) min_out[i] = bv;
    }
    free(ysq);
}

/* ------------------------------------------------------------------------ */
/* BisectingKMeans -- ferrolearn-cluster/src/bisecting_kmeans.rs            */
/* ------------------------------------------------------------------------ */

/* :174-178: acc + (a - b) * (a - b), left to right. */
static double or_bk_sqe(const double *a, const double *b, int64_t d) {
    double s = 0.0;
    for (int64_t k = 0; k < d; ++k) {
        const double t = a[k] - b[k];
        s = s + t * t;
    }
    return s;
}

/* :225-322 (run_2means) on the rows idx[0..n): two distinct random rows as
 * centres (random_range(0..n), redrawn while equal when n > 1), then until
 * no label changes or max_iter: d0 <= d1 -> 0 else 1; centres = sums in
 * member order / counts (a centre with no member keeps its value); the
 * total SSE is the left fold over the members in order. */
static double or_bk_2means(const double *X, int64_t d, const int64_t *idx, int64_t n,
                           int64_t max_iter, or_stdrng *g, uint8_t *lab, double *c0,
                           double *c1) {
    const int64_t i0 = or_stdrng_range_from(g, 0, n);
    int64_t i1 = or_stdrng_range_from(g, 0, n);
    if (n > 1)
        while (i1 == i0) i1 = or_stdrng_range_from(g, 0, n);
    memcpy(c0, X + idx[i0] * d, (size_t)d * sizeof(double));
    memcpy(c1, X + idx[i1] * d, (size_t)d * sizeof(double));
    memset(lab, 0, (size_t)n);
    double *s0 = (double *)malloc((size_t)d * sizeof(double));
    double *s1 = (double *)malloc((size_t)d * sizeof(double));
    for (int64_t it = 0; it < max_iter; ++it) {
        int changed = 0;
        for (int64_t li = 0; li < n; ++li) {
            const double *x = X + idx[li] * d;
            const double d0 = or_bk_sqe(x, c0, d), d1 = or_bk_sqe(x, c1, d);
            const uint8_t nl = d0 <= d1 ? 0 : 1;
            if (nl != lab[li]) {
                lab[li] = nl;
                changed = 1;
            }
        }
        if (!changed) break;
        memset(s0, 0, (size_t)d * sizeof(double));
        memset(s1, 0, (size_t)d * sizeof(double));
        double n0 = 0.0, n1 = 0.0;
        for (int64_t li = 0; li < n; ++li) {
            const double *x = X + idx[li] * d;
            if (lab[li] == 0) {
                n0 = n0 + 1.0;
                for (int64_t j = 0; j < d; ++j) s0[j] = s0[j] + x[j];
            } else {
                n1 = n1 + 1.0;
                for (int64_t j = 0; j < d; ++j) s1[j] = s1[j] + x[j];
            }
        }
        if (n0 > 0.0)
            for (int64_t j = 0; j < d; ++j) c0[j] = s0[j] / n0;
        if (n1 > 0.0)
            for (int64_t j = 0; j < d; ++j) c1[j] = s1[j] / n1;
    }
    free(s0);
    free(s1);
    double tot = 0.0;
    for (int64_t li = 0; li < n; ++li)
        tot = tot + or_bk_sqe(X + idx[li] * d, lab[li] == 0 ? c0 : c1, d);
    return tot;
}

typedef struct {
    int64_t *idx;
    int64_t n;
    double *center;
    double sse;
} or_bk_cluster;

/* :326-505 (fit).  strategy 0 = LargestCluster (max_by_key: the last of
 * equal sizes), 1 = LargestSSE (max_by with partial_cmp, unordered as
 * equal: the last of equal or unordered).  Returns the number of clusters,
 * or -1 where the reference panics (no run with SSE below f64::MAX), or -2
 * after max_stall bisections in a row that left one side empty (the
 * reference would keep drawing). */
int64_t or_bisecting_kmeans(const double *X, int64_t n, int64_t d, int64_t n_clusters,
                            int64_t max_iter, int64_t n_init, uint64_t seed, int strategy,
                            int64_t max_stall, double *centers_out, int64_t *labels_out,
                            double *inertia_out) {
    or_bk_cluster *cl = (or_bk_cluster *)calloc((size_t)n_clusters + 1, sizeof(or_bk_cluster));
    int64_t nc = 1;
    cl[0].idx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) cl[0].idx[i] = i;
    cl[0].n = n;
    cl[0].center = (double *)calloc((size_t)d, sizeof(double));
    /* :192-223 compute_cluster_stats */
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < d; ++j) cl[0].center[j] = cl[0].center[j] + X[i * d + j];
    for (int64_t j = 0; j < d; ++j) cl[0].center[j] = cl[0].center[j] / (double)n;
    double sse = 0.0;
    for (int64_t i = 0; i < n; ++i) sse = sse + or_bk_sqe(X + i * d, cl[0].center, d);
    cl[0].sse = sse;

    uint64_t split_count = 0;
    int64_t stall = 0, ret = 0;
    uint8_t *lab = (uint8_t *)malloc((size_t)n + 1);
    uint8_t *blab = (uint8_t *)malloc((size_t)n + 1);
    double *c0 = (double *)malloc((size_t)d * sizeof(double));
    double *c1 = (double *)malloc((size_t)d * sizeof(double));
    while (nc < n_clusters) {
        int64_t t = -1;
        for (int64_t i = 0; i < nc; ++i) {
            if (cl[i].n < 2) continue;
            if (t < 0) {
                t = i;
            } else if (strategy == 0) {
                if (!(cl[t].n > cl[i].n)) t = i;
            } else {
                if (!(cl[t].sse > cl[i].sse)) t = i;
            }
        }
        if (t < 0) break;
        const int64_t tn = cl[t].n;
        double best = 1.7976931348623157e308;
        int have = 0;
        double *b0 = (double *)malloc((size_t)d * sizeof(double));
        double *b1 = (double *)malloc((size_t)d * sizeof(double));
        for (int64_t run = 0; run < n_init; ++run) {
            or_stdrng g;
            or_stdrng_seed(&g, seed + split_count * 1000u + (uint64_t)run);
            const double s = or_bk_2means(X, d, cl[t].idx, tn, max_iter, &g, lab, c0, c1);
            if (s < best) {
                best = s;
                have = 1;
                memcpy(blab, lab, (size_t)tn);
                memcpy(b0, c0, (size_t)d * sizeof(double));
                memcpy(b1, c1, (size_t)d * sizeof(double));
            }
        }
        ++split_count;
        if (!have) {
            free(b0);
            free(b1);
            ret = -1;
            break;
        }
        int64_t n0 = 0;
        for (int64_t li = 0; li < tn; ++li) n0 += blab[li] == 0;
        const int64_t n1 = tn - n0;
        or_bk_cluster k0 = {NULL, n0, b0, 0.0}, k1 = {NULL, n1, b1, 0.0};
        k0.idx = (int64_t *)malloc((size_t)(n0 > 0 ? n0 : 1) * sizeof(int64_t));
        k1.idx = (int64_t *)malloc((size_t)(n1 > 0 ? n1 : 1) * sizeof(int64_t));
        int64_t p0 = 0, p1 = 0;
        for (int64_t li = 0; li < tn; ++li) {
            if (blab[li] == 0)
                k0.idx[p0++] = cl[t].idx[li];
            else
                k1.idx[p1++] = cl[t].idx[li];
        }
        for (int64_t i = 0; i < n0; ++i) k0.sse = k0.sse + or_bk_sqe(X + k0.idx[i] * d, b0, d);
        for (int64_t i = 0; i < n1; ++i) k1.sse = k1.sse + or_bk_sqe(X + k1.idx[i] * d, b1, d);
        free(cl[t].idx);
        free(cl[t].center);
        for (int64_t i = t; i + 1 < nc; ++i) cl[i] = cl[i + 1];
        --nc;
        if (n0 > 0) {
            cl[nc++] = k0;
        } else {
            free(k0.idx);
            free(k0.center);
        }
        if (n1 > 0) {
            cl[nc++] = k1;
        } else {
            free(k1.idx);
            free(k1.center);
        }
        if (n0 == 0 || n1 == 0) {
            if (++stall >= max_stall) {
                ret = -2;
                break;
            }
        } else {
            stall = 0;
        }
    }
    if (ret == 0) {
        double inertia = 0.0;
        for (int64_t c = 0; c < nc; ++c) {
            memcpy(centers_out + c * d, cl[c].center, (size_t)d * sizeof(double));
            for (int64_t i = 0; i < cl[c].n; ++i) labels_out[cl[c].idx[i]] = c;
            inertia = inertia + cl[c].sse;
        }
        *inertia_out = inertia;
        ret = nc;
    }
    for (int64_t c = 0; c < nc; ++c) {
        free(cl[c].idx);
        free(cl[c].center);
    }
    free(cl);
    free(lab);
    free(blab);
    free(c0);
    free(c1);
    return ret;
}
