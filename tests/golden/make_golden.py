"""Writes tests/golden/golden_vectors.json with pure-Python float arithmetic.

Python floats are IEEE-754 binary64 and every `+ - * / math.sqrt` on them is
one correctly rounded op, evaluated left to right, with no FMA contraction --
the same contract as the reference's Rust f64 code.  So these loops compute
the values the reference computes, independently of the C oracle, and the
expected outputs are stored exactly (float.hex).

Sources of each case:
  fold            ferrolearn-cluster/src/kmeans.rs:159-163, kdtree.rs:342-348
  argmin/ties     kmeans.rs:295-309
  lloyd           kmeans.rs:358-401, :459-495 (serial path: n*d < 100000,
                  so the inertia is a plain sequential fold and fully pinned)
  knn             kdtree.rs:364-391 with the (distance, index) tie order
  vote / mean     knn.rs:388-440, :952-988
  minibatch       mini_batch_kmeans.rs:410-447
  rng             SplitMix64 + xoshiro256** (the library's own generator)
  stdrng_*        the reference's StdRng of rand 0.9.4: ChaCha12 keyed by
                  rand_core 0.9.5's PCG32 seed_from_u64, random_range over
                  usize (UniformUsize, Canon's method) and random::<f64>;
                  k-means++ / MiniBatch / KMeans fits drawing from it with
                  plain left folds (kmeans.rs:169-253, :451-503;
                  mini_batch_kmeans.rs:230-320, :451-472, :528-592)

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def hx(v: float) -> str:
    return float(v).hex()


def fold(a, b):
    acc = 0.0
    for x, y in zip(a, b):
        t = x - y
        acc = acc + t * t
    return acc


def nearest(row, centers):
    best, bd = 0, 1.7976931348623157e308
    for c, ctr in enumerate(centers):
        v = fold(row, ctr)
        if v < bd:
            bd, best = v, c
    return best, bd


def lloyd(X, init, max_iter, tol):
    k, d = len(init), len(init[0])
    cur = [list(r) for r in init]
    labels = [0] * len(X)
    inertia = 1.7976931348623157e308
    n_iter = 0
    for it in range(max_iter):
        inertia = 0.0
        for i, row in enumerate(X):
            labels[i], v = nearest(row, cur)
            inertia = inertia + v
        sums = [[0.0] * d for _ in range(k)]
        cnt = [0.0] * k
        for i, row in enumerate(X):
            l = labels[i]
            cnt[l] = cnt[l] + 1.0
            for j in range(d):
                sums[l][j] = sums[l][j] + row[j]
        new = []
        for c in range(k):
            if cnt[c] > 0.0:
                new.append([s / cnt[c] for s in sums[c]])
            else:
                new.append(list(cur[c]))
        mx = 0.0
        for c in range(k):
            s = fold(new[c], cur[c])
            if s > mx:
                mx = s
        shift = math.sqrt(mx)
        cur = new
        n_iter = it + 1
        if shift < tol:
            break
    return cur, labels, inertia, n_iter


def knn(data, q, k):
    ds = [(math.sqrt(fold(row, q)), i) for i, row in enumerate(data)]
    ds.sort()
    return ds[: min(k, len(data))]


def class_scores(neigh, ypos, n_classes, weights):
    sc = [0.0] * n_classes
    if weights == 0:
        for d, i in neigh:
            sc[ypos[i]] = sc[ypos[i]] + 1.0
        return sc
    zero = any(d < 1e-15 for d, _ in neigh)
    for d, i in neigh:
        if zero:
            if d < 1e-15:
                sc[ypos[i]] = sc[ypos[i]] + 1.0
        else:
            sc[ypos[i]] = sc[ypos[i]] + 1.0 / d
    return sc


def vote(sc):
    b, bs = 0, sc[0]
    for c in range(1, len(sc)):
        if sc[c] > bs:
            b, bs = c, sc[c]
    return b


def wmean(neigh, y, weights):
    if weights == 0:
        s = 0.0
        for _, i in neigh:
            s = s + y[i]
        return s / float(len(neigh))
    z = [(d, i) for d, i in neigh if d < 1e-15]
    if z:
        s = 0.0
        for _, i in z:
            s = s + y[i]
        return s / float(len(z))
    ws = wt = 0.0
    for d, i in neigh:
        w = 1.0 / d
        ws = ws + w * y[i]
        wt = wt + w
    return ws / wt


def mb_update(X, bidx, blab, C, counts):
    C = [list(r) for r in C]
    old = [list(r) for r in C]
    counts = list(counts)
    for i, l in zip(bidx, blab):
        counts[l] += 1
        lr = 1.0 / float(counts[l])
        for j in range(len(C[l])):
            C[l][j] = C[l][j] + lr * (X[i][j] - C[l][j])
    mx = 0.0
    for c in range(len(C)):
        s = math.sqrt(fold(C[c], old[c]))
        if s > mx:
            mx = s
    return C, counts, mx


def rng_stream(seed, count):
    x = seed & M64
    s = []
    for _ in range(4):
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        s.append(z ^ (z >> 31))
    rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M64
    out = []
    for _ in range(count):
        out.append(rotl((s[1] * 5) & M64, 7) * 9 & M64)
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


class Gen:
    """The library's generator: SplitMix64 seeding of xoshiro256**;
    range(n) = high 64 bits of next() * n; f64() = (next() >> 11) * 2^-53."""

    def __init__(self, seed):
        x = seed & M64
        self.s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    def next(self):
        s = self.s
        rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M64
        out = rotl((s[1] * 5) & M64, 7) * 9 & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return out

    def range_from(self, lo, hi):
        return lo + self.range(hi - lo)

    def range(self, n):
        return (self.next() * n) >> 64

    def f64(self):
        return float(self.next() >> 11) * 2.0 ** -53


BLK = 4096   # the library's two-level sum: sequential within blocks, then over blocks


def blocked_sum(v):
    tot = 0.0
    for b in range(0, len(v), BLK):
        s = 0.0
        for x in v[b:b + BLK]:
            s = s + x
        tot = tot + s
    return tot


def blocked_search(v, thr):
    run = 0.0
    for b in range(0, len(v), BLK):
        s = 0.0
        for x in v[b:b + BLK]:
            s = s + x
        if run + s >= thr:
            c = run
            for i in range(b, min(b + BLK, len(v))):
                c = c + v[i]
                if c >= thr:
                    return i
        run = run + s
    return len(v) - 1


def plain_sum(v):
    """kmeans.rs:200: the left fold from 0."""
    s = 0.0
    for x in v:
        s = s + x
    return s


def plain_search(v, thr):
    """kmeans.rs:205-214: first index whose running left fold reaches thr."""
    c = 0.0
    for i, x in enumerate(v):
        c = c + x
        if c >= thr:
            return i
    return len(v) - 1


def kmeanspp(X, k, g, total_of=None, search=None):
    """kmeans.rs:169-253 with the library's generator g and sums: the first
    row uniform; n_trials = 2 + floor(ln k); per centre the total of the
    minimum distances; a uniform row if it is <= 0, else n_trials rows drawn
    by D^2 thresholds, keeping the first with the strictly smallest
    potential.  Returns the picked row indices.  For n <= 4096 every sum is
    the plain left fold of the reference."""
    total_of = total_of or blocked_sum
    search = search or blocked_search
    n = len(X)
    picks = [g.range(n)]
    if k == 1:
        return picks
    md = [fold(x, X[picks[0]]) for x in X]
    n_trials = max(2 + int(math.floor(math.log(k))), 1)
    for _ in range(1, k):
        total = total_of(md)
        if total <= 0.0:
            picks.append(g.range(n))
            continue
        best, best_pot, best_nd = 0, None, None
        for _ in range(n_trials):
            thr = g.f64() * total
            cand = search(md, thr)
            nd = []
            for i, x in enumerate(X):
                v = fold(x, X[cand])
                nd.append(v if v < md[i] else md[i])
            pot = total_of(nd)
            if best_pot is None or pot < best_pot:
                best, best_pot, best_nd = cand, pot, nd
        picks.append(best)
        md = best_nd
    return picks


def sample_batch(g, n, batch):
    """mini_batch_kmeans.rs:451-472: a full Fisher-Yates shuffle when
    batch >= n, else the first `batch` steps of one over 0..n."""
    pool = list(range(n))
    if batch >= n:
        for i in range(n):
            j = g.range_from(i, n)
            pool[i], pool[j] = pool[j], pool[i]
        return pool
    out = []
    for i in range(batch):
        j = g.range_from(i, n)
        pool[i], pool[j] = pool[j], pool[i]
        out.append(pool[i])
    return out


def minibatch_fit(X, k, batch, max_iter, tol, n_init, base, make=None, sums=()):
    """mini_batch_kmeans.rs:478-592: run r draws from the generator seeded
    base + r*1000003 (k-means++ first, then the batches); counts start at 1;
    per iteration sample, assign (:381-404), update online (:410-447),
    n_iter += 1, stop when the shift < tol; then label every row, inertia the
    left fold in row order (:343-376); keep the strictly smallest inertia."""
    best = None
    for r in range(n_init):
        g = (make or Gen)(base + r * 1000003)
        C = [list(X[i]) for i in kmeanspp(X, k, g, *sums)]
        counts = [1] * k
        it = 0
        for _ in range(max_iter):
            bidx = sample_batch(g, len(X), batch)
            blab = [nearest(X[i], C)[0] for i in bidx]
            C, counts, sh = mb_update(X, bidx, blab, C, counts)
            it += 1
            if sh < tol:
                break
        labels, inertia = [], 0.0
        for x in X:
            l, v = nearest(x, C)
            labels.append(l)
            inertia = inertia + v
        if best is None or inertia < best[2]:
            best = (C, labels, inertia, it)
    return best


class StdRng:
    """StdRng::seed_from_u64 of rand 0.9.4, written independently of the C
    oracle.  Seed: eight PCG32 (XSH-RR) outputs, each taken after one LCG
    step, little-endian into the 32-byte ChaCha key.  Stream: ChaCha with 12
    rounds, 64-bit block counter from zero in words 12-13, zero stream id,
    each block's 16 words in order."""
    SIGMA = (0x61707865, 0x3320646E, 0x79622D32, 0x6B206574)
    COLUMNS = ((0, 4, 8, 12), (1, 5, 9, 13), (2, 6, 10, 14), (3, 7, 11, 15))
    DIAGONALS = ((0, 5, 10, 15), (1, 6, 11, 12), (2, 7, 8, 13), (3, 4, 9, 14))

    def __init__(self, seed):
        state = seed & M64
        key = b""
        for _ in range(8):
            state = (state * 6364136223846793005 + 11634580027462260723) & M64
            xorshifted = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
            rot = state >> 59
            word = ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & 0xFFFFFFFF
            key += word.to_bytes(4, "little")
        self.key = [int.from_bytes(key[4 * i:4 * i + 4], "little") for i in range(8)]
        self.block = 0
        self.words = []

    @classmethod
    def chacha(cls, inp, rounds):
        w = list(inp)
        def rotl(v, c):
            return ((v << c) | (v >> (32 - c))) & 0xFFFFFFFF
        for _ in range(rounds // 2):
            for group in (cls.COLUMNS, cls.DIAGONALS):
                for a, b, c, d in group:
                    w[a] = (w[a] + w[b]) & 0xFFFFFFFF
                    w[d] = rotl(w[d] ^ w[a], 16)
                    w[c] = (w[c] + w[d]) & 0xFFFFFFFF
                    w[b] = rotl(w[b] ^ w[c], 12)
                    w[a] = (w[a] + w[b]) & 0xFFFFFFFF
                    w[d] = rotl(w[d] ^ w[a], 8)
                    w[c] = (w[c] + w[d]) & 0xFFFFFFFF
                    w[b] = rotl(w[b] ^ w[c], 7)
        return [(x + y) & 0xFFFFFFFF for x, y in zip(w, inp)]

    def u32(self):
        if not self.words:
            inp = list(self.SIGMA) + self.key + [self.block & 0xFFFFFFFF, self.block >> 32, 0, 0]
            self.words = self.chacha(inp, 12)[::-1]
            self.block += 1
        return self.words.pop()

    def u64(self):
        lo = self.u32()
        return (self.u32() << 32) | lo

    def f64(self):
        """random::<f64>(): 53 high bits over 2^53."""
        return (self.u64() >> 11) / 9007199254740992.0

    def range_from(self, lo, hi):
        """random_range(lo..hi) over usize: UniformUsize samples u32 when
        hi - 1 < 2^32, else u64; Canon's method in that width."""
        high = hi - 1
        bits = 32 if high < (1 << 32) else 64
        mask = (1 << bits) - 1
        draw = self.u32 if bits == 32 else self.u64
        span = (high - lo + 1) & mask
        if span == 0:
            return (lo + draw()) & mask
        prod = draw() * span
        result, low = prod >> bits, prod & mask
        if low > ((1 << bits) - span):
            extra = (draw() * span) >> bits
            if low + extra > mask:
                result += 1
        return (lo + result) & mask

    def range(self, n):
        return self.range_from(0, n)


def kmeans_fit_stdrng(X, k, max_iter, tol, n_init, base):
    """kmeans.rs:451-503: run r seeds StdRng with base + r, k-means++ with
    plain folds, Lloyd; the strictly smallest inertia wins."""
    best = None
    for r in range(n_init):
        picks = kmeanspp(X, k, StdRng(base + r), plain_sum, plain_search)
        C, L, inertia, it = lloyd(X, [list(X[i]) for i in picks], max_iter, tol)
        if best is None or inertia < best[2]:
            best = (C, L, inertia, it)
    return best


def bk_sqe(a, b):
    """bisecting_kmeans.rs:174-178."""
    s = 0.0
    for x, y in zip(a, b):
        s = s + (x - y) * (x - y)
    return s


def bk_2means(X, idx, max_iter, g):
    """bisecting_kmeans.rs:225-322."""
    n = len(idx)
    i0 = g.range(n)
    i1 = g.range(n)
    if n > 1:
        while i1 == i0:
            i1 = g.range(n)
    c0, c1 = list(X[idx[i0]]), list(X[idx[i1]])
    lab = [0] * n
    for _ in range(max_iter):
        changed = False
        for li, r in enumerate(idx):
            nl = 0 if bk_sqe(X[r], c0) <= bk_sqe(X[r], c1) else 1
            if nl != lab[li]:
                lab[li] = nl
                changed = True
        if not changed:
            break
        d = len(c0)
        s0, s1, n0, n1 = [0.0] * d, [0.0] * d, 0.0, 0.0
        for li, r in enumerate(idx):
            if lab[li] == 0:
                n0 = n0 + 1.0
                s0 = [a + b for a, b in zip(s0, X[r])]
            else:
                n1 = n1 + 1.0
                s1 = [a + b for a, b in zip(s1, X[r])]
        if n0 > 0.0:
            c0 = [v / n0 for v in s0]
        if n1 > 0.0:
            c1 = [v / n1 for v in s1]
    tot = 0.0
    for li, r in enumerate(idx):
        tot = tot + bk_sqe(X[r], c0 if lab[li] == 0 else c1)
    return lab, c0, c1, tot


def bisecting_kmeans(X, n_clusters, max_iter, n_init, seed, strategy):
    """bisecting_kmeans.rs:326-505; strategy "size" (LargestCluster) or
    "sse" (LargestSSE), the last of equal candidates winning as Rust's
    max_by_key / max_by do."""
    n, d = len(X), len(X[0])
    cen = [0.0] * d
    for r in X:
        cen = [a + b for a, b in zip(cen, r)]
    cen = [v / float(n) for v in cen]
    sse = 0.0
    for r in X:
        sse = sse + bk_sqe(r, cen)
    clusters = [(list(range(n)), cen, sse)]
    split = 0
    while len(clusters) < n_clusters:
        t = None
        for i, c in enumerate(clusters):
            if len(c[0]) < 2:
                continue
            if t is None:
                t = i
            elif strategy == "size":
                if not len(clusters[t][0]) > len(c[0]):
                    t = i
            elif not clusters[t][2] > c[2]:
                t = i
        if t is None:
            break
        idx = clusters[t][0]
        best = None
        for run in range(n_init):
            g = StdRng((seed + split * 1000 + run) & M64)
            lab, c0, c1, s = bk_2means(X, idx, max_iter, g)
            if s < (best[3] if best else 1.7976931348623157e308):
                best = (lab, c0, c1, s)
        split += 1
        lab, c0, c1, _ = best
        i0 = [r for li, r in enumerate(idx) if lab[li] == 0]
        i1 = [r for li, r in enumerate(idx) if lab[li] == 1]
        s0 = 0.0
        for r in i0:
            s0 = s0 + bk_sqe(X[r], c0)
        s1 = 0.0
        for r in i1:
            s1 = s1 + bk_sqe(X[r], c1)
        del clusters[t]
        if i0:
            clusters.append((i0, c0, s0))
        if i1:
            clusters.append((i1, c1, s1))
    labels = [0] * n
    inertia = 0.0
    for ci, c in enumerate(clusters):
        for r in c[0]:
            labels[r] = ci
        inertia = inertia + c[2]
    return [c[1] for c in clusters], labels, inertia


def main():
    rnd = random.Random(20260101)
    out = {}

    # --- folds: random and adversarial ---------------------------------
    folds = []
    for d in (1, 2, 3, 5, 8, 31, 32, 64, 127, 128):
        a = [rnd.uniform(-10, 10) for _ in range(d)]
        b = [rnd.uniform(-10, 10) for _ in range(d)]
        folds.append((a, b))
    # cancellation: huge common offset
    folds.append(([1e16 + 2.0, 3.0, -1e16], [1e16, 3.5, -1e16 + 4.0]))
    # subnormal differences
    folds.append(([5e-324, 1e-310, 0.0], [0.0, -1e-310, 5e-324]))
    # order-sensitive: big term then many small ones
    folds.append(([1.0] + [2.0 ** -27] * 40, [0.0] * 41))
    folds.append(([2.0 ** -27] * 40 + [1.0], [0.0] * 41))
    # overflow to +inf
    folds.append(([1e200, 0.0], [-1e200, 0.0]))
    # exact 3-4-5 (kdtree.rs:486-491)
    folds.append(([0.0, 0.0], [3.0, 4.0]))
    out["fold"] = [
        {"a": [hx(v) for v in a], "b": [hx(v) for v in b], "sq": hx(fold(a, b)),
         "dist": hx(math.sqrt(fold(a, b)))}
        for a, b in folds
    ]

    # --- argmin with planted exact ties --------------------------------
    ties = []
    for trial in range(6):
        d = rnd.choice([2, 3, 7])
        k = rnd.choice([3, 5, 9])
        centers = [[float(rnd.randint(-3, 3)) for _ in range(d)] for _ in range(k)]
        centers[k - 1] = list(centers[0])          # duplicate centre -> exact tie
        if k > 3:
            centers[2] = list(centers[1])
        rows = [[float(rnd.randint(-4, 4)) for _ in range(d)] for _ in range(40)]
        rows += [list(centers[0]), list(centers[1])]
        labs, dists = zip(*[nearest(r, centers) for r in rows])
        ties.append({"centers": [[hx(v) for v in c] for c in centers],
                     "rows": [[hx(v) for v in r] for r in rows],
                     "labels": list(labs), "dist": [hx(v) for v in dists]})
    out["argmin"] = ties

    # --- Lloyd runs (serial path) --------------------------------------
    runs = []
    for (n, d, k, max_iter, tol, spread) in ((120, 3, 4, 50, 1e-4, 1.0),
                                              (300, 5, 6, 100, 1e-4, 2.5),
                                              (64, 2, 8, 3, 0.0, 4.0),
                                              (50, 4, 5, 0, 1e-4, 1.0),
                                              (90, 3, 3, 100, 1e-4, 0.2)):
        cen = [[rnd.uniform(-10, 10) for _ in range(d)] for _ in range(k)]
        X = [[cen[i % k][j] + rnd.gauss(0, spread) for j in range(d)] for i in range(n)]
        init = [list(X[rnd.randrange(n)]) for _ in range(k)]
        # one centre far away so a cluster can empty out
        init[-1] = [1e3] * d
        C, L, inertia, it = lloyd(X, init, max_iter, tol)
        runs.append({"X": [[hx(v) for v in r] for r in X],
                     "init": [[hx(v) for v in r] for r in init],
                     "max_iter": max_iter, "tol": tol,
                     "centers": [[hx(v) for v in r] for r in C], "labels": L,
                     "inertia": hx(inertia), "n_iter": it})
    out["lloyd"] = runs

    # --- kNN with duplicate points (exact ties) -------------------------
    knns = []
    for (N, d, nq, k) in ((40, 2, 12, 5), (60, 3, 10, 7), (25, 4, 6, 25), (30, 1, 8, 4)):
        data = [[float(rnd.randint(-3, 3)) for _ in range(d)] for _ in range(N)]
        Q = [[rnd.choice([float(rnd.randint(-3, 3)), rnd.uniform(-3, 3)]) for _ in range(d)]
             for _ in range(nq)]
        Q[0] = list(data[3])                                  # exact match
        y = [rnd.randrange(4) * 3 for _ in range(N)]          # classes {0,3,6,9}
        yr = [rnd.uniform(-50, 50) for _ in range(N)]
        classes = sorted(set(y))
        ypos = [classes.index(v) for v in y]
        res = []
        for q in Q:
            nb = knn(data, q, k)
            r = {"idx": [i for _, i in nb], "dist": [hx(dv) for dv, _ in nb]}
            for w in (0, 1):
                sc = class_scores(nb, ypos, len(classes), w)
                r[f"vote_w{w}"] = classes[vote(sc)]
                r[f"mean_w{w}"] = hx(wmean(nb, yr, w))
            res.append(r)
        knns.append({"data": [[hx(v) for v in r] for r in data],
                     "Q": [[hx(v) for v in r] for r in Q], "k": k,
                     "y": y, "yr": [hx(v) for v in yr], "results": res})
    out["knn"] = knns

    # --- MiniBatch online update ----------------------------------------
    mbs = []
    for (n, d, k, b) in ((80, 3, 4, 20), (50, 2, 3, 50)):
        X = [[rnd.uniform(-5, 5) for _ in range(d)] for _ in range(n)]
        C = [list(X[i]) for i in range(k)]
        counts = [1] * k
        bidx = [rnd.randrange(n) for _ in range(b)]
        blab = [nearest(X[i], C)[0] for i in bidx]
        C2, cnt2, sh = mb_update(X, bidx, blab, C, counts)
        mbs.append({"X": [[hx(v) for v in r] for r in X], "C": [[hx(v) for v in r] for r in C],
                    "counts": counts, "bidx": bidx, "blab": blab,
                    "C_out": [[hx(v) for v in r] for r in C2], "counts_out": cnt2,
                    "shift": hx(sh)})
    out["minibatch"] = mbs

    # --- generator ------------------------------------------------------
    out["rng"] = {str(s): [str(v) for v in rng_stream(s, 8)] for s in (0, 1, 42, 2 ** 63 + 5)}

    # --- k-means++ (library generator) ------------------------------------
    # n <= 4096: one block, so every sum is the reference's plain left fold;
    # n = 9000: three blocks, which pins the library's two-level order;
    # identical rows: the zero-potential branch; k = 1; k = n.
    rk = random.Random(20261204)
    kpp = []
    for (n, d, k, seed, spread, dup) in ((300, 3, 5, 7, 1.0, False), (1000, 2, 12, 3, 3.0, False),
                                          (50, 4, 1, 5, 1.0, False), (40, 2, 6, 9, 0.0, True),
                                          (9000, 2, 6, 11, 2.0, False), (12, 3, 12, 2, 1.0, False),
                                          (2000, 5, 17, 2 ** 40 + 3, 0.5, False)):
        cen = [[rk.uniform(-10, 10) for _ in range(d)] for _ in range(max(k, 2))]
        if dup:
            X = [[1.5] * d for _ in range(n - 3)] + [[2.5] * d, [1.5] * d, [-1.0] * d]
        else:
            X = [[cen[i % len(cen)][j] + rk.gauss(0, spread) for j in range(d)] for i in range(n)]
        picks = kmeanspp(X, k, Gen(seed))
        kpp.append({"X": [[hx(v) for v in r] for r in X], "k": k, "seed": str(seed),
                    "picks": picks})
    out["kmeanspp"] = kpp

    # --- whole MiniBatchKMeans fits (k-means++ and batches from each run's
    # generator) -------------------------------------------------------------
    mbf = []
    for (n, d, k, batch, max_iter, tol, n_init, base) in ((200, 3, 4, 50, 10, 0.0, 2, 11),
                                                          (60, 2, 3, 100, 5, 0.0, 3, 5),
                                                          (150, 2, 5, 40, 30, 0.05, 2, 8),
                                                          (90, 4, 3, 30, 6, 0.0, 1, 0)):
        cen = [[rk.uniform(-8, 8) for _ in range(d)] for _ in range(k)]
        X = [[cen[i % k][j] + rk.gauss(0, 1.5) for j in range(d)] for i in range(n)]
        C, L, inertia, it = minibatch_fit(X, k, batch, max_iter, tol, n_init, base)
        mbf.append({"X": [[hx(v) for v in r] for r in X], "k": k, "batch": batch,
                    "max_iter": max_iter, "tol": tol, "n_init": n_init, "seed": base,
                    "centers": [[hx(v) for v in r] for r in C], "labels": L,
                    "inertia": hx(inertia), "n_iter": it})
    out["minibatch_fit"] = mbf

    # --- the reference's StdRng stream ------------------------------------
    stream = {}
    for seed in (0, 1, 42, 12345678901234, M64):
        g = StdRng(seed)
        stream[str(seed)] = [g.u32() for _ in range(40)]
    out["stdrng_stream"] = stream
    spans = [(0, 10), (0, 3), (5, 9), (0, 2 ** 31 - 5), (0, 2 ** 32), (7, 2 ** 32 + 1),
             (0, 2 ** 40), (2 ** 33, 2 ** 62 + 3), (0, 1), (100, 1000)] * 3
    g = StdRng(7)
    out["stdrng_ranges"] = {"seed": 7, "lo": [str(a) for a, _ in spans],
                            "hi": [str(b) for _, b in spans],
                            "out": [str(g.range_from(a, b)) for a, b in spans]}
    g = StdRng(99)
    out["stdrng_f64"] = {"seed": 99, "out": [hx(g.f64()) for _ in range(20)]}

    rs = random.Random(20261205)
    kps = []
    for (n, d, k, seed, spread, dup) in ((300, 3, 5, 7, 1.0, False), (50, 4, 1, 5, 1.0, False),
                                          (40, 2, 6, 9, 0.0, True), (9000, 2, 6, 11, 2.0, False),
                                          (12, 3, 12, 2, 1.0, False),
                                          (1500, 5, 17, 2 ** 40 + 3, 0.5, False)):
        cen = [[rs.uniform(-10, 10) for _ in range(d)] for _ in range(max(k, 2))]
        if dup:
            X = [[1.5] * d for _ in range(n - 3)] + [[2.5] * d, [1.5] * d, [-1.0] * d]
        else:
            X = [[cen[i % len(cen)][j] + rs.gauss(0, spread) for j in range(d)] for i in range(n)]
        picks = kmeanspp(X, k, StdRng(seed), plain_sum, plain_search)
        kps.append({"X": [[hx(v) for v in r] for r in X], "k": k, "seed": str(seed),
                    "picks": picks})
    out["kmeanspp_stdrng"] = kps

    kfs = []
    for (n, d, k, max_iter, tol, n_init, base) in ((400, 3, 5, 50, 1e-4, 3, 17),
                                                   (250, 2, 4, 30, 0.0, 2, 0)):
        cen = [[rs.uniform(-8, 8) for _ in range(d)] for _ in range(k)]
        X = [[cen[i % k][j] + rs.gauss(0, 1.2) for j in range(d)] for i in range(n)]
        C, L, inertia, it = kmeans_fit_stdrng(X, k, max_iter, tol, n_init, base)
        kfs.append({"X": [[hx(v) for v in r] for r in X], "k": k, "max_iter": max_iter,
                    "tol": tol, "n_init": n_init, "seed": base,
                    "centers": [[hx(v) for v in r] for r in C], "labels": L,
                    "inertia": hx(inertia), "n_iter": it})
    out["kmeans_fit_stdrng"] = kfs

    mbs2 = []
    for (n, d, k, batch, max_iter, tol, n_init, base) in ((200, 3, 4, 50, 10, 0.0, 2, 11),
                                                          (60, 2, 3, 100, 5, 0.0, 2, 5)):
        cen = [[rs.uniform(-8, 8) for _ in range(d)] for _ in range(k)]
        X = [[cen[i % k][j] + rs.gauss(0, 1.5) for j in range(d)] for i in range(n)]
        C, L, inertia, it = minibatch_fit(X, k, batch, max_iter, tol, n_init, base, StdRng,
                                          (plain_sum, plain_search))
        mbs2.append({"X": [[hx(v) for v in r] for r in X], "k": k, "batch": batch,
                     "max_iter": max_iter, "tol": tol, "n_init": n_init, "seed": base,
                     "centers": [[hx(v) for v in r] for r in C], "labels": L,
                     "inertia": hx(inertia), "n_iter": it})
    out["minibatch_fit_stdrng"] = mbs2

    rb = random.Random(20261204)
    bks = []
    for (n, d, blobs, n_clusters, max_iter, n_init, seed, strategy) in (
            (120, 2, 4, 4, 300, 3, 42, "size"),
            (150, 3, 5, 5, 300, 2, 7, "sse"),
            (90, 2, 3, 6, 4, 1, 0, "sse"),
            (64, 4, 2, 3, 300, 10, 2 ** 64 - 3, "size")):
        cen = [[rb.uniform(-10, 10) for _ in range(d)] for _ in range(blobs)]
        X = [[cen[i % blobs][j] + rb.gauss(0, 1.0) for j in range(d)] for i in range(n)]
        C, L, inertia = bisecting_kmeans(X, n_clusters, max_iter, n_init, seed, strategy)
        bks.append({"X": [[hx(v) for v in r] for r in X], "n_clusters": n_clusters,
                    "max_iter": max_iter, "n_init": n_init, "seed": str(seed),
                    "strategy": strategy, "centers": [[hx(v) for v in r] for r in C],
                    "labels": L, "inertia": hx(inertia)})
    out["bisecting"] = bks

    with open(os.path.join(HERE, "golden_vectors.json"), "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))


if __name__ == "__main__":
    main()
