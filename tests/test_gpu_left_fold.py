"""The workgroup left fold (fl_left_fold_device) against Python's own
serial loop of IEEE binary64 adds, bit for bit: the sums and every running
sum, on data built to hit each branch of the method -- long binade phases,
exact round-half-to-even ties, subnormal sums, crossings by huge addends, a
sum that reaches 2^53 ulps, infinities, NaN, negative addends and signed
zeros -- at sizes inside one window and across many."""
import ctypes as C
import math

import numpy as np
import pytest

import ferrolearn_amd as F
from ferrolearn_amd import _lib as L
from ferrolearn_amd.errors import check

pytestmark = pytest.mark.gpu


def serial(v):
    s = 0.0
    out = np.empty(len(v))
    for i, x in enumerate(v.tolist()):
        s = s + x
        out[i] = s
    return out


def device_fold(seqs):
    n = seqs.shape[1]
    count = seqs.shape[0]
    dv = F.to_device(np.ascontiguousarray(seqs))
    sums = F.device.empty((count,), np.float64)
    pre = F.device.empty((count, n), np.float64)
    check(L.lib().fl_left_fold_device(C.c_void_p(dv.data_ptr()), n, count, n,
                                      C.c_void_p(sums.data_ptr()), C.c_void_p(pre.data_ptr())))
    return sums.numpy(), pre.numpy()


def same_bits(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(((a.view(np.int64) == b.view(np.int64)) | both_nan).all())


def cases(n, rng):
    out = {}
    out["squares"] = rng.random(n) ** 2 * 100.0
    out["distances"] = rng.gamma(2.0, 40.0, n) ** 2
    ties = rng.choice([0.0, 1.0, 2.0, 3.0, 5.0, 7.0, 0.5, 1.5], n)
    ties[0] = 2.0 ** 53 + 2.0          # u = 2: odd integers are exact ties
    out["ties"] = ties
    ties4 = rng.choice([2.0, 6.0, 1.0, 3.0, 4.0], n)
    ties4[0] = 2.0 ** 60
    out["ties_large"] = ties4
    out["subnormal"] = rng.random(n) * 2.0 ** -1060
    out["mixed_scale"] = rng.random(n) * 2.0 ** rng.integers(-40, 40, n)
    out["integers"] = rng.integers(0, 2 ** 20, n).astype(np.float64) * 2.0 ** rng.integers(0, 30, n)
    zeros = rng.choice([0.0, -0.0, 1e-300, 3.0], n)
    out["signed_zeros"] = zeros
    out["negative"] = rng.normal(0, 10, n)
    out["coordinates"] = rng.normal(-7.0, 1.0, n)       # a centroid's feature sum
    half = rng.uniform(1, 2, n // 2) * 2.0 ** rng.integers(-30, 30, n // 2)
    out["cancelling"] = np.concatenate([half, -half, np.ones(n - 2 * (n // 2))])
    tn = rng.choice([-1.0, -3.0, 1.0, 3.0, -0.5, 2.0], n)
    tn[0] = -(2.0 ** 53 + 2.0)
    out["negative_ties"] = tn
    inf = rng.random(n)
    if n > 10:
        inf[n // 3] = np.inf
        inf[2 * n // 3] = np.nan
    out["inf_then_nan"] = inf
    big = np.ones(n)
    big[: min(3, n)] = 1e308
    out["overflow"] = big
    out["zeros"] = np.zeros(n)
    return out


@pytest.mark.parametrize("n", [1, 7, 1000, 4096, 4097, 50_000])
def test_left_fold_bit_exact(n):
    rng = np.random.default_rng(n)
    cs = cases(n, rng)
    names = list(cs)
    seqs = np.stack([cs[k] for k in names])
    sums, pre = device_fold(seqs)
    for b, name in enumerate(names):
        want = serial(seqs[b])
        assert same_bits(pre[b], want), (name, n)
        assert same_bits(sums[b], want[-1]), (name, n)


def test_left_fold_long_sequence_and_count():
    # several sequences at once (ld = n), 300k elements each: the k-means++
    # shape, with distances whose sum crosses many binades
    rng = np.random.default_rng(7)
    seqs = (rng.standard_normal((4, 300_000)) * 3.0) ** 2
    seqs[1, :1000] = 0.0                  # a run of zeros first
    seqs[2, 5] = 1e12                     # one huge addend early
    sums, pre = device_fold(seqs)
    for b in range(4):
        want = serial(seqs[b])
        assert same_bits(pre[b], want) and same_bits(sums[b], want[-1]), b
    assert math.isfinite(float(sums[0]))


def device_fold_ld(seqs, ld):
    """Sequences placed ld elements apart (ld >= n) in one device array."""
    count, n = seqs.shape
    buf = np.full((count, ld), 12345.0)
    buf[:, :n] = seqs
    dv = F.to_device(buf)
    sums = F.device.empty((count,), np.float64)
    pre = F.device.empty((count, ld), np.float64)
    check(L.lib().fl_left_fold_device(C.c_void_p(dv.data_ptr()), n, count, ld,
                                      C.c_void_p(sums.data_ptr()), C.c_void_p(pre.data_ptr())))
    return sums.numpy(), pre.numpy()[:, :n]


@pytest.mark.parametrize("n", [8191, 8192, 8193, 2048 * 9 + 1, 200_000])
def test_left_fold_split_sizes(n):
    # sequences of 8192 elements and more are cut into 2048-element blocks
    # done in parallel; every case of the small sizes again at these sizes
    rng = np.random.default_rng(n + 11)
    cs = cases(n, rng)
    names = list(cs)
    seqs = np.stack([cs[k] for k in names])
    sums, pre = device_fold(seqs)
    for b, name in enumerate(names):
        want = serial(seqs[b])
        assert same_bits(pre[b], want), (name, n)
        assert same_bits(sums[b], want[-1]), (name, n)


def test_left_fold_split_block_edge_cases():
    n = 2048 * 40 + 123
    rng = np.random.default_rng(99)
    seqs = []
    # an exact tie deep inside a middle block (the block is folded serially)
    a = np.ones(n)
    a[0] = 2.0 ** 53
    a[2048 * 17 + 900] = 1.0
    seqs.append(a)
    # NaN in a middle block, then infinities
    b = rng.random(n) * 10.0
    b[2048 * 5 + 3] = np.nan
    b[2048 * 30] = np.inf
    seqs.append(b)
    # negative running sums throughout
    seqs.append(-(rng.random(n) * 50.0) ** 2)
    # the running sum lands on a power of two exactly at a block boundary
    c = np.full(n, 0.25)
    c[:2048 * 8] = 2.0 ** 11 / (2048 * 8) * 4.0
    seqs.append(c)
    # sums that cancel to tiny values, so the plain block sums guess a wrong
    # binade for the blocks after the cancellation
    d = rng.uniform(1e6, 2e6, n)
    d[1::2] = -d[0::2][: n // 2]
    d[-1] = 1e-300
    seqs.append(d)
    # mixed magnitudes spanning the grid limit of a block (|x/u| near 2^51)
    e = rng.random(n)
    e[2048 * 3 + 7] = 2.0 ** 40
    e[2048 * 20 + 1] = 2.0 ** 45
    seqs.append(e)
    seqs = np.stack(seqs)
    sums, pre = device_fold(seqs)
    for i in range(seqs.shape[0]):
        want = serial(seqs[i])
        assert same_bits(pre[i], want), i
        assert same_bits(sums[i], want[-1]), i


def test_left_fold_split_strided():
    # count > 1 with ld > n: only each sequence's n elements are read and
    # only its n running sums written
    n, ld = 30_000, 30_037
    rng = np.random.default_rng(5)
    seqs = rng.gamma(2.0, 3.0, (3, n)) ** 2
    sums, pre = device_fold_ld(seqs, ld)
    for i in range(3):
        want = serial(seqs[i])
        assert same_bits(pre[i], want) and same_bits(sums[i], want[-1]), i
