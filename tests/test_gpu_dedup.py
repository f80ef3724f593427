"""Exact duplicate queries answered once (knnj_opts.dedup): the join runs
for the distinct queries and copies each duplicate's output row from its
leader's.  Every result must equal the join without folding, bit for bit,
and the FP64 oracle."""
import numpy as np
import pytest

import knn_sweet_amd as kj
from knn_sweet_amd import datagen
from helpers import assert_matches_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = kj.Context(0)
    yield c
    c.close()


def _distinct(X):
    """Bitwise-distinct rows (the folding compares bit patterns: -0.0 and
    0.0 are different rows to it)."""
    return np.unique(np.ascontiguousarray(X).view(np.uint32), axis=0).shape[0]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:2], b[:2]) if x is not None)


@pytest.mark.parametrize("k", [1, 50, 200])
def test_integer_selfjoin_folds_duplicates(ctx, k):
    """Skin-like integer self-join: about 4 rows per distinct triple."""
    X = datagen.skin_like(30_000, seed=701)
    a = ctx.join(X, X, k)  # auto: on for integer-mode VALU joins
    b = ctx.join(X, X, k, dedup=1)
    distinct = _distinct(X)
    assert a[2]["n_distinct_queries"] == distinct, (a[2]["n_distinct_queries"], distinct)
    assert b[2]["n_distinct_queries"] == X.shape[0]
    assert _same(a, b)
    rows = np.random.default_rng(7).choice(X.shape[0], 800, replace=False)
    assert_matches_oracle(X, X, k, a[0], a[1], rows=rows)


def test_reference_skin_rows(ctx):
    """The reference's own Skin rows at d = 3 and d = 4 (label included)."""
    import os
    rows = np.load(os.path.join(os.path.dirname(__file__), "golden", "skin_rows_u8.npz"))["rows"][:60_000]
    for d in (3, 4):
        X = np.ascontiguousarray(rows[:, :d].astype(np.float32))
        a = ctx.join(X, X, 100)
        b = ctx.join(X, X, 100, dedup=1)
        assert a[2]["n_distinct_queries"] < X.shape[0] // 2
        assert _same(a, b)
        sel = np.random.default_rng(8 + d).choice(X.shape[0], 600, replace=False)
        assert_matches_oracle(X, X, 100, a[0], a[1], rows=sel)


def test_outputs_aos_squared_and_device(ctx):
    import torch
    X = datagen.skin_like(12_000, seed=702)
    Q = np.concatenate([X[:3000], X[:3000]])  # every query twice, non-self join
    aos, _, st = ctx.join(X, Q, 40, out_aos=True)
    assert st["n_distinct_queries"] <= 3000
    ref = ctx.join(X, Q, 40, dedup=1)
    assert np.array_equal(aos[..., 0], ref[0])
    assert np.array_equal(aos[..., 1].copy().view(np.float32), ref[1])
    sq = ctx.join(X, Q, 40, out_squared=True)
    assert_matches_oracle(X, Q, 40, sq[0], sq[1], squared=True, rows=np.arange(0, 6000, 5))
    Xd, Qd = torch.from_numpy(X).cuda(), torch.from_numpy(Q).cuda()
    di, dd, _ = ctx.join(Xd, Qd, 40)
    torch.cuda.synchronize()
    assert np.array_equal(di.cpu().numpy(), ref[0]) and np.array_equal(dd.cpu().numpy(), ref[1])


@pytest.mark.parametrize("kernel,prec", [(kj.KERNEL_MFMA, kj.PREC_FP32), (kj.KERNEL_MFMA, kj.PREC_BF16),
                                         (kj.KERNEL_VALU, kj.PREC_FP32)])
def test_forced_on_real_data_with_planted_duplicates(ctx, kernel, prec):
    """dedup = 2 on real-valued data (filter keys, re-ranked) whose queries
    repeat: same answer as without folding."""
    mix = datagen.mixture64_model(seed=703, d=64, comps=20)
    R = mix.sample(datagen.rng(704), 15_000)
    Q0 = mix.sample(datagen.rng(705), 700)
    Q = Q0[np.random.default_rng(9).integers(0, 700, 2000)]
    a = ctx.join(R, Q, 32, kernel=kernel, precision=prec, dedup=2)
    b = ctx.join(R, Q, 32, kernel=kernel, precision=prec, dedup=1)
    assert a[2]["n_distinct_queries"] == _distinct(Q)
    assert _same(a, b)
    assert_matches_oracle(R, Q, 32, a[0], a[1], rows=np.arange(0, 2000, 3))


def test_forced_overflow_reruns_then_copies(ctx):
    """Overflowed leaders are re-run before their rows are copied."""
    mix = datagen.mixture64_model(seed=706, d=128, comps=16)
    R = mix.sample(datagen.rng(707), 15_000)
    Q = mix.sample(datagen.rng(708), 300)
    Q = np.concatenate([Q, Q, Q])
    a = ctx.join(R, Q, 60, precision=kj.PREC_BF16X1, kcap=64, dedup=2)
    assert a[2]["n_overflow"] > 0 and a[2]["n_distinct_queries"] == 300
    b = ctx.join(R, Q, 60, precision=kj.PREC_BF16X1, dedup=1)
    assert _same(a, b)
    assert_matches_oracle(R, Q, 60, a[0], a[1], rows=np.arange(0, 900, 4))


def test_batched_integer_joins_fold_duplicates(ctx):
    """knnj_join_batch fills every instance's kernel arguments before it
    queues the index builds: the leaders' side must already be in them
    (regression: it was filled by the dedup launch, after the arguments)."""
    X = datagen.skin_like(9000, seed=709)
    Y = datagen.skin_like(5000, seed=710)
    res, st = ctx.join_batch([(X, X, 64), (Y, X[:2000], 30), (Y, Y, 200)])
    for (R, Q, k), (idx, dist) in zip([(X, X, 64), (Y, X[:2000], 30), (Y, Y, 200)], res):
        a_idx, a_dist, a_st = ctx.join(R, Q, k)
        assert a_st["n_distinct_queries"] < Q.shape[0]
        assert np.array_equal(idx, a_idx) and np.array_equal(dist, a_dist)
        assert_matches_oracle(R, Q, k, idx, dist, rows=np.arange(0, Q.shape[0], 9))
