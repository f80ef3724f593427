"""Pins the CPU oracle: closed forms, an independent numpy restatement of the
key, brute force on tiny inputs, and the CPU triangle-inequality join."""
import numpy as np
import pytest

import oracle
from knn_sweet_amd import datagen


def np_sq64(R, q):
    """The key's binary64 sequence, dimension by dimension (numpy never fuses)."""
    R = R.astype(np.float64)
    q = q.astype(np.float64)
    s = (q[0] - R[:, 0]) * (q[0] - R[:, 0])
    for j in range(1, R.shape[1]):
        t = q[j] - R[:, j]
        s = s + t * t
    return s


def np_knn(R, Q, k):
    idx = np.empty((Q.shape[0], k), np.int32)
    sq = np.empty((Q.shape[0], k), np.float64)
    ar = np.arange(R.shape[0])
    for i, q in enumerate(Q):
        s = np_sq64(R, q)
        o = np.lexsort((ar, s))[:k]
        idx[i] = o
        sq[i] = s[o]
    return idx, sq


def test_all_identical_points():
    R = np.full((50, 4), 7.25, np.float32)
    idx, dist, sq = oracle.bruteforce(R, R[:5], 10)
    assert (idx == np.arange(10)[None, :]).all()
    assert (dist == 0).all() and (sq == 0).all()


def test_integer_line_tie_break():
    n = 40
    R = np.arange(n, dtype=np.float32)[:, None]
    Q = np.array([[20.0], [0.0], [39.0]], np.float32)
    idx, dist, _ = oracle.bruteforce(R, Q, 7)
    # equal distances on both sides: the lower index comes first
    assert idx[0].tolist() == [20, 19, 21, 18, 22, 17, 23]
    assert dist[0].tolist() == [0, 1, 1, 2, 2, 3, 3]
    assert idx[1].tolist() == list(range(7))
    assert idx[2].tolist() == list(range(39, 32, -1))


def test_lattice_shells():
    g = np.arange(5, dtype=np.float32)
    R = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    center = np.array([[2.0, 2.0, 2.0]], np.float32)
    idx, _, sq = oracle.bruteforce(R, center, 1 + 6 + 12 + 8)
    shells = np.bincount(sq[0].astype(np.int64))
    assert shells.tolist() == [1, 6, 12, 8]
    # within a shell, ascending row index
    for s in range(4):
        rows = idx[0][sq[0] == s]
        assert (np.diff(rows) > 0).all()


def test_k_equals_n_full_sort():
    g = np.random.default_rng(1)
    R = g.standard_normal((300, 5)).astype(np.float32)
    Q = g.standard_normal((4, 5)).astype(np.float32)
    idx, _, sq = oracle.bruteforce(R, Q, 300)
    ref_idx, ref_sq = np_knn(R, Q, 300)
    assert (idx == ref_idx).all()
    assert (sq == ref_sq).all()


@pytest.mark.parametrize("d,k", [(1, 3), (3, 50), (17, 8), (64, 32), (128, 100)])
def test_random_matches_numpy_restatement(d, k):
    g = np.random.default_rng(d * 100 + k)
    R = g.standard_normal((700, d)).astype(np.float32)
    Q = g.standard_normal((40, d)).astype(np.float32)
    idx, dist, sq = oracle.bruteforce(R, Q, k)
    ref_idx, ref_sq = np_knn(R, Q, k)
    assert (idx == ref_idx).all()
    assert (sq == ref_sq).all()
    assert (dist == np.sqrt(ref_sq).astype(np.float32)).all()


def test_duplicates_heavy_integer():
    X = datagen.skin_like(2000, seed=5)
    idx, _, sq = oracle.bruteforce(X, X[:200], 200)
    ref_idx, ref_sq = np_knn(X, X[:200], 200)
    assert (idx == ref_idx).all() and (sq == ref_sq).all()
    # integer inputs: binary64 keys are the exact integers
    assert (sq == np.rint(sq)).all()


def test_pair_sq64_matches_key():
    g = np.random.default_rng(3)
    R = g.standard_normal((100, 33)).astype(np.float32)
    Q = g.standard_normal((10, 33)).astype(np.float32)
    qi = g.integers(0, 10, 500).astype(np.int32)
    ri = g.integers(0, 100, 500).astype(np.int32)
    got = oracle.pair_sq64(R, Q, qi, ri)
    ref = np.array([np_sq64(R[r:r + 1], Q[q])[0] for q, r in zip(qi, ri)])
    assert (got == ref).all()


def test_certify_detects_misorder():
    g = np.random.default_rng(4)
    R = g.standard_normal((500, 16)).astype(np.float32)
    Q = g.standard_normal((20, 16)).astype(np.float32)
    idx, _, _ = oracle.bruteforce(R, Q, 30)
    assert oracle.certify(R, Q, idx) == 0
    bad = idx.copy()
    bad[:, [3, 4]] = bad[:, [4, 3]]  # planted mistake: swap two neighbours
    assert oracle.certify(R, Q, bad) == 20


@pytest.mark.parametrize("case", ["int3", "real16", "real64", "tiny", "dup"])
def test_ti_join_equals_bruteforce(case):
    g = np.random.default_rng(11)
    if case == "int3":
        R = datagen.skin_like(3000, seed=9)
        Q, k = R, 100
    elif case == "real16":
        mix = datagen.mixture64_model(seed=3, d=16, comps=20)
        R = mix.sample(datagen.rng(1), 4000)
        Q = mix.sample(datagen.rng(2), 300)
        k = 16
    elif case == "real64":
        R, Q = datagen.mixture64(3000, 200, seed=8)
        k = 32
    elif case == "tiny":
        R = g.standard_normal((7, 2)).astype(np.float32)
        Q = g.standard_normal((5, 2)).astype(np.float32)
        k = 7
    else:
        R = np.repeat(g.integers(0, 5, (30, 3)).astype(np.float32), 20, axis=0)
        Q, k = R[::7], 60
    idx, dist, sq = oracle.bruteforce(R, Q, k)
    tidx, tdist, tsq, ev = oracle.ti_join(R, Q, k)
    assert (tidx == idx).all()
    assert (tsq == sq).all()
    assert ev <= R.shape[0] * Q.shape[0]


def test_generators_deterministic():
    a = datagen.skin_like(1000, seed=2016)
    b = datagen.skin_like(1000, seed=2016)
    assert datagen.sha256(a) == datagen.sha256(b)
    R1, Q1 = datagen.mixture128(2000, 100)
    nr = np.linalg.norm(R1.astype(np.float64), axis=1)
    assert nr.min() >= 0.5 - 1e-6 and nr.max() <= 2.0 + 1e-6
    assert a.min() >= 0 and a.max() <= 255 and (a == np.rint(a)).all()
