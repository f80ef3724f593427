"""GPU parity of the serving path: a resident index (knnj_index_*) answers
exactly what knnj_join answers, and the continuous-batching server gives
every request the same result alone and fused with others."""
import concurrent.futures as cf

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


@pytest.fixture(scope="module")
def served():
    """A small version of the serving workload: 96-d Gaussian mixture."""
    mix = datagen.serving_model(seed=41, d=96, comps=64)
    R = mix.sample(datagen.rng(42), 30_000)
    batches = [mix.sample(datagen.rng(100 + i), n) for i, n in enumerate([1, 2, 5, 17, 64, 130, 333, 700, 1500])]
    return R, batches


def test_index_query_equals_join(ctx, served):
    R, batches = served
    ix = ctx.build_index(R)
    info = ix.info()
    assert info["n"] == R.shape[0] and info["d"] == 96 and info["bytes"] > 0
    for Q in batches[::2]:
        for k in (1, 50):
            a_idx, a_dist, st = ix.query(Q, k)
            b_idx, b_dist, _ = ctx.join(R, Q, k)
            assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    Q = np.concatenate(batches[-3:])
    idx, dist, st = ix.query(Q, 50)
    assert st["precision_used"] == info["precision"]
    assert_matches_oracle(R, Q, 50, idx, dist)
    ix.close()


def test_index_query_seeds_by_assignment(ctx, served, monkeypatch):
    """An index of 512 or more representatives starts each query row of the
    MFMA join at its nearest representative (an assignment of the query rows
    per call) instead of searching the cluster list; both give the oracle's
    answer, and the no-allocation entry point (which keeps the search) the
    same."""
    R, batches = served
    ix = ctx.build_index(R, m_r=1024)
    assert ix.info()["m_r"] == 1024
    Q = np.concatenate(batches[-3:])
    a_idx, a_dist, _ = ix.query(Q, 50)
    monkeypatch.setenv("KNNJ_SEED_LIST", "1")
    b_idx, b_dist, _ = ix.query(Q, 50)
    monkeypatch.delenv("KNNJ_SEED_LIST")
    assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    c_idx, c_dist, _ = ix.query(Q[:700], 50, no_alloc=True)
    assert np.array_equal(c_idx, a_idx[:700]) and np.array_equal(c_dist, a_dist[:700])
    assert_matches_oracle(R, Q, 50, a_idx, a_dist)
    ix.close()


@pytest.mark.parametrize("d,k,prec", [(32, 16, kj.PREC_AUTO), (64, 10, kj.PREC_FP32), (96, 200, kj.PREC_BF16)])
def test_index_precisions(ctx, d, k, prec):
    mix = datagen.mixture64_model(seed=d + 7, d=d, comps=30)
    R = mix.sample(datagen.rng(d), 8000)
    Q = mix.sample(datagen.rng(d + 1), 600)
    ix = ctx.build_index(R, precision=prec)
    idx, dist, _ = ix.query(Q, k)
    assert_matches_oracle(R, Q, k, idx, dist)
    ix.close()


def test_index_integer_exact_mode(ctx):
    X = datagen.skin_like(6000, seed=43)
    Q = datagen.skin_like(900, seed=44)
    ix = ctx.build_index(X)
    idx, dist, st = ix.query(Q, 64)
    assert st["exact_mode"] == 1
    assert_matches_oracle(X, Q, 64, idx, dist)
    # a non-integer query set against the same index falls back to the filter form
    Qf = Q + np.float32(0.25)
    idx, dist, st = ix.query(Qf, 64)
    assert st["exact_mode"] == 0
    assert_matches_oracle(X, Qf, 64, idx, dist)
    ix.close()


def test_index_shared_by_contexts(ctx, served):
    R, batches = served
    ix = ctx.build_index(R)
    other = kj.Context(0)
    a = ix.query(batches[5], 20)
    b = ix.query(batches[5], 20, ctx=other)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    other.close()
    ix.close()


def test_server_alone_equals_in_mix(ctx, served):
    """Requests submitted concurrently from several threads are fused into
    batches; each result equals the same request served alone."""
    R, batches = served
    ix = ctx.build_index(R)
    k = 50
    alone = [ix.query(Q, k)[:2] for Q in batches]
    srv = kj.Server(ix, k, slots=3, max_batch_rows=2048)
    reqs = [batches[i % len(batches)] for i in range(40)]

    def one(Q):
        t = srv.submit(Q)
        return srv.wait(t)

    with cf.ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(one, reqs))
    fused = 0
    for i, (idx, dist, info) in enumerate(outs):
        ref_idx, ref_dist = alone[i % len(batches)]
        assert np.array_equal(idx, ref_idx), f"request {i} indices differ in the mix"
        assert np.array_equal(dist, ref_dist), f"request {i} distances differ in the mix"
        assert info["latency_ms"] >= info["queue_ms"] >= 0.0
        assert info["t_submit_ms"] <= info["t_start_ms"] <= info["t_end_ms"]
        fused += info["batch_requests"] > 1
    nb, nrows = srv.counters()
    assert nrows == sum(Q.shape[0] for Q in reqs)
    assert nb <= len(reqs)
    # a real mix: some batch fused several requests, and batches of two
    # different slots were in flight at the same time
    assert fused >= 1, "no request was fused with another"
    spans = {(inf["slot"], inf["t_start_ms"], inf["t_end_ms"]) for _, _, inf in outs}
    overlap = any(s1 != s2 and a1 < b2 and a2 < b1 for (s1, a1, b1) in spans for (s2, a2, b2) in spans)
    assert overlap, f"no two slots' batches overlapped: {sorted(spans)[:6]}"
    srv.close()
    ix.close()


def test_server_rejects_oversized_request(ctx, served):
    R, batches = served
    ix = ctx.build_index(R)
    srv = kj.Server(ix, 5, slots=1, max_batch_rows=100)
    with pytest.raises(kj.KnnjError) as e:
        srv.submit(batches[7])  # 700 rows
    assert e.value.status == kj.KNNJ_EUNSUPPORTED
    t = srv.submit(batches[3])
    idx, dist, info = srv.wait(t)
    assert idx.shape == (17, 5) and info["status"] == kj.KNNJ_OK
    srv.close()
    ix.close()
