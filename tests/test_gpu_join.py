"""GPU parity: the engine's join against the FP64 oracle, bit-identical
indices under the (squared distance, row) key.  Sizes are chosen so the
oracle finishes in seconds."""
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


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_integer_selfjoin_k200_duplicates(ctx):
    X = datagen.skin_like(3000, seed=21)
    idx, dist, st = ctx.join(X, X, 200)
    assert st["exact_mode"] == 1
    assert_matches_oracle(X, X, 200, idx, dist)


def test_integer_selfjoin_bruteforce_mode(ctx):
    X = datagen.skin_like(2500, seed=22)
    idx, dist, st = ctx.join(X, X, 64, brute_force=True)
    assert st["m_r"] == 1 and st["m_q"] == 1
    assert_matches_oracle(X, X, 64, idx, dist)


@pytest.mark.parametrize("k", [1, 7, 64, 256])
def test_integer_d3_various_k(ctx, k):
    R = datagen.skin_like(4000, seed=30 + k)
    Q = datagen.skin_like(500, seed=60 + k)
    idx, dist, st = ctx.join(R, Q, k)
    assert st["exact_mode"] == 1
    assert_matches_oracle(R, Q, k, idx, dist)


@pytest.mark.parametrize("d,k", [(2, 5), (16, 16), (64, 32), (128, 10), (128, 100), (96, 50)])
def test_real_mixture(ctx, d, k):
    mix = datagen.mixture64_model(seed=d, d=d, comps=24)
    R = mix.sample(datagen.rng(d * 3 + 1), 6000)
    Q = mix.sample(datagen.rng(d * 3 + 2), 400)
    idx, dist, st = ctx.join(R, Q, k)
    assert st["exact_mode"] == 0
    assert_matches_oracle(R, Q, k, idx, dist)


def test_real_mixture128_normed(ctx):
    R, Q = datagen.mixture128(8000, 300, seed=77)
    idx, dist, st = ctx.join(R, Q, 100)
    assert_matches_oracle(R, Q, 100, idx, dist)
    assert st["filter_max_abs_err"] >= 0


def test_device_inputs_and_outputs(ctx):
    import torch
    R, Q = datagen.mixture64(5000, 300, seed=5)
    idx, dist, _ = ctx.join(_dev(R), _dev(Q), 32)
    assert idx.is_cuda and dist.is_cuda
    torch.cuda.synchronize()
    assert_matches_oracle(R, Q, 32, idx.cpu().numpy(), dist.cpu().numpy())


def test_squared_and_aos_outputs(ctx):
    X = datagen.skin_like(2000, seed=8)
    idx, dist, _ = ctx.join(X, X[:300], 20, out_squared=True)
    assert_matches_oracle(X, X[:300], 20, idx, dist, squared=True)
    aos, none, _ = ctx.join(X, X[:300], 20, out_aos=True)
    assert none is None and aos.shape == (300, 20, 2)
    assert (aos[..., 0] == idx).all()
    d_aos = aos[..., 1].copy().view(np.float32)
    assert_matches_oracle(X, X[:300], 20, aos[..., 0], d_aos)


def test_k_equals_n_and_tiny(ctx):
    g = np.random.default_rng(3)
    R = g.standard_normal((50, 3)).astype(np.float32)
    Q = g.standard_normal((9, 3)).astype(np.float32)
    idx, dist, _ = ctx.join(R, Q, 50)
    assert_matches_oracle(R, Q, 50, idx, dist)
    idx, dist, _ = ctx.join(R[:1], Q, 1)
    assert (idx == 0).all()


def test_all_identical_points(ctx):
    X = np.full((700, 5), 3.0, np.float32)
    idx, dist, st = ctx.join(X, X, 64)
    assert (idx == np.arange(64)[None, :]).all() and (dist == 0).all()


def test_integer_line_ties(ctx):
    R = np.arange(500, dtype=np.float32)[:, None]
    Q = np.array([[250.0], [0.0], [499.0], [100.5]], np.float32)
    idx, dist, _ = ctx.join(R, Q, 9)
    assert idx[0].tolist() == [250, 249, 251, 248, 252, 247, 253, 246, 254]
    assert_matches_oracle(R, Q, 9, idx, dist)


def test_empty_query_set(ctx):
    R = np.ones((10, 3), np.float32)
    Q = np.zeros((0, 3), np.float32)
    idx, dist, _ = ctx.join(R, Q, 3)
    assert idx.shape == (0, 3)


def test_errors(ctx):
    R = np.ones((10, 3), np.float32)
    with pytest.raises(kj.KnnjError) as e:
        ctx.join(R, R, 11)
    assert e.value.status == kj.KNNJ_EINVAL
    bad = R.copy()
    bad[3, 1] = np.nan
    with pytest.raises(kj.KnnjError) as e:
        ctx.join(bad, R, 2)
    assert e.value.status == kj.KNNJ_EINVAL
    big = np.ones((2000, 2), np.float32)
    with pytest.raises(kj.KnnjError) as e:
        ctx.join(big, big[:3], 1500)
    assert e.value.status == kj.KNNJ_EUNSUPPORTED


def test_integer_selfjoin_50k_sampled_rows(ctx):
    X = datagen.skin_like(50_000, seed=2016)
    idx, dist, st = ctx.join(X, X, 200)
    rows = np.random.default_rng(0).choice(X.shape[0], 1500, replace=False)
    assert_matches_oracle(X, X, 200, idx, dist, rows=rows)
    assert st["pruning_ratio"] > 0.5


@pytest.mark.parametrize("d,k", [(32, 16), (64, 32), (128, 100), (200, 20)])
def test_mfma_fp32_real(ctx, d, k):
    mix = datagen.mixture64_model(seed=d + 5, d=d, comps=30)
    R = mix.sample(datagen.rng(d * 5 + 1), 7000)
    Q = mix.sample(datagen.rng(d * 5 + 2), 300)
    idx, dist, st = ctx.join(R, Q, k, kernel=kj.KERNEL_MFMA, precision=kj.PREC_FP32)
    assert st["kernel_used"] == kj.KERNEL_MFMA and st["precision_used"] == kj.PREC_FP32
    assert_matches_oracle(R, Q, k, idx, dist)


@pytest.mark.parametrize("prec", [kj.PREC_BF16, kj.PREC_BF16X1])
@pytest.mark.parametrize("d,k", [(64, 10), (96, 50), (128, 100), (128, 10), (250, 30)])
def test_mfma_bf16_filter(ctx, d, k, prec):
    R, Q = datagen.mixture128(9000, 300, seed=d + k, d=d)
    idx, dist, st = ctx.join(R, Q, k, precision=prec)
    assert st["kernel_used"] == kj.KERNEL_MFMA and st["precision_used"] == prec
    assert_matches_oracle(R, Q, k, idx, dist)


def test_mfma_exact_integer(ctx):
    g = np.random.default_rng(12)
    R = g.integers(0, 16, (5000, 32)).astype(np.float32)
    Q = g.integers(0, 16, (300, 32)).astype(np.float32)
    idx, dist, st = ctx.join(R, Q, 40, kernel=kj.KERNEL_MFMA)
    assert st["exact_mode"] == 1 and st["kernel_used"] == kj.KERNEL_MFMA
    assert_matches_oracle(R, Q, 40, idx, dist)


def test_valu_and_mfma_agree(ctx):
    R, Q = datagen.mixture64(6000, 400, seed=31)
    a = ctx.join(R, Q, 32, kernel=kj.KERNEL_VALU)
    b = ctx.join(R, Q, 32, kernel=kj.KERNEL_MFMA, precision=kj.PREC_FP32)
    c = ctx.join(R, Q, 32, precision=kj.PREC_BF16)
    assert (a[0] == b[0]).all() and (a[0] == c[0]).all()
    assert (a[1] == b[1]).all() and (a[1] == c[1]).all()


@pytest.mark.parametrize("refine", [-1, 2])
def test_refinement_options(ctx, refine):
    R, Q = datagen.mixture128(6000, 200, seed=5, d=96)
    idx, dist, _ = ctx.join(R, Q, 20, refine_iters=refine)
    assert_matches_oracle(R, Q, 20, idx, dist)
    X = datagen.skin_like(3000, seed=4)
    idx, dist, _ = ctx.join(X, X[:400], 50, refine_iters=refine)
    assert_matches_oracle(X, X[:400], 50, idx, dist)


def test_xcd_tile_placement(ctx):
    R, Q = datagen.mixture128(8000, 700, seed=41, d=64)
    a = ctx.join(R, Q, 16)
    b = ctx.join(R, Q, 16, xcd_tiles=True)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    assert_matches_oracle(R, Q, 16, b[0], b[1])


@pytest.mark.parametrize("d,k,prec", [(64, 32, kj.PREC_FP32), (96, 50, kj.PREC_BF16), (128, 200, kj.PREC_BF16),
                                      (64, 16, kj.PREC_BF16X1)])
def test_split_mode_matches_single_wave_tiles(ctx, d, k, prec):
    """Split mode (the waves of a workgroup share one tile's cluster list,
    thresholds shared through LDS, lists merged at the end) and one-wave
    tiles give the oracle's answer."""
    mix = datagen.mixture64_model(seed=300 + d, d=d, comps=40)
    R = mix.sample(datagen.rng(d + 301), 12000)
    Q = mix.sample(datagen.rng(d + 302), 700)
    a_idx, a_dist, sa = ctx.join(R, Q, k, precision=prec, tile_split=4)
    b_idx, b_dist, sb = ctx.join(R, Q, k, precision=prec, tile_split=1)
    assert sa["tile_q"] == 16 and sb["tile_q"] == 16
    assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    assert_matches_oracle(R, Q, k, a_idx, a_dist)


def test_split_mode_exact_integer(ctx):
    R = datagen.skin_like(9000, seed=310)
    Q = datagen.skin_like(800, seed=311)
    # integer inputs small enough for exact MFMA keys
    idx, dist, st = ctx.join(R / 1, Q / 1, 100, kernel=kj.KERNEL_MFMA, tile_split=4)
    assert st["exact_mode"] == 1
    assert_matches_oracle(R, Q, 100, idx, dist)


@pytest.mark.parametrize("d,k", [(3, 50), (64, 16)])
def test_autotune_gives_the_same_answer(ctx, d, k):
    """The autotuner times its candidates on the first call of a shape and
    reuses the choice; every call returns the oracle's answer."""
    mix = datagen.mixture64_model(seed=400 + d, d=d, comps=30)
    R = mix.sample(datagen.rng(d + 401), 9000)
    Q = mix.sample(datagen.rng(d + 402), 800)
    tctx = kj.Context(0)
    a_idx, a_dist, sa = tctx.join(R, Q, k, autotune=True)
    b_idx, b_dist, sb = tctx.join(R, Q, k, autotune=True)
    assert sa["tuned"] >= 1 and sb["tuned"] == sa["tuned"]
    assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    assert_matches_oracle(R, Q, k, a_idx, a_dist)
    # a later call of the shape replays the choice with its own per-call
    # options (here squared distances out), not the first call's
    c_idx, c_dist, sc = tctx.join(R, Q, k, autotune=True, out_squared=True)
    assert sc["tuned"] == sa["tuned"]
    assert np.array_equal(c_idx, a_idx)
    np.testing.assert_allclose(np.sqrt(c_dist.astype(np.float64)), a_dist, rtol=1e-6)
    tctx.close()


@pytest.mark.parametrize("prec", [kj.PREC_BF16, kj.PREC_FP32])
def test_seeds_by_assignment(ctx, prec, monkeypatch):
    """With 512 or more reference representatives the MFMA join starts each
    query row at its nearest representative from an assignment of the query
    rows (in a self-join, the row's own cluster) instead of searching its
    cluster list.  Both starts give the oracle's answer, bit for bit."""
    mix = datagen.mixture64_model(seed=510, d=64, comps=40)
    R = mix.sample(datagen.rng(511), 20000)
    Q = mix.sample(datagen.rng(512), 1500)
    a_idx, a_dist, sa = ctx.join(R, Q, 32, precision=prec, m_r=1024)
    assert sa["m_r"] == 1024 and sa["kernel_used"] == kj.KERNEL_MFMA
    monkeypatch.setenv("KNNJ_SEED_LIST", "1")
    b_idx, b_dist, _ = ctx.join(R, Q, 32, precision=prec, m_r=1024)
    monkeypatch.delenv("KNNJ_SEED_LIST")
    assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    assert_matches_oracle(R, Q, 32, a_idx, a_dist)
    X = np.ascontiguousarray(R[:8000])
    c_idx, c_dist, sc = ctx.join(X, X, 16, precision=prec, m_r=600)
    assert sc["m_r"] == 600 and sc["kernel_used"] == kj.KERNEL_MFMA
    assert_matches_oracle(X, X, 16, c_idx, c_dist)


@pytest.mark.parametrize("d", [3, 64])
def test_many_tiny_reference_clusters(ctx, d):
    """About 1.5 rows per reference cluster and k = 500: gathering k bound
    candidates needs several hundred clusters, more than the one-at-a-time
    pass takes, so the full-sort bounds pass answers these query clusters."""
    mix = datagen.mixture64_model(seed=d + 90, d=d, comps=20)
    R = mix.sample(datagen.rng(d + 91), 3000)
    Q = mix.sample(datagen.rng(d + 92), 300)
    idx, dist, st = ctx.join(R, Q, 500, m_r=2000)
    assert st["m_r"] == 2000
    assert_matches_oracle(R, Q, 500, idx, dist)


@pytest.mark.parametrize("d,extra,shift", [(64, 4, 0), (64, 3, 0), (32, 4, 1), (96, 0, 0)])
def test_strided_device_inputs(ctx, d, extra, shift):
    """Device row views with ld > d and unaligned starts: the 16-byte index
    kernels (ld % 4 == 0, 16-byte aligned rows) and the scalar ones give the
    oracle's answer."""
    import torch
    mix = datagen.mixture64_model(seed=d + 70 + extra, d=d, comps=12)
    R = mix.sample(datagen.rng(d + 71), 5000)
    Q = mix.sample(datagen.rng(d + 72), 700)

    def view(X):
        W = torch.zeros((X.shape[0], d + extra + shift), dtype=torch.float32, device="cuda")
        W[:, shift:shift + d] = torch.from_numpy(X).cuda()
        return W[:, shift:shift + d]

    Rv, Qv = view(R), view(Q)
    assert Rv.stride(0) == d + extra + shift
    idx, dist, _ = ctx.join(Rv, Qv, 16)
    torch.cuda.synchronize()
    assert_matches_oracle(R, Q, 16, idx.cpu().numpy(), dist.cpu().numpy())
