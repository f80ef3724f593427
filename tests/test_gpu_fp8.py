"""The FP8 (e4m3) filter: centred rows scaled per row by a power of two, one
fp8 product per term on v_mfma_f32_16x16x32_fp8_fp8, a rigorous error bound
that widens each query's threshold, and the binary64 rerank.  Indices must
be the FP64 oracle's and the bound must hold on every re-scored candidate;
the filter's candidate counts are compared with the BF16 filter's (the
survivor inflation its wider bound costs)."""
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


@pytest.mark.parametrize("d,k", [(64, 10), (64, 50), (96, 32), (96, 100), (128, 10), (128, 100), (40, 16)])
def test_fp8_filter_matches_oracle(ctx, d, k):
    R, Q = datagen.mixture128(9000, 300, seed=800 + d + k, d=d)
    idx, dist, st = ctx.join(R, Q, k, precision=kj.PREC_FP8, filter_stats=True)
    assert st["precision_used"] == kj.PREC_FP8 and st["kernel_used"] == kj.KERNEL_MFMA
    assert_matches_oracle(R, Q, k, idx, dist)
    assert 0 < st["filter_err_ratio"] <= 1.0, st
    b = ctx.join(R, Q, k, precision=kj.PREC_BF16)
    assert np.array_equal(b[0], idx) and np.array_equal(b[1], dist)
    assert st["n_appended"] >= b[2]["n_appended"]


def test_fp8_split_mode_and_overflow_reruns(ctx):
    """Few tiles (split mode) and a capacity barely above k: the re-run
    ladder on the fp8 kernel, then the VALU kernel."""
    R, Q = datagen.mixture128(12_000, 200, seed=811)
    idx, dist, st = ctx.join(R, Q, 60, precision=kj.PREC_FP8, kcap=64, kcap_mfma_max=128)
    assert st["n_overflow"] > 0
    assert_matches_oracle(R, Q, 60, idx, dist)


def test_fp8_index_and_batch(ctx):
    R, Q = datagen.mixture128(10_000, 400, seed=812, d=96)
    ix = ctx.build_index(R, precision=kj.PREC_FP8)
    assert ix.info()["precision"] == kj.PREC_FP8
    idx, dist, st = ix.query(Q, 20)
    assert st["precision_used"] == kj.PREC_FP8
    assert_matches_oracle(R, Q, 20, idx, dist)
    ix.close()
    R2, Q2 = datagen.mixture128(6000, 250, seed=813, d=64)
    res, _ = ctx.join_batch([(R, Q, 20), (R2, Q2, 7)], precision=kj.PREC_FP8)
    assert_matches_oracle(R, Q, 20, res[0][0], res[0][1])
    assert_matches_oracle(R2, Q2, 7, res[1][0], res[1][1])


def test_fp8_argument_checks(ctx):
    R = np.random.default_rng(1).standard_normal((500, 32)).astype(np.float32)
    with pytest.raises(kj.KnnjError) as e:
        ctx.join(R, R[:10], 5, precision=kj.PREC_FP8)
    assert e.value.status == kj.KNNJ_EUNSUPPORTED
    R = np.random.default_rng(2).standard_normal((500, 64)).astype(np.float32)
    with pytest.raises(kj.KnnjError) as e:
        ctx.join(R, R[:10], 5, precision=kj.PREC_FP8, kernel=kj.KERNEL_VALU)
    assert e.value.status == kj.KNNJ_EINVAL
