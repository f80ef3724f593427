"""GPU checks of what the join's exactness rests on, from the device's own
state (knnj_ctx_copy_plan_part) and its counters:

  - the cluster index: offsets and permutation, distances to the
    representative ascending per cluster, radius = largest member distance
    (the reference never computes the reference-side radius,
    src/knnjoin.cu:95, 391-392, and uses it at knnjoin.cu:515);
  - theta_A >= the true k-th neighbour distance of every query of cluster A
    (the reference's tie-broken upper-bound list, knnjoin.cu:484-494, can
    underestimate it);
  - no true neighbour lies in a pruned (query cluster, reference cluster)
    pair, and each listed cluster's lower bound is <= d(a, s) for all its
    members s (knnjoin.cu:506-525);
  - the filter's rigorous error bound holds on every re-scored candidate;
  - every stage of the overflow re-run ladder (the MFMA kernel at twice the
    capacity, then the VALU kernel at 1024) gives the oracle's answer;
  - the multi-wave lockstep workgroup gives the one-wave tiles' answer.
"""
import numpy as np
import pytest

import knn_sweet_amd as kj
import oracle
from knn_sweet_amd import datagen
from helpers import assert_matches_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = kj.Context(0)
    yield c
    c.close()


def _clusters(off, n):
    """Cluster id of every position of a side."""
    cl = np.empty(n, np.int64)
    for b in range(off.shape[0] - 1):
        cl[off[b]:off[b + 1]] = b
    return cl


def _dist64(A, B):
    A = A.astype(np.float64)
    B = B.astype(np.float64)
    return np.sqrt(((A - B) ** 2).sum(axis=1))


def check_plan(ctx, R, Q, k, rows, self_join=False):
    """Every invariant above on the context's last join, for the sampled
    query rows `rows` (original indices into Q)."""
    P = ctx.last_plan()
    assert P["valid"] == 1 and P["n_r"] == R.shape[0] and P["n_q"] == Q.shape[0] and P["k"] == k
    d = R.shape[1]
    sides = {}
    for side, X in (("r", R), ("q", Q)):
        off = ctx.plan_part(f"{side}_offsets").astype(np.int64)
        perm = ctx.plan_part(f"{side}_perm").astype(np.int64)
        pdist = ctx.plan_part(f"{side}_pdist")
        reps = ctx.plan_part(f"{side}_reps")[:, :d]
        rad = ctx.plan_part(f"{side}_radius")
        n, m = X.shape[0], off.shape[0] - 1
        assert off[0] == 0 and off[-1] == n and (np.diff(off) >= 0).all()
        assert np.array_equal(np.sort(perm), np.arange(n)), f"{side}: perm is not a permutation"
        cl = _clusters(off, n)
        dd = _dist64(X[perm], reps[cl])
        # distances to the representative: binary32 of the binary64 value, ascending per cluster
        np.testing.assert_allclose(pdist, dd, rtol=1e-6, atol=1e-30)
        same = cl[1:] == cl[:-1]
        assert (pdist[1:][same] >= pdist[:-1][same]).all(), f"{side}: pdist not ascending in a cluster"
        # radius = largest member distance (binary64; summation order differs from numpy's)
        mx = np.zeros(m)
        np.maximum.at(mx, cl, dd)
        np.testing.assert_allclose(rad, mx, rtol=1e-12, atol=0)
        sides[side] = (off, perm, cl, reps, rad)
    r_off, r_perm, r_cl, r_reps, _ = sides["r"]
    q_off, q_perm, q_cl, q_reps, _ = sides["q"]
    theta = ctx.plan_part("theta")
    pcnt = ctx.plan_part("pair_count")
    plist = ctx.plan_part("pair_list")
    plb = ctx.plan_part("pair_lb")
    r_pos = np.empty_like(r_perm)
    r_pos[r_perm] = np.arange(r_perm.shape[0])
    q_pos = np.empty_like(q_perm)
    q_pos[q_perm] = np.arange(q_perm.shape[0])
    ref_idx, _, ref_sq = oracle.bruteforce(R, Q[rows], k)
    kth = np.sqrt(ref_sq[:, k - 1])
    checked_lb = set()
    for j, q in enumerate(rows):
        A = int(q_cl[q_pos[q]])
        assert theta[A] >= kth[j], f"theta[{A}] = {theta[A]!r} below query {q}'s k-th distance {kth[j]!r}"
        listed = set(plist[A, :pcnt[A]].tolist())
        nb_clusters = set(r_cl[r_pos[ref_idx[j]]].tolist())
        lost = nb_clusters - listed
        assert not lost, f"query {q} (cluster {A}): true neighbours in pruned clusters {sorted(lost)[:5]}"
        if A not in checked_lb:
            checked_lb.add(A)
            lbs = plb[A, :pcnt[A]]
            assert (np.diff(lbs) >= 0).all(), f"cluster {A}: list not ascending by lower bound"
            for t, b in enumerate(plist[A, :pcnt[A]][:64]):
                mem = R[r_perm[r_off[b]:r_off[b + 1]]]
                if mem.shape[0]:
                    dmin = _dist64(mem, np.broadcast_to(q_reps[A], mem.shape)).min()
                    assert lbs[t] <= dmin, f"pair ({A}, {b}): lower bound {lbs[t]} above d(a, s) {dmin}"
    if self_join:
        assert np.array_equal(r_perm, q_perm)
    return P


def test_plan_invariants_c3_like(ctx):
    """d = 128 mixture at the benchmarked plan's choices: 4,096 reference
    representatives and two Lloyd steps."""
    R, Q = datagen.mixture128(300_000, 20_000, seed=501)
    idx, dist, st = ctx.join(R, Q, 100, m_r=4096, refine_iters=2)
    assert st["m_r"] == 4096
    rows = np.random.default_rng(5).choice(Q.shape[0], 300, replace=False)
    check_plan(ctx, R, Q, 100, rows)
    assert_matches_oracle(R, Q, 100, idx, dist, rows=rows)


def test_plan_invariants_integer_selfjoin(ctx):
    """d = 3 integer self-join with many duplicates, k = 200."""
    X = datagen.skin_like(60_000, seed=502)
    idx, dist, st = ctx.join(X, X, 200)
    assert st["exact_mode"] == 1
    rows = np.random.default_rng(6).choice(X.shape[0], 400, replace=False)
    check_plan(ctx, X, X, 200, rows, self_join=True)
    assert_matches_oracle(X, X, 200, idx, dist, rows=rows)


@pytest.mark.parametrize("d,k,prec,kern", [
    (128, 100, kj.PREC_BF16, kj.KERNEL_AUTO),
    (128, 10, kj.PREC_BF16X1, kj.KERNEL_AUTO),
    (64, 32, kj.PREC_FP32, kj.KERNEL_MFMA),
    (64, 32, kj.PREC_FP32, kj.KERNEL_VALU),
    (96, 50, kj.PREC_BF16, kj.KERNEL_AUTO),
])
def test_filter_error_within_its_bound(ctx, d, k, prec, kern):
    """Over every re-scored candidate, the filter value's error against
    binary64 is within the bound the pruning used (ratio <= 1), and the
    error itself is reported apart from the bound."""
    R, Q = datagen.mixture128(20_000, 500, seed=503 + d, d=d)
    idx, dist, st = ctx.join(R, Q, k, precision=prec, kernel=kern, filter_stats=True)
    assert_matches_oracle(R, Q, k, idx, dist)
    plain = ctx.join(R, Q, k, precision=prec, kernel=kern)
    assert np.array_equal(plain[0], idx) and np.array_equal(plain[1], dist)
    assert plain[2]["filter_err_max"] == 0  # measured only on request
    assert st["filter_bound_max"] > 0 and st["filter_err_max"] > 0
    assert st["filter_err_ratio"] <= 1.0, st
    assert st["filter_err_max"] <= st["filter_bound_max"]


@pytest.mark.parametrize("kcap,mfma_max,prec", [
    (64, 64, kj.PREC_BF16X1),    # overflow -> straight to the VALU kernel at 1024
    (64, 512, kj.PREC_BF16X1),   # MFMA at 128, 256, ... in split mode, then VALU if needed
    (64, 128, kj.PREC_BF16X1),   # one MFMA stage, then VALU
])
def test_overflow_ladder_stages(ctx, kcap, mfma_max, prec):
    """A first-pass capacity barely above k makes most queries overflow with
    the single-product BF16 filter (its bound is ~2^-7 relative: many
    candidates tie within it); every re-run stage must still give the
    oracle's answer.  (The compensated and FP32 filters' bounds are tight
    enough that capacity 64 holds k = 60 without overflowing.)"""
    mix = datagen.mixture64_model(seed=504, d=128, comps=16)
    R = mix.sample(datagen.rng(505), 20_000)
    Q = mix.sample(datagen.rng(506), 600)
    k = 60
    idx, dist, st = ctx.join(R, Q, k, precision=prec, kernel=kj.KERNEL_MFMA, kcap=kcap, kcap_mfma_max=mfma_max)
    assert st["kcap"] == kcap
    assert st["n_overflow"] > 0, st
    assert st["rerun_stages"] >= 1
    assert_matches_oracle(R, Q, k, idx, dist)
    # the same join at the automatic capacity
    a_idx, a_dist, _ = ctx.join(R, Q, k, precision=prec, kernel=kj.KERNEL_MFMA)
    assert np.array_equal(idx, a_idx) and np.array_equal(dist, a_dist)


def test_overflow_ladder_reaches_valu_stage(ctx):
    """With the MFMA re-runs capped at the first capacity, every overflowed
    query is answered by the VALU kernel at 1024 (one stage)."""
    mix = datagen.mixture64_model(seed=507, d=96, comps=12)
    R = mix.sample(datagen.rng(508), 15_000)
    Q = mix.sample(datagen.rng(509), 400)
    idx, dist, st = ctx.join(R, Q, 64, precision=kj.PREC_BF16X1, kcap=64, kcap_mfma_max=64)
    assert st["n_overflow"] > 0 and st["rerun_stages"] == 1, st
    assert_matches_oracle(R, Q, 64, idx, dist)


def test_small_capacity_exact_mode_never_overflows(ctx):
    """Integer keys are exact: a capacity equal to k is always enough."""
    X = datagen.skin_like(8000, seed=510)
    idx, dist, st = ctx.join(X, X[:700], 64, kcap=64)
    assert st["exact_mode"] == 1 and st["n_overflow"] == 0
    assert_matches_oracle(X, X[:700], 64, idx, dist)


@pytest.mark.parametrize("prec,integer", [(kj.PREC_FP32, False), (kj.PREC_BF16, False), (kj.PREC_FP32, True)])
def test_lockstep_multiwave_tiles(ctx, prec, integer):
    """tile_waves = 4: four waves own 64 queries and walk the union of their
    cluster lists with a barrier per cluster; trailing waves of a cluster's
    last tile own fewer than 16 queries or none.  Same answer as one-wave
    tiles and as the oracle."""
    if integer:
        g = np.random.default_rng(511)
        R = g.integers(0, 16, (9000, 32)).astype(np.float32)
        Q = g.integers(0, 16, (1300, 32)).astype(np.float32)
    else:
        mix = datagen.mixture64_model(seed=512, d=64, comps=25)
        R = mix.sample(datagen.rng(513), 12_000)
        Q = mix.sample(datagen.rng(514), 1300)
    k = 24
    a_idx, a_dist, sa = ctx.join(R, Q, k, precision=prec, kernel=kj.KERNEL_MFMA, tile_waves=4, m_q=37)
    b_idx, b_dist, sb = ctx.join(R, Q, k, precision=prec, kernel=kj.KERNEL_MFMA, tile_waves=1, tile_split=1,
                                 m_q=37)
    assert sa["tile_q"] == 64 and sb["tile_q"] == 16, (sa["tile_q"], sb["tile_q"])
    assert sa["exact_mode"] == int(integer)
    assert np.array_equal(a_idx, b_idx) and np.array_equal(a_dist, b_dist)
    assert_matches_oracle(R, Q, k, a_idx, a_dist)
