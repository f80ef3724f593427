"""GPU parity of the batched join (knnj_join_batch): every instance's result
is bit-identical to the same join run alone (knnj_join) and to the FP64
oracle, whatever the mix of sizes, dimensions, k, kernels and precisions
that share the fused launches."""
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


def _host(t):
    return t.cpu().numpy() if hasattr(t, "cpu") else np.asarray(t)


def _assert_same(a, b, what):
    a, b = _host(a), _host(b)
    assert a.shape == b.shape, what
    if a.size == 0:
        return
    bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[0]}"


def test_batch_uneven_mixture_matches_alone_and_oracle(ctx):
    """Small version of the batched workload: uneven |R_i|, |Q_i| = |R_i|/4,
    d_i in {16, 32, 96}, 8-64 Gaussians each, k = 16."""
    inst = datagen.batch_instances(n_inst=10, seed=91, r_lo=300, r_hi=12000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    res, st = ctx.join_batch(dev)
    assert st["n_launches"] <= 3  # one per kernel instantiation in the mix
    for i, ((R, Q), (idx, dist), (Rd, Qd, k)) in enumerate(zip(inst, res, dev)):
        a_idx, a_dist, _ = ctx.join(Rd, Qd, k)
        _assert_same(idx, a_idx, f"instance {i} idx (d={R.shape[1]})")
        _assert_same(dist, a_dist, f"instance {i} dist")
        assert_matches_oracle(R, Q, k, _host(idx), _host(dist))


def test_batch_mixed_k_kernels_precisions_and_hosts(ctx):
    """Different k, integer and real inputs, VALU / MFMA-FP32 / BF16 filter,
    a self-join, host and device buffers, an empty query set."""
    X = datagen.skin_like(3000, seed=92)
    m96 = datagen.mixture64_model(seed=93, d=96, comps=20)
    R96, Q96 = m96.sample(datagen.rng(1), 5000), m96.sample(datagen.rng(2), 700)
    m32 = datagen.mixture64_model(seed=94, d=32, comps=12)
    R32, Q32 = m32.sample(datagen.rng(3), 4000), m32.sample(datagen.rng(4), 333)
    m5 = datagen.mixture64_model(seed=95, d=5, comps=9)
    R5, Q5 = m5.sample(datagen.rng(5), 2000), m5.sample(datagen.rng(6), 100)
    Xd = _dev(X)
    cases = [
        (Xd, Xd, 200),                      # integer self-join, k = 200
        (R96, Q96, 100),                    # host, BF16 filter, k = 100
        (_dev(R96), _dev(Q96), 1),          # device, k = 1
        (_dev(R32), _dev(Q32), 16),         # MFMA FP32
        (R5, Q5, 37),                       # VALU, real
        (_dev(R5), _dev(Q5[:0]), 3),        # empty Q
        (_dev(R32), _dev(Q32), 600),        # k above the MFMA capacity: VALU at 1024
    ]
    res, st = ctx.join_batch(cases)
    oracle_inputs = [(X, X), (R96, Q96), (R96, Q96), (R32, Q32), (R5, Q5), (R5, Q5[:0]), (R32, Q32)]
    for i, (case, (idx, dist), (R, Q)) in enumerate(zip(cases, res, oracle_inputs)):
        a_idx, a_dist, _ = ctx.join(*case)
        _assert_same(idx, a_idx, f"case {i} idx")
        _assert_same(dist, a_dist, f"case {i} dist")
        if Q.shape[0]:
            assert_matches_oracle(R, Q, case[2], _host(idx), _host(dist))


def test_batch_repeated_calls_reuse_workspace(ctx):
    """Two batches in a row, the second smaller: no state leaks between calls."""
    big = datagen.batch_instances(n_inst=6, seed=96, r_lo=2000, r_hi=9000)
    small = datagen.batch_instances(n_inst=3, seed=97, r_lo=300, r_hi=1500)
    for inst in (big, small):
        res, _ = ctx.join_batch([(_dev(R), _dev(Q), 16) for R, Q in inst])
        for (R, Q), (idx, dist) in zip(inst, res):
            assert_matches_oracle(R, Q, 16, _host(idx), _host(dist))


def test_batch_errors_name_the_instance(ctx):
    R = np.zeros((10, 4), np.float32)
    with pytest.raises(kj.KnnjError) as e:
        ctx.join_batch([(R, R, 3), (R, R, 11)])
    assert e.value.status == kj.KNNJ_EINVAL and "instance 1" in str(e.value)
    bad = R.copy()
    bad[3, 2] = np.nan
    with pytest.raises(kj.KnnjError) as e:
        ctx.join_batch([(R, R, 3), (bad, R, 2)])
    assert "instance 1" in str(e.value) and "non-finite" in str(e.value)
    res, st = ctx.join_batch([])
    assert res == [] and st["n_dist_evals"] == 0


def test_batch_graph_replays_match_eager(ctx):
    """The captured plan replays bit-identically: three replays, with eager
    work and synchronisations between them, each equal to the eager batch;
    a changed plan is captured anew."""
    import torch
    inst = datagen.batch_instances(n_inst=8, seed=98, r_lo=500, r_hi=8000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    eager, st0 = ctx.join_batch(dev)
    assert st0["graph"] == 0
    outs, st1 = ctx.join_batch(dev, graph=True)
    assert st1["graph"] == 1
    X = _dev(inst[0][0])
    for rep in range(3):
        for o in outs:
            o[0].fill_(-1)
            o[1].fill_(-1.0)
        ctx.join(X, X, 5)  # eager work between replays
        torch.cuda.synchronize()
        _, st = ctx.join_batch(dev, outputs=outs, graph=True)
        assert st["graph"] == 2, rep
        for i, ((a_i, a_d), (e_i, e_d)) in enumerate(zip(outs, eager)):
            _assert_same(a_i, e_i, f"replay {rep} instance {i} idx")
            _assert_same(a_d, e_d, f"replay {rep} instance {i} dist")
    # a different plan (one instance fewer) is captured again and still exact
    res, st = ctx.join_batch(dev[:-1], outputs=outs[:-1], graph=True)
    assert st["graph"] == 1
    for (R, Q), (idx, dist) in zip(inst[:-1], res):
        assert_matches_oracle(R, Q, 16, _host(idx), _host(dist))


def test_batch_autotune_picks_and_reuses_an_exact_plan(ctx):
    """The batch autotuner times its candidates (Lloyd step, VALU tile, then
    the doubled k-select capacity and the lockstep MFMA tile) on the first
    call of a batch shape and replays the choice on later calls; every
    candidate writes the same exact answer, equal to the untuned batch."""
    inst = datagen.batch_instances(n_inst=9, seed=99, r_lo=400, r_hi=6000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    plain, _ = ctx.join_batch(dev)
    first, st1 = ctx.join_batch(dev, autotune=True)
    again, st2 = ctx.join_batch(dev, autotune=True)
    assert st1["tuned"] >= 1 and st2["tuned"] == st1["tuned"]
    for i, (p, a, b) in enumerate(zip(plain, first, again)):
        _assert_same(a[0], p[0], f"tuned instance {i} idx")
        _assert_same(a[1], p[1], f"tuned instance {i} dist")
        _assert_same(b[0], p[0], f"replayed choice, instance {i} idx")
    for (R, Q), (idx, dist) in zip(inst, again):
        assert_matches_oracle(R, Q, 16, _host(idx), _host(dist))


@pytest.mark.parametrize("tile_waves", [0, 4])
def test_batch_work_queue_matches_one_workgroup_per_slot(ctx, tile_waves):
    """The persistent launches (opts.work_queue: as many single-wave
    workgroups as the device holds, taking tile slots from eight per-XCD
    device counters) give the answer of the launches with one workgroup per
    slot, bit for bit, and the oracle's: uneven real instances of three
    widths, an integer self-join, twice in a row (the counters start from
    zero each call), and with multi-wave MFMA tiles (tile_waves=4), which
    keep one workgroup per slot."""
    inst = datagen.batch_instances(n_inst=9, seed=120 + tile_waves, r_lo=400, r_hi=15000)
    X = datagen.skin_like(4000, seed=121)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst] + [(_dev(X), _dev(X), 50)]
    opts = {"tile_waves": tile_waves} if tile_waves else {}
    a1, sa = ctx.join_batch(dev, work_queue=True, **opts)
    a2, _ = ctx.join_batch(dev, work_queue=True, **opts)
    b, sb = ctx.join_batch(dev, **opts)
    assert sa["n_launches"] == sb["n_launches"]
    assert sb["queue_launches"] == 0
    assert 1 <= sa["queue_launches"] <= sa["n_launches"]
    if tile_waves:
        assert sa["queue_launches"] < sa["n_launches"]  # the multi-wave MFMA launches keep their slots
    for i, ((x1, d1), (x2, d2), (y, e)) in enumerate(zip(a1, a2, b)):
        _assert_same(x1, y, f"instance {i} idx")
        _assert_same(d1, e, f"instance {i} dist")
        _assert_same(x2, y, f"instance {i} idx, second call")
    for (R, Q), (idx, dist) in zip(inst + [(X, X)], a1):
        k = 50 if R is X else 16
        assert_matches_oracle(R, Q, k, _host(idx), _host(dist))


def test_batch_work_queue_graph_replays(ctx):
    """The persistent form captured as a HIP graph: its counters are cleared
    by a fill kernel inside the graph, so every replay (with eager work and
    synchronisations between them) equals the eager batch."""
    import torch
    inst = datagen.batch_instances(n_inst=7, seed=130, r_lo=500, r_hi=9000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    eager, _ = ctx.join_batch(dev)
    outs, st1 = ctx.join_batch(dev, graph=True, work_queue=True)
    assert st1["graph"] == 1 and st1["queue_launches"] >= 1
    X = _dev(inst[0][0])
    for rep in range(3):
        for o in outs:
            o[0].fill_(-1)
            o[1].fill_(-1.0)
        ctx.join(X, X, 5)
        torch.cuda.synchronize()
        _, st = ctx.join_batch(dev, outputs=outs, graph=True, work_queue=True)
        assert st["graph"] == 2, rep
        for i, ((a_i, a_d), (e_i, e_d)) in enumerate(zip(outs, eager)):
            _assert_same(a_i, e_i, f"replay {rep} instance {i} idx")
            _assert_same(a_d, e_d, f"replay {rep} instance {i} dist")


def test_batch_rep_scale_matches_default(ctx):
    """rep_scale multiplies each instance's default representative count
    (2 sqrt(n), at most 4096): the answers stay the default's, bit for bit,
    and the oracle's; a single join takes the option too."""
    inst = datagen.batch_instances(n_inst=6, seed=140, r_lo=800, r_hi=20000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    base, _ = ctx.join_batch(dev)
    for sc in (2, 4):
        res, _ = ctx.join_batch(dev, rep_scale=sc)
        for i, ((x, d), (y, e)) in enumerate(zip(res, base)):
            _assert_same(x, y, f"rep_scale {sc} instance {i} idx")
            _assert_same(d, e, f"rep_scale {sc} instance {i} dist")
    R, Q = inst[-1]
    idx, dist, st = ctx.join(R, Q, 16, rep_scale=4)
    assert st["m_r"] == min(4096, R.shape[0], 4 * int(round(2 * np.sqrt(R.shape[0]))))
    assert_matches_oracle(R, Q, 16, idx, dist)
    with pytest.raises(kj.KnnjError):
        ctx.join(R, Q, 16, rep_scale=9)


def test_batch_bf16_filter_on_every_instance(ctx):
    """The BF16 hi/lo filter forced on every instance (the batch autotuner's
    precision candidate), d = 16 and 32 included: the oracle's answer."""
    inst = datagen.batch_instances(n_inst=8, seed=150, r_lo=600, r_hi=15000)
    dev = [(_dev(R), _dev(Q), 16) for R, Q in inst]
    res, st = ctx.join_batch(dev, precision=kj.PREC_BF16)
    base, _ = ctx.join_batch(dev)
    for i, ((R, Q), (idx, dist), (bi, bd)) in enumerate(zip(inst, res, base)):
        _assert_same(idx, bi, f"instance {i} idx (d={R.shape[1]})")
        assert_matches_oracle(R, Q, 16, _host(idx), _host(dist))
