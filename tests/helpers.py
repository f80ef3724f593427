"""Shared checks: the engine's answer against the oracle's."""
import numpy as np

import oracle


def assert_matches_oracle(R, Q, k, idx, dist, squared=False, rows=None, rtol=1e-5):
    """Indices bit-identical to the FP64 oracle; distances within rtol
    (bit-identical is expected and reported separately)."""
    R = np.ascontiguousarray(R, np.float32)
    Q = np.ascontiguousarray(Q, np.float32)
    idx = np.asarray(idx)
    dist = np.asarray(dist)
    if rows is not None:
        Q = Q[rows]
        idx = idx[rows]
        dist = dist[rows]
    ref_idx, ref_dist, ref_sq = oracle.bruteforce(R, Q, k)
    bad = np.nonzero((idx != ref_idx).any(axis=1))[0]
    assert bad.size == 0, (
        f"{bad.size} of {Q.shape[0]} rows differ; first row {bad[0]}: "
        f"got {idx[bad[0]][:12].tolist()} want {ref_idx[bad[0]][:12].tolist()}")
    want = ref_sq.astype(np.float32) if squared else ref_dist
    np.testing.assert_allclose(dist, want, rtol=rtol, atol=0)
    return int((dist != want).sum())
