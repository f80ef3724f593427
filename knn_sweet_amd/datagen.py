"""Seeded synthetic inputs for the five join workloads.

Every generator draws from ``numpy.random.Generator(numpy.random.Philox(seed))``
(Philox4x32-10, Salmon et al. 2011) so the arrays are reproducible on any
host; ``sha256`` records a digest of each array.  The default seeds are 2015
(the reference's seed, include/common8.h:94) plus the workload number.

Workloads (sizes given by the defaults; every function takes smaller sizes
for tests):
  1. ``skin_like``: d=3 integer coordinates 0..255 stored as FP32, a mixture
     of 16 isotropic Gaussians (means U[0,255]^3, sigma U[8,40]) rounded and
     clipped, self-join of 245,057 points, k=200.
  2. ``mixture64``: d=64, 256 Gaussians, means N(0, 4^2), sigma log-U[0.25,2],
     |R|=1,000,000, |Q|=100,000 independent draws, k=32.
  3. ``mixture128``: d=128, 1,024 Gaussians, means N(0,1), sigma
     log-U[0.05,0.5], rows scaled by one global factor (the reciprocal of the
     median over components of the expected row norm) and each row's norm
     then clamped into [0.5, 2]; |R|=2,000,000, |Q|=200,000, k=10 and 100.
  4. ``batch_instances``: 512 independent joins, |R_i| log-U[1e3, 2e5],
     |Q_i| = |R_i|/4, d_i from {16, 32, 96}, 8-64 Gaussians each, k=16.
  5. ``serving_set`` / ``serving_batches``: resident 4,000,000 x 96 set from
     2,048 Gaussians (means N(0,1), sigma log-U[0.05,0.5]); query batches of
     log-U[1, 4096] points from the same mixture, k=50.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@dataclass
class Mixture:
    means: np.ndarray   # [c, d] float64
    sigma: np.ndarray   # [c] float64
    scale: float = 1.0  # global row scale applied after sampling
    clamp: tuple | None = None  # (lo, hi) row-norm clamp after scaling

    def sample(self, g: np.random.Generator, n: int, chunk: int = 1 << 20) -> np.ndarray:
        c, d = self.means.shape
        out = np.empty((n, d), np.float32)
        for s in range(0, n, chunk):
            m = min(chunk, n - s)
            comp = g.integers(0, c, size=m)
            x = g.standard_normal((m, d))
            x *= self.sigma[comp, None]
            x += self.means[comp]
            if self.scale != 1.0:
                x *= self.scale
            if self.clamp is not None:
                lo, hi = self.clamp
                nrm = np.sqrt(np.einsum("ij,ij->i", x, x))
                f = np.ones_like(nrm)
                small = (nrm < lo) & (nrm > 0)
                f[small] = lo / nrm[small]
                big = nrm > hi
                f[big] = hi / nrm[big]
                x *= f[:, None]
            out[s:s + m] = x
        return out


def _loguniform(g, lo, hi, size):
    return np.exp(g.uniform(np.log(lo), np.log(hi), size=size))


def skin_like(n: int = 245_057, seed: int = 2016) -> np.ndarray:
    """Workload 1: integer-valued d=3 points (self-join: R = Q = the result)."""
    g = rng(seed)
    means = g.uniform(0.0, 255.0, size=(16, 3))
    sigma = g.uniform(8.0, 40.0, size=16)
    comp = g.integers(0, 16, size=n)
    x = means[comp] + sigma[comp, None] * g.standard_normal((n, 3))
    return np.clip(np.rint(x), 0, 255).astype(np.float32)


def mixture64_model(seed: int = 2017, d: int = 64, comps: int = 256) -> Mixture:
    g = rng(seed)
    return Mixture(means=4.0 * g.standard_normal((comps, d)), sigma=_loguniform(g, 0.25, 2.0, comps))


def mixture64(n_r: int = 1_000_000, n_q: int = 100_000, seed: int = 2017, d: int = 64):
    """Workload 2: (R, Q) FP32, d=64."""
    mix = mixture64_model(seed, d)
    return mix.sample(rng(seed * 7 + 1), n_r), mix.sample(rng(seed * 7 + 2), n_q)


def mixture128_model(seed: int = 2018, d: int = 128, comps: int = 1024) -> Mixture:
    g = rng(seed)
    means = g.standard_normal((comps, d))
    sigma = _loguniform(g, 0.05, 0.5, comps)
    expected = np.sqrt(np.einsum("ij,ij->i", means, means) + d * sigma**2)
    return Mixture(means=means, sigma=sigma, scale=float(1.0 / np.median(expected)), clamp=(0.5, 2.0))


def mixture128(n_r: int = 2_000_000, n_q: int = 200_000, seed: int = 2018, d: int = 128):
    """Workload 3: (R, Q) FP32, d=128, row norms in [0.5, 2]."""
    mix = mixture128_model(seed, d)
    return mix.sample(rng(seed * 7 + 1), n_r), mix.sample(rng(seed * 7 + 2), n_q)


def batch_instances(n_inst: int = 512, seed: int = 2019, r_lo: float = 1e3, r_hi: float = 2e5,
                    dims=(16, 32, 96)):
    """Workload 4: list of (R_i, Q_i) pairs of uneven sizes and dimensions."""
    g = rng(seed)
    out = []
    for i in range(n_inst):
        n_r = int(np.floor(_loguniform(g, r_lo, r_hi, 1)[0]))
        n_q = max(1, n_r // 4)
        d = int(dims[g.integers(0, len(dims))])
        comps = int(g.integers(8, 65))
        mix = Mixture(means=4.0 * g.standard_normal((comps, d)), sigma=_loguniform(g, 0.25, 2.0, comps))
        sub = rng(seed * 1000 + i)
        out.append((mix.sample(sub, n_r), mix.sample(sub, n_q)))
    return out


def serving_model(seed: int = 2020, d: int = 96, comps: int = 2048) -> Mixture:
    g = rng(seed)
    return Mixture(means=g.standard_normal((comps, d)), sigma=_loguniform(g, 0.05, 0.5, comps))


def serving_set(n: int = 4_000_000, seed: int = 2020, d: int = 96) -> np.ndarray:
    """Workload 5: the resident reference set."""
    return serving_model(seed, d).sample(rng(seed * 7 + 1), n)


def serving_batches(n_batches: int, seed: int = 2020, d: int = 96, max_batch: int = 4096):
    """Workload 5: query batches with sizes floor(exp U[0, ln max_batch])."""
    mix = serving_model(seed, d)
    g = rng(seed * 7 + 3)
    sizes = np.floor(np.exp(g.uniform(0.0, np.log(max_batch), size=n_batches))).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    return [mix.sample(g, int(s)) for s in sizes]
