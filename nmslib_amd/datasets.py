"""Seeded synthetic vector sets for the HNSW query-path workloads.

Every generator is a pure function of its seed, so the same base vectors can
be regenerated on any host bit-for-bit (numpy's PCG64 ``default_rng``).

* ``uniform``        i.i.d. U[0,1) float32 (10,000 x 128 l2 workload).
* ``halfnormal_u8``  integer coordinates 0..255 from a clipped, shifted
                     half-normal with sigma 40: ``clip(rint(|N(0,40)| - 15),
                     0, 255)``; the shift of 15 leaves P(x == 0) =
                     P(|Z| < 15.5/40) ~= 0.30 (1,000,000 x 128 l2 workload).
* ``decaying_normal`` N(0, 1/(1+0.05 i)) per dimension i, L2-normalised
                     (2,000,000 x 100 cosinesimil workload).
* ``gmm``            64-component Gaussian mixture, centres N(0,1), component
                     sigma 0.3, L2-normalised (10,000,000 x 96 negdotprod).
"""
from __future__ import annotations

import numpy as np

HALFNORMAL_SIGMA = 40.0
HALFNORMAL_SHIFT = 15.0


def uniform(n: int, dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.random((n, dim), dtype=np.float32)


def halfnormal_u8(n: int, dim: int, seed: int, chunk: int = 1 << 16) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = rng.standard_normal((e - s, dim), dtype=np.float32)
        v = np.rint(np.abs(z) * np.float32(HALFNORMAL_SIGMA) - np.float32(HALFNORMAL_SHIFT))
        out[s:e] = np.clip(v, 0.0, 255.0)
    return out


def _normalise(x: np.ndarray) -> np.ndarray:
    nrm = np.sqrt((x.astype(np.float64) ** 2).sum(axis=1, keepdims=True))
    nrm[nrm == 0] = 1.0
    return (x / nrm).astype(np.float32)


def decaying_normal(n: int, dim: int, seed: int, chunk: int = 1 << 16) -> np.ndarray:
    rng = np.random.default_rng(seed)
    scale = (1.0 / np.sqrt(1.0 + 0.05 * np.arange(dim))).astype(np.float32)
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = _normalise(rng.standard_normal((e - s, dim), dtype=np.float32) * scale)
    return out


def gmm(n: int, dim: int, seed: int, components: int = 64, sigma: float = 0.3,
        chunk: int = 1 << 16) -> np.ndarray:
    # centres depend only on the mixture (fixed seed), samples on `seed`
    centres = np.random.default_rng(12345).standard_normal((components, dim), dtype=np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, components, size=e - s)
        x = centres[comp] + np.float32(sigma) * rng.standard_normal((e - s, dim), dtype=np.float32)
        out[s:e] = _normalise(x)
    return out


# name -> (generator, n, dim, base seed, query seed, nq)
WORKLOADS = {
    "uniform10k": (uniform, 10_000, 128, 0, 1, 1_000),
    "halfnormal1m": (halfnormal_u8, 1_000_000, 128, 2, 3, 10_000),
    "decay2m": (decaying_normal, 2_000_000, 100, 4, 5, 8_192),
    "gmm10m": (gmm, 10_000_000, 96, 6, 7, 10_000),
    "gmm2m": (gmm, 2_000_000, 96, 6, 7, 10_000),
}


def base(name: str, n: int | None = None) -> np.ndarray:
    gen, n0, dim, seed, _, _ = WORKLOADS[name]
    return gen(n0 if n is None else n, dim, seed)


def queries(name: str, nq: int | None = None) -> np.ndarray:
    gen, _, dim, _, qseed, nq0 = WORKLOADS[name]
    return gen(nq0 if nq is None else nq, dim, qseed)
