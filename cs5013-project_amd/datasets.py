"""Seeded synthetic inputs of the benchmark configurations.

integer_mixture: K-component Gaussian mixture in d dimensions, centres uniform in
[0, 128]^d, standard deviation sigma, rounded and clipped to [0, 255] and stored
as float32 (integer-valued, so every squared-L2 partial sum at d <= 128 is an
exact float). The component centres come from `seed`; `stream` selects an
independent draw of points from the same mixture (0 = base, 1 = queries,
2.. = insert batches).

unit_mixture: K-component mixture with centres drawn on the unit sphere,
sigma applied before normalising each point to unit length.
"""
import numpy as np


def _rng(seed, stream):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, stream])))


def integer_mixture(n, d=128, k=64, sigma=20.0, seed=1234, stream=0, chunk=1 << 17):
    centres = _rng(seed, 1000).uniform(0.0, 128.0, size=(k, d)).astype(np.float32)
    rng = _rng(seed, stream)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, k, size=e - s)
        x = centres[comp] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        np.clip(np.rint(x, out=x), 0.0, 255.0, out=x)
        out[s:e] = x
    return out


def unit_mixture(n, d=96, k=1024, sigma=0.1, seed=1234, stream=0, chunk=1 << 17):
    c = _rng(seed, 1000).standard_normal((k, d)).astype(np.float32)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    rng = _rng(seed, stream)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, k, size=e - s)
        x = c[comp] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x
    return out


# Benchmark configurations (sizes and parameters).
CONFIG1 = dict(n=10_000, d=128, k=64, nq=100, R=32, L_build=64, alphas=(1.2, 1.2), k_nn=10,
               L_search=(10, 20, 40, 80))
CONFIG2 = dict(n=1_000_000, d=128, k=256, nq=10_000, R=64, L_build=128, alphas=(1.0, 1.2),
               k_nn=10, L_search=(10, 20, 40, 60, 80, 100, 120, 150, 200))
