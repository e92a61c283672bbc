"""Seeded synthetic datasets for the benchmark workloads and the parity tests.

All generators are numpy (PCG64 ``default_rng``) so a given seed yields the
same float32 bytes on every machine with this numpy; the GPU benchmark uses the
torch variants, which generate on the device.

* ``mixture``: Gaussian mixture, ``n_centers`` centers drawn uniformly on the
  unit sphere plus isotropic noise of standard deviation ``sigma``
  (the 100k x 128 flat workload: 256 centers, sigma 0.25, not normalized).
* ``anisotropic``: x_d ~ N(0, (d+1)^-0.8), then L2-normalized
  (the 1M x 1536 embedding-scale workload).
* ``zipf_mixture``: centers uniform on the unit sphere, cluster sizes by a
  Zipf rank law (cluster r drawn with probability r^-s / H), isotropic noise
  of RMS norm ``sigma`` (each coordinate sigma / sqrt(dim)): the bucketed
  workload (10M x 768, 20,000 centers, s = 1.1, sigma 0.2). With a
  per-coordinate sigma of 0.2 the 768-d noise would have norm 5.5 against
  unit-norm centers and no bucket structure would survive.
"""
from __future__ import annotations

import numpy as np


def sphere_centers(n_centers: int, dim: int, seed: int) -> np.ndarray:
    r = np.random.default_rng(seed)
    c = r.standard_normal((n_centers, dim))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    return c


def mixture(n: int, dim: int, n_centers: int, sigma: float, center_seed: int, seed: int) -> np.ndarray:
    c = sphere_centers(n_centers, dim, center_seed)
    r = np.random.default_rng(seed)
    z = r.integers(0, n_centers, n)
    return (c[z] + sigma * r.standard_normal((n, dim))).astype(np.float32)


def zipf_weights(n_centers: int, s: float) -> np.ndarray:
    w = np.arange(1, n_centers + 1, dtype=np.float64) ** -s
    return w / w.sum()


def zipf_mixture(n: int, dim: int, n_centers: int, zipf_s: float, sigma: float, center_seed: int,
                 seed: int) -> np.ndarray:
    c = sphere_centers(n_centers, dim, center_seed)
    r = np.random.default_rng(seed)
    z = r.choice(n_centers, size=n, p=zipf_weights(n_centers, zipf_s))
    return (c[z] + (sigma / np.sqrt(dim)) * r.standard_normal((n, dim))).astype(np.float32)


def anisotropic(n: int, dim: int, seed: int, normalize: bool = True) -> np.ndarray:
    r = np.random.default_rng(seed)
    std = (np.arange(1, dim + 1, dtype=np.float64)) ** -0.4
    x = r.standard_normal((n, dim)) * std
    if normalize:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x.astype(np.float32)


def flat100k_db(n: int = 100_000) -> np.ndarray:
    """Flat CPU-runnable workload: 100k x 128, 256-center mixture, sigma 0.25."""
    return mixture(n, 128, 256, 0.25, 1001, 1002)


def flat100k_queries(nq: int = 1000) -> np.ndarray:
    return mixture(nq, 128, 256, 0.25, 1001, 1003)


# ---- device generators (torch; the benchmark's embedding-scale data) -------
def anisotropic_torch(n: int, dim: int, seed: int, device="cuda", chunk: int = 1 << 16):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    std = torch.arange(1, dim + 1, dtype=torch.float32, device=device).pow(-0.4)
    out = torch.empty((n, dim), dtype=torch.float32, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = torch.randn((e - s, dim), generator=g, device=device, dtype=torch.float32) * std
        out[s:e] = x / x.norm(dim=1, keepdim=True)
    return out


def zipf_mixture_torch(n: int, dim: int, n_centers: int, zipf_s: float, sigma: float, center_seed: int, seed: int,
                       device="cuda", chunk: int = 1 << 18):
    """zipf_mixture generated on the device (torch generators: the same seeds
    give the same rows on every run with this torch, not the numpy rows)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(center_seed)
    c = torch.randn((n_centers, dim), generator=g, device=device, dtype=torch.float32)
    c = c / c.norm(dim=1, keepdim=True)
    w = torch.from_numpy(zipf_weights(n_centers, zipf_s)).to(device=device, dtype=torch.float64)
    g.manual_seed(seed)
    out = torch.empty((n, dim), dtype=torch.float32, device=device)
    scale = sigma / float(np.sqrt(dim))
    for s0 in range(0, n, chunk):
        e = min(n, s0 + chunk)
        z = torch.multinomial(w, e - s0, replacement=True, generator=g)
        out[s0:e] = c[z] + scale * torch.randn((e - s0, dim), generator=g, device=device, dtype=torch.float32)
    return out
