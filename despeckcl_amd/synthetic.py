"""Seeded synthetic interferograms for tests and benchmarks.

The speckle model is the reference's own InSAR simulator
(src/compute/nlinsar/insarsim_simu.cpp:34-60):

    z1 = sqrt(R) * n1
    z2 = sqrt(R) * (gamma * exp(i*phi) * n1 + sqrt(1 - gamma**2) * n2),  n1, n2 ~ CN(0, 1)
    a1 = |z1|,  a2 = |z2|,  phase = arg(z1 * conj(z2))

Reflectivity R is piecewise constant over random rectangles (log-uniform values), the
coherence gamma is uniform per rectangle, and the true phase is a linear ramp along the
columns plus a Gaussian bump.  polsar() draws dim-channel scattering vectors
k = chol(Sigma) n, n ~ CN(0, I), with a random Hermitian covariance Sigma per rectangle, and
returns the single-look covariance matrices k k^H.  Everything is float32 and drawn from numpy's PCG64 with the
given seed, so a (shape, seed) pair always yields the same image.
"""
from __future__ import annotations

import numpy as np

__all__ = ["interferogram", "homogeneous_area", "polsar", "tile_shapes", "served_stream", "SEED_BASE"]

SEED_BASE = 20260


def _regions(rng, height, width, n_regions, refl_range, coh_range):
    lo, hi = np.log(refl_range[0]), np.log(refl_range[1])
    refl = np.full((height, width), np.exp(rng.uniform(lo, hi)), dtype=np.float64)
    coh = np.full((height, width), rng.uniform(*coh_range), dtype=np.float64)
    for _ in range(n_regions):
        h0, h1 = np.sort(rng.integers(0, height + 1, size=2))
        w0, w1 = np.sort(rng.integers(0, width + 1, size=2))
        if h1 - h0 < 2:
            h1 = min(height, h0 + max(2, height // 8))
        if w1 - w0 < 2:
            w1 = min(width, w0 + max(2, width // 8))
        refl[h0:h1, w0:w1] = np.exp(rng.uniform(lo, hi))
        coh[h0:h1, w0:w1] = rng.uniform(*coh_range)
    return refl, coh


def interferogram(height, width, seed, n_regions=8, coh_range=(0.3, 0.9),
                  refl_range=(1.0, 100.0), ramp=0.05, bump=np.pi,
                  scatterer_fraction=0.0, scatterer_gain=50.0):
    """Return (ampl_master, ampl_slave, phase) as C-contiguous float32 arrays."""
    rng = np.random.Generator(np.random.PCG64(seed))
    refl, coh = _regions(rng, height, width, n_regions, refl_range, coh_range)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    cy, cx = rng.uniform(0.25, 0.75) * height, rng.uniform(0.25, 0.75) * width
    sigma = max(4.0, min(height, width) / 6.0)
    phi = ramp * xx + bump * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2.0 * sigma ** 2))
    n1 = (rng.standard_normal((height, width)) + 1j * rng.standard_normal((height, width))) / np.sqrt(2)
    n2 = (rng.standard_normal((height, width)) + 1j * rng.standard_normal((height, width))) / np.sqrt(2)
    sr = np.sqrt(refl)
    z1 = sr * n1
    z2 = sr * (coh * np.exp(1j * phi) * n1 + np.sqrt(1.0 - coh ** 2) * n2)
    if scatterer_fraction > 0:
        mask = rng.random((height, width)) < scatterer_fraction
        z1 = np.where(mask, z1 * scatterer_gain, z1)
        z2 = np.where(mask, z2 * scatterer_gain, z2)
    a1 = np.abs(z1).astype(np.float32)
    a2 = np.abs(z2).astype(np.float32)
    ph = np.angle(z1 * np.conj(z2)).astype(np.float32)
    return np.ascontiguousarray(a1), np.ascontiguousarray(a2), np.ascontiguousarray(ph)


def homogeneous_area(size, seed, reflectivity=10.0, coherence=0.6, phase=0.5):
    """Constant-statistics area (the NL-SAR training input): same model, one region."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n1 = (rng.standard_normal((size, size)) + 1j * rng.standard_normal((size, size))) / np.sqrt(2)
    n2 = (rng.standard_normal((size, size)) + 1j * rng.standard_normal((size, size))) / np.sqrt(2)
    sr = np.sqrt(reflectivity)
    z1 = sr * n1
    z2 = sr * (coherence * np.exp(1j * phase) * n1 + np.sqrt(1.0 - coherence ** 2) * n2)
    return (np.ascontiguousarray(np.abs(z1).astype(np.float32)),
            np.ascontiguousarray(np.abs(z2).astype(np.float32)),
            np.ascontiguousarray(np.angle(z1 * np.conj(z2)).astype(np.float32)))


def tile_shapes(n, seed, lo=256, hi=2048, max_aspect=4.0):
    """n (height, width) pairs with both sides log-uniform in [lo, hi] and aspect ratio at
    most max_aspect (pairs beyond it are redrawn)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    while len(out) < n:
        h, w = np.exp(rng.uniform(np.log(lo), np.log(hi), 2))
        h, w = int(round(h)), int(round(w))
        if max(h, w) <= max_aspect * min(h, w):
            out.append((h, w))
    return out


def polsar(height, width, seed, dim=3, n_regions=6, refl_range=(1.0, 100.0)):
    """Return single-look covariance matrices, complex64 of shape (dim, dim, height, width),
    the layout the covariance NL-SAR entry point takes (before interlacing)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lab = np.zeros((height, width), np.int64)
    for r in range(1, n_regions + 1):
        h0, h1 = np.sort(rng.integers(0, height + 1, size=2))
        w0, w1 = np.sort(rng.integers(0, width + 1, size=2))
        lab[h0:max(h1, h0 + 2), w0:max(w1, w0 + 2)] = r
    lo, hi = np.log(refl_range[0]), np.log(refl_range[1])
    chols = []
    for _ in range(n_regions + 1):
        g = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
        sig = g @ g.conj().T + 0.2 * np.eye(dim)
        sig *= np.exp(rng.uniform(lo, hi)) / np.real(np.trace(sig)) * dim
        chols.append(np.linalg.cholesky(sig))
    n = (rng.standard_normal((dim, height, width)) + 1j * rng.standard_normal((dim, height, width))) / np.sqrt(2)
    L = np.stack(chols)[lab]                      # (H, W, dim, dim)
    k = np.einsum("hwij,jhw->ihw", L, n)          # (dim, H, W)
    cov = k[:, None] * np.conj(k[None, :])        # (dim, dim, H, W)
    return np.ascontiguousarray(cov.astype(np.complex64))


def served_stream(n, rate=40.0, scene_side=4096, n_scenes=4, lo=128, hi=4096):
    """The seeded mixed request stream: n requests, 30 % NL-InSAR (kind 1) and 70 % NL-SAR
    (kind 0), sides log-uniform in [lo, hi] (aspect <= 4:1), each a crop of one of n_scenes
    scene_side^2 interferograms (request i uses scene i % n_scenes, made by
    interferogram(scene_side, scene_side, SEED_BASE + 40 + k, n_regions=16,
    coh_range=(0.2, 0.95))), and Poisson arrival times at `rate` requests/s.  Returns
    (kinds, shapes, origins, arrivals_s)."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 4))
    kinds = (rng.random(n) < 0.3).astype(int)
    shapes = tile_shapes(n, SEED_BASE + 4, lo=lo, hi=hi)
    origins = []
    for h, w in shapes:
        y0, x0 = int(rng.integers(0, scene_side - h + 1)), int(rng.integers(0, scene_side - w + 1))
        origins.append((y0, x0))
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    return kinds, shapes, origins, arrivals


def served_scene(k, scene_side=4096):
    """Scene k of served_stream."""
    return interferogram(scene_side, scene_side, SEED_BASE + 40 + k, n_regions=16, coh_range=(0.2, 0.95))
