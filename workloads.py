"""Synthetic inputs for the benchmark and parity workloads (numpy PCG64, generated in float64
and cast to the working dtype; the GPU and the CPU oracle always consume the same arrays).

Workloads (seed = workload number - 1):

1. ``plain_cpu``   d=3 fp32, 100,000 reference points N(0,1)*1e3, 10,000 queries from the same
                   distribution (references drawn first), k=5.
2. ``plain_gpu``   d=3 fp32, 4,000,000 reference points U[0,1]^3; 16,000,000 queries: the first
                   half U[-0.1,1.1]^3, the second half from a 32-component isotropic Gaussian
                   mixture (sigma 0.02, centres U[0,1]^3, equal weights); k=16.
3. ``diff``        d=3 fp32, 1,000,000 points on a unit sphere whose radius is perturbed by
                   0.05 * sum of 8 real spherical harmonics (degree <= 6, (l, m) uniform over the
                   49 pairs, N(0,1) coefficients); queries = points + N(0, 0.01^2); k=8;
                   loss = sum(w * dists), w ~ U(0,1).
4. ``batch``       256 d=3 fp32 clouds, N_c log-uniform in [1e3, 5e5], Q_c = U(0.5,2)*N_c, each
                   cloud U[0,1]^3 or an 8-component Gaussian mixture (sigma 0.02) by its own seed
                   (1000 + c), queries from the same kind, k_c uniform in {1, 4, 16}.
5a. ``d8``         d=8 fp32, 2,000,000 points and 2,000,000 queries from a 64-component
                   anisotropic mixture: covariance R diag(v) R^T, v ~ logU(0.01, 1) (variances),
                   R from the QR of a Gaussian matrix with the sign fix; centres U[0,1]^8; k=32.
5b. ``fp64``       d=3 fp64, 500,000 points and 500,000 queries N(0,1)*1e3, k=16.

Every workload accepts smaller sizes (``n``, ``q``) for parity tests; the draws keep their order,
so a smaller instance is not a prefix of the full one.
"""
from __future__ import annotations

import math

import numpy as np


def plain_cpu(n=100_000, q=10_000, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    ref = (rng.standard_normal((n, 3)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((q, 3)) * 1e3).astype(dtype)
    return dict(ref=ref, queries=qry, k=5, name="plain_cpu")


def _gmm(rng, count, centers, sigma):
    lab = rng.integers(0, centers.shape[0], count)
    return centers[lab] + sigma * rng.standard_normal((count, centers.shape[1]))


def plain_gpu(n=4_000_000, q=16_000_000, seed=1, dtype=np.float32):
    rng = np.random.default_rng(seed)
    ref = rng.random((n, 3)).astype(dtype)
    q_uni = q // 2
    q_gmm = q - q_uni
    uni = rng.uniform(-0.1, 1.1, (q_uni, 3))
    centers = rng.random((32, 3))
    gmm = _gmm(rng, q_gmm, centers, 0.02)
    qry = np.concatenate([uni, gmm]).astype(dtype)
    return dict(ref=ref, queries=qry, k=16, name="plain_gpu", halves=(q_uni, q_gmm))


def _assoc_legendre(l, m, x):
    """P_l^m(x) for m >= 0 by the standard three-term recurrence (no Condon-Shortley phase)."""
    pmm = np.ones_like(x)
    if m > 0:
        somx2 = np.sqrt(np.maximum(0.0, 1.0 - x * x))
        fact = 1.0
        for _ in range(m):
            pmm = pmm * fact * somx2
            fact += 2.0
    if l == m:
        return pmm
    pmmp1 = x * (2 * m + 1) * pmm
    if l == m + 1:
        return pmmp1
    pll = pmmp1
    for ll in range(m + 2, l + 1):
        pll = ((2 * ll - 1) * x * pmmp1 - (ll + m - 1) * pmm) / (ll - m)
        pmm, pmmp1 = pmmp1, pll
    return pll


def real_sph_harm(l, m, theta, phi):
    """Real orthonormal spherical harmonic Y_lm(theta, phi), theta polar, phi azimuth."""
    am = abs(m)
    k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = _assoc_legendre(l, am, np.cos(theta))
    if m > 0:
        return math.sqrt(2.0) * k * np.cos(m * phi) * p
    if m < 0:
        return math.sqrt(2.0) * k * np.sin(am * phi) * p
    return k * p


def diff_sphere(n=1_000_000, seed=2, dtype=np.float32):
    rng = np.random.default_rng(seed)
    pairs = [(l, m) for l in range(7) for m in range(-l, l + 1)]
    pick = rng.integers(0, len(pairs), 8)
    coef = rng.standard_normal(8)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    rad = np.ones(n)
    for (l, m), c in zip((pairs[i] for i in pick), coef):
        rad += 0.05 * c * real_sph_harm(l, m, theta, phi)
    ref = v * rad[:, None]
    qry = ref + rng.normal(0.0, 0.01, (n, 3))
    q = qry.astype(dtype)
    w = rng.random((n, 8)).astype(dtype)
    return dict(ref=ref.astype(dtype), queries=q, k=8, weights=w, name="diff")


def batch_clouds(n_clouds=256, seed=3, n_min=1_000, n_max=500_000, dtype=np.float32):
    """List of dicts (ref, queries, k) — each cloud drawn from its own PCG64 stream 1000 + c."""
    meta = np.random.default_rng(seed)
    sizes = np.round(np.exp(meta.uniform(math.log(n_min), math.log(n_max), n_clouds))).astype(np.int64)
    qfac = meta.uniform(0.5, 2.0, n_clouds)
    ks = meta.choice(np.array([1, 4, 16]), n_clouds)
    clouds = []
    for c in range(n_clouds):
        rng = np.random.default_rng(1000 + c)
        n = int(sizes[c])
        q = int(round(qfac[c] * n))
        if rng.random() < 0.5:
            ref = rng.random((n, 3))
            qry = rng.random((q, 3))
            kind = "uniform"
        else:
            centers = rng.random((8, 3))
            ref = _gmm(rng, n, centers, 0.02)
            qry = _gmm(rng, q, centers, 0.02)
            kind = "gmm"
        clouds.append(dict(ref=ref.astype(dtype), queries=qry.astype(dtype), k=int(ks[c]), kind=kind))
    return clouds


def _random_rotation(rng, d):
    a = rng.standard_normal((d, d))
    qm, r = np.linalg.qr(a)
    return qm * np.sign(np.diag(r))[None, :]


def d8_mixture(n=2_000_000, q=2_000_000, seed=4, dtype=np.float32, comps=64):
    rng = np.random.default_rng(seed)
    d = 8
    centers = rng.random((comps, d))
    chol = []
    for _ in range(comps):
        rot = _random_rotation(rng, d)
        var = np.exp(rng.uniform(math.log(0.01), math.log(1.0), d))
        chol.append(rot * np.sqrt(var)[None, :])
    chol = np.stack(chol)

    def draw(count):
        lab = rng.integers(0, comps, count)
        z = rng.standard_normal((count, d))
        return centers[lab] + np.einsum("nij,nj->ni", chol[lab], z)

    ref = draw(n).astype(dtype)
    qry = draw(q).astype(dtype)
    return dict(ref=ref, queries=qry, k=32, name="d8")


def fp64_normal(n=500_000, q=500_000, seed=4, dtype=np.float64):
    rng = np.random.default_rng(seed)
    ref = (rng.standard_normal((n, 3)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((q, 3)) * 1e3).astype(dtype)
    return dict(ref=ref, queries=qry, k=16, name="fp64")
