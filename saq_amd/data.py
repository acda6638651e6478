"""Seeded synthetic data and IVF preparation for the SAQ workloads.

The reference prepares its inputs offline with faiss (python/pca.py, python/ivf.py):
PCA rotation of the base and query vectors, k-means centroids, and cluster ids.
faiss is not available, so this module restates those steps:

* mixture(): Gaussian mixture, centres ~ N(0, 4I), per-dimension std
  1/sqrt(1 + i/tau), uniform or Zipf(1.1) component weights.
* pca_fit(): PCA with an FP64 covariance (eigenvectors by decreasing variance),
  returning the per-dimension variance that SAQ's bit allocation consumes.
* kmeans(): Lloyd's algorithm, k-means++-free deterministic init (seeded sample).
* random_rotation(): a seeded random orthogonal matrix (Rotator::orthogonalize,
  saqlib/utils/rotator.hpp:30-35, which takes Q of a QR of a random matrix).

numpy is used for small inputs; torch (on the GPU when present) for large ones.
"""
import numpy as np


def mixture(n, d, comps, tau, seed, zipf=None, queries=0, chunk=1 << 18):
    """Returns (base n x d, queries q x d) float32, both drawn from the same mixture."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, 2.0, size=(comps, d)).astype(np.float32)
    std = (1.0 / np.sqrt(1.0 + np.arange(d) / tau)).astype(np.float32)
    if zipf:
        w = 1.0 / np.arange(1, comps + 1) ** zipf
        w = w / w.sum()
    else:
        w = np.full(comps, 1.0 / comps)
    out = []
    for total in (n, queries):
        x = np.empty((total, d), np.float32)
        for s in range(0, total, chunk):
            e = min(total, s + chunk)
            comp = rng.choice(comps, size=e - s, p=w)
            x[s:e] = centers[comp] + rng.standard_normal((e - s, d), dtype=np.float32) * std
        out.append(x)
    return out[0], out[1]


def pca_fit(x):
    """PCA with an FP64 covariance.  Returns (mean, components d x d, variance per output dim)."""
    x64 = x.astype(np.float64)
    mean = x64.mean(0)
    xc = x64 - mean
    cov = xc.T @ xc / x.shape[0]
    w, v = np.linalg.eigh(cov)
    order = np.argsort(w)[::-1]
    return mean.astype(np.float32), v[:, order].astype(np.float32), np.maximum(w[order], 0).astype(np.float32)


def pca_apply(x, mean, comps):
    return ((x - mean) @ comps).astype(np.float32)


def kmeans(x, k, iters=10, seed=0):
    """Lloyd's k-means (float32 distances, float64 centroid sums).  Returns (centroids, ids)."""
    rng = np.random.default_rng(seed)
    cent = x[rng.choice(x.shape[0], size=k, replace=False)].astype(np.float32)
    ids = None
    for _ in range(iters):
        ids = assign(x, cent)
        sums = np.zeros((k, x.shape[1]), np.float64)
        np.add.at(sums, ids, x)
        cnt = np.bincount(ids, minlength=k)
        nz = cnt > 0
        cent[nz] = (sums[nz] / cnt[nz, None]).astype(np.float32)
    ids = assign(x, cent)
    return cent, ids.astype(np.uint32)


def assign(x, cent, chunk=1 << 15):
    cn = (cent.astype(np.float32) ** 2).sum(1)
    out = np.empty(x.shape[0], np.int64)
    for s in range(0, x.shape[0], chunk):
        xs = x[s:s + chunk]
        d = cn[None, :] - 2.0 * (xs @ cent.T)
        out[s:s + chunk] = np.argmin(d, axis=1)
    return out


def random_rotation(d, seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    q = q * np.sign(np.diag(r))[None, :]
    return np.ascontiguousarray(q.T.astype(np.float32))


def exact_topk(base, queries, k, chunk=4096):
    """Exact FP64 L2 ground truth (ids, distances)."""
    b64 = base.astype(np.float64)
    bn = (b64 ** 2).sum(1)
    ids = np.empty((queries.shape[0], k), np.int64)
    dist = np.empty((queries.shape[0], k), np.float64)
    for s in range(0, queries.shape[0], chunk):
        q = queries[s:s + chunk].astype(np.float64)
        d = bn[None, :] - 2.0 * q @ b64.T + (q ** 2).sum(1)[:, None]
        part = np.argpartition(d, k, axis=1)[:, :k]
        pd = np.take_along_axis(d, part, 1)
        o = np.argsort(pd, axis=1)
        ids[s:s + chunk] = np.take_along_axis(part, o, 1)
        dist[s:s + chunk] = np.take_along_axis(pd, o, 1)
    return ids, dist


def recall(found, gt, k):
    hit = 0
    for f, g in zip(found, gt):
        hit += len(set(f[:k].tolist()) & set(g[:k].tolist()))
    return hit / (len(gt) * k)
