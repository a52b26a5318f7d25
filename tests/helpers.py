"""Shared test helpers: seeded ray batches and image statistics."""
import numpy as np

from radiantium_amd import abi


def random_rays(n, seed, center=(0.0, 0.5, 0.0), radius=2.0, target_radius=0.6, tmin=1e-4, tmax=3.4e38):
    """Rays from a sphere of `radius` around `center` toward random points
    within `target_radius` of it (so most rays hit the scene)."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    o = rng.normal(size=(n, 3))
    o = c + radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    t = c + target_radius * rng.uniform(-1, 1, size=(n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, abi.ray_dtype())
    r["o"] = o.astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def edge_rays(desc, n, seed):
    """Rays aimed exactly at shared triangle edges and vertices of the scene:
    the tie cases of closest-hit (first-visited primitive wins)."""
    a = desc.arrays()
    pos, idx = a["positions"], a["indices"]
    rng = np.random.default_rng(seed)
    tri = rng.integers(0, idx.shape[0], size=n)
    k = rng.integers(0, 3, size=n)
    p0 = pos[idx[tri, k]].astype(np.float64)
    p1 = pos[idx[tri, (k + 1) % 3]].astype(np.float64)
    w = rng.uniform(0, 1, size=(n, 1))
    w[rng.uniform(size=n) < 0.2] = 0.0   # exact vertices
    tgt = p0 * (1 - w) + p1 * w
    o = tgt + rng.normal(size=(n, 3)) * 1.5
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, abi.ray_dtype())
    r["o"] = o.astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-4
    r["tmax"] = 3.4e38
    return r


def z_test(mean_a, sum2_a, n_a, mean_b, sum2_b, n_b, floor=1e-6):
    """Per-channel z scores of two independent per-pixel means."""
    var_a = np.maximum(sum2_a / n_a - mean_a.astype(np.float64) ** 2, 0) / n_a
    var_b = np.maximum(sum2_b / n_b - mean_b.astype(np.float64) ** 2, 0) / n_b
    return (mean_a - mean_b) / np.sqrt(var_a + var_b + floor)


def rel_mse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))
