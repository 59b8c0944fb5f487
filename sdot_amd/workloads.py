"""Synthetic inputs of the benchmark workloads (splitmix64 -> xoshiro256** streams)."""
import numpy as np

from . import rng_uniform


def uniform_diracs(n, dim, seed=0):
    """n diracs uniform in [0,1)^dim, drawn dirac-major (x0, y0[, z0], x1, ...)."""
    return rng_uniform(seed, n * dim).reshape(n, dim).copy()


def lognormal_image(size, dim=3, seed=1, sigma=0.5):
    """Piecewise-constant density exp(sigma z), z ~ N(0,1) (Box-Muller on the
    seeded stream), x fastest, normalised to mean 1 (total mass 1 on the unit box)."""
    m = size ** dim
    u = rng_uniform(seed, 2 * m)
    z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])
    img = np.exp(sigma * z).reshape((size,) * dim)
    return img / img.mean()


def serving_schedule(rate, duration, seed=7, frac_2d=0.7, lo=1e3, hi=2e6):
    """Open-loop serving load: Poisson arrivals at `rate` per second over `duration`
    seconds; each request 2D with probability frac_2d (else 3D), N = floor(exp(U[ln lo,
    ln hi])).  Returns (arrival times, dims, sizes); request k's diracs are
    uniform_diracs(sizes[k], dims[k], seed=10000 + k)."""
    m = int(rate * duration * 2 + 64)
    u = rng_uniform(seed, 3 * m).reshape(m, 3)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    keep = t < duration
    dims = np.where(u[:, 1] < frac_2d, 2, 3)
    sizes = np.floor(np.exp(np.log(lo) + u[:, 2] * (np.log(hi) - np.log(lo)))).astype(np.int64)
    return t[keep], dims[keep], sizes[keep]


def batch_sizes(count, lo=1e3, hi=1e5, seed=0):
    """Problem sizes drawn log-uniform in [lo, hi] (one value per problem)."""
    u = rng_uniform(seed, count)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
