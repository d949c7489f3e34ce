"""Seeded synthetic inputs for the benchmark configurations.

theta: per-element log-diffusion field on an n^dim element lattice, i.i.d.
N(0,1) from a counter-based generator (SplitMix64 + Box-Muller, one counter per
element in lattice order), then `passes` rounds of the box average over the
3^dim neighbouring elements (clipped at the domain boundary). k = exp(theta).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x + _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
    return x ^ (x >> np.uint64(31))


def _unit_open(h):
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gauss_from_counter(seed, counters):
    """Standard normal samples for the given counters (Box-Muller)."""
    counters = np.asarray(counters, np.uint64)
    base = splitmix64(np.uint64(seed)) ^ (counters * np.uint64(2))
    u1 = _unit_open(splitmix64(base))
    u2 = _unit_open(splitmix64(base ^ np.uint64(1)))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def _box3(a, axis):
    """sum over the 3-neighbourhood along one axis, clipped at the ends"""
    s = a.copy()
    sl_lo = [slice(None)] * a.ndim
    sl_hi = [slice(None)] * a.ndim
    sl_lo[axis] = slice(0, -1)
    sl_hi[axis] = slice(1, None)
    s[tuple(sl_hi)] += a[tuple(sl_lo)]
    s[tuple(sl_lo)] += a[tuple(sl_hi)]
    return s


def make_theta(n, seed, passes=3, dim=3):
    """theta[element] in lattice order (x fastest), float64, length n**dim."""
    t = gauss_from_counter(seed, np.arange(n ** dim, dtype=np.uint64)).reshape((n,) * dim)
    cnt = np.ones((n,) * dim)
    for ax in range(dim):
        cnt = _box3(cnt, ax)
    for _ in range(passes):
        s = t
        for ax in range(dim):
            s = _box3(s, ax)
        t = s / cnt
    return np.ascontiguousarray(t.reshape(-1))


def poisson_source_lattice(n, seed, sigma=0.1):
    """Source of the 2D Poisson configuration per lattice vertex (x fastest):
    2 pi^2 sin(pi x) sin(pi y) plus N(0, sigma^2) noise, the noise counter being
    the vertex's lattice index. Returns a float64 array of length (n+1)^2."""
    W = n + 1
    idx = np.arange(W * W, dtype=np.uint64)
    x = (idx % np.uint64(W)).astype(np.float64) / n
    y = (idx // np.uint64(W)).astype(np.float64) / n
    f = 2.0 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    return f + sigma * gauss_from_counter(seed, idx)
