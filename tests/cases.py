"""Seeded synthetic inputs shared by the parity tests and bench.py.

Arrays are Fortran order (x fastest), data float32 with -99 at uninformed
cells, mask uint8.  The generators follow the workload descriptions used
for measurement (grid, informed fraction, value law, covariance), at any
size.
"""
import numpy as np


class Grid:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    @property
    def shape(self):
        return (self.x, self.y, self.z)

    @property
    def size(self):
        return self.x * self.y * self.z


class Cov:
    """Duck-typed CovarianceModel (type, ranges, angles, sill, nugget)."""

    def __init__(self, type, ranges, angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.0):
        self.type, self.ranges, self.angles, self.sill, self.nugget = type, tuple(ranges), tuple(angles), sill, nugget


def scattered(shape, frac, seed, law="lognormal"):
    """A property with round(frac * N) informed cells chosen uniformly."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    k = max(1, int(round(frac * n)))
    idx = rng.choice(n, k, replace=False)
    mask = np.zeros(n, np.uint8)
    mask[idx] = 1
    if law == "lognormal":
        vals = rng.lognormal(0.0, 1.0, k)
    else:
        vals = rng.standard_normal(k)
    data = np.full(n, -99.0, np.float32)
    data[idx] = vals.astype(np.float32)
    return data.reshape(shape, order="F"), mask.reshape(shape, order="F")


def categories(data, mask, quantiles=(0.1, 0.3, 0.5, 0.7, 0.9), k=5):
    """Categorical property from thresholds at data quantiles: category =
    min(searchsorted(t, v), k - 1); marginals are the category frequencies."""
    v = data[mask == 1]
    t = np.quantile(v, quantiles)
    cat = np.zeros(data.shape, np.uint8)
    c = np.minimum(np.searchsorted(t, data[mask == 1]), k - 1).astype(np.uint8)
    cat[mask == 1] = c
    marg = [float(np.mean(c == i)) for i in range(k)]
    return cat, marg


def smooth_field(shape, corr, seed):
    """Zero-mean unit-variance Gaussian random field by FFT filtering of
    seeded white noise with an exponential-like spectrum of length `corr`."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(s) for s in shape]
    kx, ky, kz = np.meshgrid(*k, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp = 1.0 / (1.0 + (2 * np.pi * corr * kk) ** 2) ** 1.0
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * amp))
    f = (f - f.mean()) / (f.std() + 1e-12)
    return np.asfortranarray(f.astype(np.float32))


def rel_err(a, b, scale):
    """max |a - b| / max(|b|, scale) over the arrays."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), scale))) if a.size else 0.0
