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
    cat = np.zeros(data.shape, np.uint8, order="F")  # the grids' Fortran order
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


def uneven_twins(shape, seed, sigma=15.0, twin_frac=0.005, twin_sd=1e-3):
    """Precision-stress property: informed density varying smoothly from
    0.1% to 5%, p = 0.001 + 0.049 / (1 + exp(-2 f)) with f white noise
    Gaussian-filtered (sigma cells, wrapped edges) and standardised; each
    cell is informed with probability p, values N(0, 1).  Then a fraction
    twin_frac of the samples get a twin in an x-adjacent uninformed cell
    (x + 1, else x - 1) with the value plus N(0, twin_sd): near-duplicate
    samples one cell apart (grid coordinates are integers, so sub-cell
    jitter cannot be expressed).  Returns data, mask and the count of twins."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    f = gaussian_filter(noise, sigma, mode="wrap")
    f = (f - f.mean()) / (f.std() + 1e-12)
    p = 0.001 + 0.049 / (1.0 + np.exp(-2.0 * f))
    mask = (rng.random(shape) < p).astype(np.uint8)
    data = np.full(shape, -99.0, np.float32)
    n_inf = int(mask.sum())
    data[mask == 1] = rng.standard_normal(n_inf).astype(np.float32)
    src = np.argwhere(mask == 1)
    pick = src[rng.choice(len(src), int(round(twin_frac * len(src))), replace=False)]
    X = shape[0]
    twins = 0
    for x, y, z in pick:
        for nx in (x + 1, x - 1):
            if 0 <= nx < X and mask[nx, y, z] == 0:
                mask[nx, y, z] = 1
                data[nx, y, z] = np.float32(data[x, y, z] + rng.normal(0.0, twin_sd))
                twins += 1
                break
    return np.asfortranarray(data), np.asfortranarray(mask), twins
