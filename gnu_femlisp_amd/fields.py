"""Seeded random coefficient fields on n^3 cells (synthetic inputs).

inclusions: two-phase random balls (SplitMix64 stream, Steele, Lea, Flood
2014), the 128^3 deployment workload.  gaussian_log_field: kappa = exp(g) with
g a smoothed Gaussian random field, the training workload.
"""
import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def inclusions(n, seed=2024, frac=0.30, rmin=3.0, rmax=8.0, contrast=1e4):
    """1 in the matrix phase, `contrast` inside random balls (centres uniform in
    the cube, radii uniform in [rmin, rmax] cells) added until the inclusion
    volume fraction reaches `frac`."""
    rng = SplitMix64(seed)
    inside = np.zeros((n, n, n), bool)
    c = np.arange(n) + 0.5
    target = frac * n ** 3
    count = 0
    while count < target:
        cx, cy, cz = (rng.uniform() * n for _ in range(3))
        r = rmin + (rmax - rmin) * rng.uniform()
        lo = [max(int(v - r), 0) for v in (cx, cy, cz)]
        hi = [min(int(v + r) + 1, n) for v in (cx, cy, cz)]
        X, Y, Z = np.ix_(c[lo[0]:hi[0]], c[lo[1]:hi[1]], c[lo[2]:hi[2]])
        ball = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
        sub = inside[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        count += int(np.count_nonzero(ball & ~sub))
        sub |= ball
    k = np.ones((n, n, n))
    k[inside] = contrast
    return k


def gaussian_log_field(n, seed, corr=4.0, sigma=2.0):
    """kappa = exp(g): g = white noise smoothed by a Gaussian kernel of width
    `corr` cells (periodic FFT filter), scaled to standard deviation `sigma`."""
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(n, n, n))
    k = np.fft.fftfreq(n) * 2 * np.pi
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * corr ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    g = (g - g.mean()) / g.std() * sigma
    return np.exp(g)
