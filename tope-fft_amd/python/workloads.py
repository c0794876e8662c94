"""Seeded synthetic inputs for the transform workloads.

Every input is produced on the host from SplitMix64 (S. Vigna, public
domain), so a CPU check and a GPU run see identical bits.  Conventions:

* stream i of a seed: z_i = mix(seed + (i+1) * 0x9E3779B97F4A7C15)
* uniform in [-1, 1):  2 * ((z >> 11) * 2^-53) - 1
* Gaussian (Box-Muller): u1 = ((z_{2i} >> 11) + 1) * 2^-53,
  u2 = (z_{2i+1} >> 11) * 2^-53, re = r cos(2 pi u2), im = r sin(2 pi u2)
* complex arrays take re from value 2k and im from value 2k+1 (uniform)

Seeds of the benchmark configurations (kept here so tests and the bench
agree): 1D 2^16 0x5EED0001, 2D 4096^2 0x5EED0002, 3D 256^3 0x5EED0003,
batch 4096x4096 0x5EED0004, uneven batch sizes 0x5EED0005 / data
0x5EED0006, 3D 512^3 0x5EED0007.
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

SEED_1D = 0x5EED0001
SEED_2D = 0x5EED0002
SEED_3D = 0x5EED0003
SEED_BATCH = 0x5EED0004
SEED_MANY_SIZES = 0x5EED0005
SEED_MANY_DATA = 0x5EED0006
SEED_3D_LARGE = 0x5EED0007

_CHUNK = 1 << 24


def splitmix64(seed, count, start=0):
    """uint64 array of SplitMix64 outputs start .. start+count-1."""
    out = np.empty(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for c0 in range(0, count, _CHUNK):
            c1 = min(count, c0 + _CHUNK)
            i = np.arange(start + c0 + 1, start + c1 + 1, dtype=np.uint64)
            z = np.uint64(seed) + i * GAMMA
            z = (z ^ (z >> np.uint64(30))) * M1
            z = (z ^ (z >> np.uint64(27))) * M2
            out[c0:c1] = z ^ (z >> np.uint64(31))
    return out


def uniform(seed, count):
    """float64 values in [-1, 1)."""
    z = splitmix64(seed, count)
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0**-53) - 1.0


def complex_uniform(seed, n):
    """complex128 array of n values, re and im uniform in [-1, 1)."""
    v = uniform(seed, 2 * n)
    return v.view(np.complex128).copy()


def complex_gaussian(seed, n):
    """complex128 array of n values, re and im standard Gaussian."""
    z = splitmix64(seed, 2 * n)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0**-53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * 2.0**-53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(n, dtype=np.complex128)
    out.real = r * np.cos(2.0 * np.pi * u2)
    out.imag = r * np.sin(2.0 * np.pi * u2)
    return out


def uneven_sizes(seed, count, lmin=6, lmax=14):
    """`count` powers of two drawn log-uniform from 2^lmin .. 2^lmax."""
    z = splitmix64(seed, count)
    return (1 << (lmin + (z % np.uint64(lmax - lmin + 1)).astype(np.int64))).astype(np.int32)


def planted_peaks(seed, dims, count=8):
    """`count` distinct frequency triples (kx, ky, kz) and phases drawn from
    the seed's second stream; dims = (nx, ny, nz)."""
    z = splitmix64(seed ^ 0xA5A5A5A5A5A5A5A5, 16 * count)
    nx, ny, nz = dims
    peaks, seen = [], set()
    i = 0
    while len(peaks) < count:
        kx = int(z[i] % np.uint64(nx)); ky = int(z[i + 1] % np.uint64(ny))
        kz = int(z[i + 2] % np.uint64(nz))
        ph = float((z[i + 3] >> np.uint64(11)).astype(np.float64) * 2.0**-53) * 2.0 * np.pi
        i += 4
        if (kx, ky, kz) in seen:
            continue
        seen.add((kx, ky, kz))
        peaks.append((kx, ky, kz, ph))
    return peaks


def _axis_phase(k, n):
    # exp(2 pi i k x / n) for x < n, phase index reduced exactly mod n
    idx = (np.arange(n, dtype=np.int64) * k) % n
    return np.exp(2j * np.pi * idx / n)


def add_peaks(x, dims, peaks):
    """x (flat, x fastest) += sum of unit exponentials exp(i(2 pi k.n/N + ph))."""
    nx, ny, nz = dims
    v = x.reshape(nz, ny, nx)
    for kx, ky, kz, ph in peaks:
        ex, ey, ez = _axis_phase(kx, nx), _axis_phase(ky, ny), _axis_phase(kz, nz)
        plane = np.exp(1j * ph) * np.outer(ey, ex)
        for zz in range(nz):
            v[zz] += ez[zz] * plane
    return x


def peak_bins(dims, peaks):
    nx, ny, _ = dims
    return np.array([kx + nx * (ky + ny * kz) for kx, ky, kz, _ in peaks], dtype=np.int64)


def planted_input(seed, dims, count=8):
    """Uniform [-1,1) re/im plus `count` planted exponentials (configs with
    spectral peaks checkable by position).  Returns (x, peaks)."""
    n = dims[0] * dims[1] * dims[2]
    x = complex_uniform(seed, n)
    peaks = planted_peaks(seed, dims, count)
    add_peaks(x, dims, peaks)
    return x, peaks


def powerlaw_spectrum(seed, dims, slope=-2.0):
    """k-space coefficients of a Gaussian random field with power |k|^slope:
    complex Gaussian(0,1) times |k|^(slope/2) (k = 0 set to 0), |k| from the
    signed frequency indices.  The real-space field is its inverse DFT."""
    nx, ny, nz = dims
    z = complex_gaussian(seed, nx * ny * nz).reshape(nz, ny, nx)
    fx = np.fft.fftfreq(nx) * nx
    fy = np.fft.fftfreq(ny) * ny
    fz = np.fft.fftfreq(nz) * nz
    for iz in range(nz):
        k2 = fz[iz] ** 2 + fy[:, None] ** 2 + fx[None, :] ** 2
        with np.errstate(divide="ignore"):
            amp = np.where(k2 > 0, k2 ** (slope / 4.0), 0.0)
        z[iz] *= amp
    return z.reshape(-1)


def interleaved(x):
    """complex128 -> contiguous float64 view (re, im, re, im, ...)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    return x.view(np.float64)
