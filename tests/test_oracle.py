"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Independent anchors: numpy/pocketfft golden fixtures (tests/golden), closed
forms (delta, constant, planted exponentials, the ramp input of the
reference's test/1d.c:70 and the flat sinusoid of test/2d.c:77), brute-force
sums at tiny n, and invariants (Parseval, linearity, round trip).  Each check
is chosen so that a plausible mistake -- wrong sign, transposed axis, missing
1/N, wrong twiddle index -- fails at least one of them (see the negative
tests at the end, which plant such mistakes in the oracle's output).
"""
import cmath
import math
from fractions import Fraction

import numpy as np
import pytest

import workloads
from conftest import rel_l2

TOL = 1e-13

CASES_1D = [1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 32, 64, 100, 128, 256, 512, 1024, 2048, 4096]
CASES_ND = [(4, 4), (8, 4), (4, 16), (16, 16), (64, 32), (2, 2, 2), (4, 4, 4), (8, 16, 4), (16, 8, 32)]


def case_seed(dims):
    h = 0x60D
    for d in dims:
        h = (h * 1000003 + d) & 0xFFFFFFFF
    return h


def key(dims):
    return "x".join(str(d) for d in dims)


@pytest.mark.parametrize("dims", [(n,) for n in CASES_1D] + CASES_ND, ids=key)
def test_naive_dft_matches_golden(oracle, golden, dims):
    n = int(np.prod(dims))
    x = workloads.complex_uniform(case_seed(dims), n)
    assert abs(x.sum() - golden[f"in_sum_{key(dims)}"][0]) < 1e-9
    fwd = oracle.dft(x, dims, oracle.FORWARD)
    inv = oracle.dft(x, dims, oracle.INVERSE)
    assert rel_l2(fwd, golden[f"fwd_{key(dims)}"]) < TOL
    assert rel_l2(inv, golden[f"inv_{key(dims)}"]) < TOL


@pytest.mark.parametrize("dims", [(n,) for n in CASES_1D if n >= 2 and (n & (n - 1)) == 0]
                         + [d for d in CASES_ND if all((v & (v - 1)) == 0 for v in d)], ids=key)
def test_refcpu_matches_golden(oracle, golden, dims):
    n = int(np.prod(dims))
    x = workloads.complex_uniform(case_seed(dims), n)
    assert rel_l2(oracle.refcpu(x, dims, oracle.FORWARD), golden[f"fwd_{key(dims)}"]) < TOL
    assert rel_l2(oracle.refcpu(x, dims, oracle.INVERSE), golden[f"inv_{key(dims)}"]) < TOL


def test_refcpu_radix_rule(oracle):
    # fft1d.c:320-326 / fft2d.c:335-354
    for l in range(1, 25):
        want = 8 if l % 3 == 0 else 4 if l % 2 == 0 else 2
        assert oracle.radix(1 << l) == want
    assert oracle.radix(12) == -1


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 16])
def test_naive_dft_brute_force(oracle, n):
    # exact rational phases, math.cos/sin, plain sums
    x = workloads.complex_uniform(0xB00 + n, n)
    for direction, sign in ((oracle.FORWARD, -1), (oracle.INVERSE, 1)):
        want = []
        for k in range(n):
            s = 0j
            for j in range(n):
                ph = Fraction((j * k) % n, n)
                s += complex(x[j]) * cmath.exp(sign * 2j * math.pi * float(ph))
            want.append(s / n if direction == oracle.INVERSE else s)
        assert rel_l2(oracle.dft(x, (n,), direction), np.array(want)) < 1e-14


@pytest.mark.parametrize("n", [2, 8, 64, 1024, 4096])
def test_closed_forms_1d(oracle, n):
    k0 = (n // 3) % n
    delta = np.zeros(n, complex); delta[0] = 1
    const = np.ones(n, complex)
    expo = np.exp(2j * np.pi * ((np.arange(n) * k0) % n) / n)
    e_k0 = np.zeros(n, complex); e_k0[k0] = n
    for f in (oracle.dft, oracle.refcpu):
        assert rel_l2(f(delta, (n,), 1), np.ones(n)) < TOL
        c = np.zeros(n, complex); c[0] = n
        assert rel_l2(f(const, (n,), 1), c) < TOL
        assert rel_l2(f(expo, (n,), 1), e_k0) < TOL


def ramp_spectrum(n):
    # x[m] = m + 1 (test/1d.c:70): X[0] = n(n+1)/2, X[k] = n / (W^k - 1)
    k = np.arange(n)
    w = np.exp(-2j * np.pi * k / n)
    out = np.empty(n, complex)
    out[0] = n * (n + 1) / 2
    out[1:] = n / (w[1:] - 1)
    return out


@pytest.mark.parametrize("n", [4, 16, 512, 4096])
def test_ramp_known_answer(oracle, n):
    x = np.arange(1, n + 1).astype(complex)
    want = ramp_spectrum(n)
    assert rel_l2(oracle.dft(x, (n,), 1), want) < TOL
    assert rel_l2(oracle.refcpu(x, (n,), 1), want) < TOL


def sinusoid_spectrum(nx, ny):
    # x[j*nx + i] = sin(2 pi (j*nx + i) / (nx*ny)) (test/2d.c:75-80): only
    # ky = 1 and ky = ny-1 are non-zero, each a geometric sum over i.
    N = nx * ny
    kx = np.arange(nx)
    out = np.zeros((ny, nx), complex)
    for s, ky in ((1, 1 % ny), (-1, (ny - 1) % ny)):
        r = np.exp(2j * np.pi * (s / N - kx / nx))
        g = (1 - r ** nx) / (1 - r)
        out[ky] += s * ny / 2j * g
    return out.reshape(-1)


@pytest.mark.parametrize("dims", [(8, 4), (16, 16), (64, 32), (4, 64)])
def test_flat_sinusoid_known_answer(oracle, dims):
    nx, ny = dims
    x = np.sin(2 * np.pi * np.arange(nx * ny) / (nx * ny)).astype(complex)
    want = sinusoid_spectrum(nx, ny)
    assert rel_l2(oracle.dft(x, dims, 1), want) < 1e-12
    assert rel_l2(oracle.refcpu(x, dims, 1), want) < 1e-12


@pytest.mark.parametrize("dims", [(64,), (16, 8), (8, 4, 16)], ids=key)
def test_invariants(oracle, dims):
    n = int(np.prod(dims))
    x = workloads.complex_uniform(7, n)
    y = workloads.complex_gaussian(8, n)
    X, Y = oracle.dft(x, dims, 1), oracle.dft(y, dims, 1)
    # Parseval
    assert abs(np.vdot(X, X).real - n * np.vdot(x, x).real) / (n * np.vdot(x, x).real) < TOL
    # linearity
    a, b = 0.3 - 1.2j, 2.5 + 0.25j
    assert rel_l2(oracle.dft(a * x + b * y, dims, 1), a * X + b * Y) < TOL
    # round trip
    assert rel_l2(oracle.dft(X, dims, 0), x) < TOL


def test_planted_exponentials(oracle):
    dims = (16, 8, 4)
    n = 16 * 8 * 4
    x = np.zeros(n, complex)
    peaks = workloads.planted_peaks(3, dims)
    workloads.add_peaks(x, dims, peaks)
    X = oracle.dft(x, dims, 1)
    want = np.zeros(n, complex)
    for (kx, ky, kz, ph), b in zip(peaks, workloads.peak_bins(dims, peaks)):
        want[b] += n * np.exp(1j * ph)
    assert rel_l2(X, want) < TOL


def test_bins_match_full_dft(oracle):
    dims = (32, 16, 8)
    n = 32 * 16 * 8
    x = workloads.complex_uniform(11, n)
    bins = np.array([0, 1, 5, 31, 32, 511, 1000, n - 1, 2345, 4095])
    for d in (1, 0):
        full = oracle.dft(x, dims, d)
        assert rel_l2(oracle.dft_bins(x, dims, bins, d), full[bins]) < TOL


def test_splitmix64_known_values():
    # SplitMix64 reference outputs for seed 0 (Vigna's published generator)
    z = workloads.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


# ---- negative controls: planted mistakes must be caught ------------------

def test_checks_reject_planted_mistakes(oracle, golden):
    dims = (64, 32)
    x = workloads.complex_uniform(case_seed(dims), 64 * 32)
    good = golden["fwd_64x32"]
    X = oracle.dft(x, dims, 1)
    assert rel_l2(X, good) < TOL
    wrong_sign = oracle.dft(x, dims, 0) * (64 * 32)          # conjugate twiddles
    transposed = oracle.dft(x.reshape(32, 64).T.copy().reshape(-1), (32, 64), 1)
    no_scale = oracle.dft(x, dims, 0)                        # inverse vs forward
    shifted = np.roll(X, 1)                                  # index error
    for bad in (wrong_sign, transposed, no_scale, shifted):
        assert rel_l2(bad, good) > 1e-3
    # ramp KAT rejects a conjugated spectrum
    r = np.arange(1, 65).astype(complex)
    assert rel_l2(np.conj(oracle.dft(r, (64,), 1)), ramp_spectrum(64)) > 1e-3
