"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bar (FP64 path): relative L2 <= 1e-12 against the oracle on the same seeded
inputs, for forward and inverse, at every size the oracle finishes in
seconds; at the benchmark sizes (2D 4096^2, 3D 256^3) against the direct DFT
at sampled and planted bins, plus size-independent properties (round trip,
Parseval).  Calls follow the reference's sequence: topeFFTInit,
tope*PlanInit (uploads), tope*Exec (transforms on the device, downloads),
tope*Destroy.
"""
import ctypes as C

import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run(_fw, dims, x, direction):
    # the reference's full call sequence per transform (test/2d.c:82-114):
    # destroying the last plan releases the framework, so each call inits
    return topefft.transform(x, dims, direction)


@pytest.fixture
def gfw():
    f = topefft.Framework()
    yield f
    f.destroy()


def seeded(dims, salt=0):
    n = int(np.prod(dims))
    h = 0x7E57 + salt
    for d in dims:
        h = (h * 1000003 + d) & 0xFFFFFFFF
    return workloads.complex_uniform(h, n)


POW2_1D = [1 << l for l in range(0, 13)]  # 1 .. 4096: single fused pass


@pytest.mark.parametrize("n", POW2_1D)
def test_1d_pow2_fused(gfw, oracle, n):
    x = seeded((n,))
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, (n,), x, d)
        assert rel_l2(got, oracle.dft(x, (n,), d)) < TOL, (n, d)


@pytest.mark.parametrize("n", [1 << 13, 1 << 14, 1 << 16])
def test_1d_pow2_four_step(gfw, oracle, n):
    x = seeded((n,))
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, (n,), x, d)
        assert rel_l2(got, oracle.dft(x, (n,), d)) < TOL, (n, d)


def test_1d_large_four_step_bins(gfw, oracle):
    n = 1 << 22
    x = seeded((n,))
    rng = np.random.default_rng(5)
    bins = np.concatenate([[0, 1, n - 1, n // 2], rng.integers(0, n, 24)])
    fwd = run(gfw, (n,), x, topefft.FORWARD)
    assert rel_l2(fwd[bins], oracle.dft_bins(x, (n,), bins, 1)) < TOL
    back = run(gfw, (n,), fwd, topefft.INVERSE)
    assert rel_l2(back, x) < TOL


def smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


MIXED_1D = [6, 12, 15, 49, 96, 100, 243, 343, 625, 1000, 2187, 2401, 3000, 3125, 3969, 4000]


@pytest.mark.parametrize("n", MIXED_1D)
def test_1d_mixed_radix(oracle, n):
    # 7-smooth lengths up to 4096: one in-place mixed-radix pass (radix
    # 16/8/4/2, 7, 5, 3) instead of the reference's O(n^2) DFT
    x = seeded((n,))
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    names = p.pass_names()
    fwd = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert names == ["mixed_radix_f64"], names
    assert rel_l2(fwd, oracle.dft(x, (n,), topefft.FORWARD)) < TOL, n
    assert rel_l2(back, x) < TOL, n
    assert rel_l2(topefft.transform(x, (n,), topefft.INVERSE), oracle.dft(x, (n,), topefft.INVERSE)) < TOL


@pytest.mark.parametrize("n", [4116, 6000, 10000, 19683])
def test_1d_mixed_four_step(oracle, n):
    # 7-smooth lengths above 4096: two mixed-radix passes (n = n1 n2, both at
    # most 4096, four-step with the W_n^{n1 k2} post-twiddle), not Bluestein
    x = seeded((n,))
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    names = p.pass_names()
    fwd = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert names == ["mixed_radix_f64", "mixed_radix_f64"], names
    assert rel_l2(fwd, oracle.dft(x, (n,), topefft.FORWARD)) < TOL, n
    assert rel_l2(back, x) < TOL, n
    assert rel_l2(topefft.transform(x, (n,), topefft.INVERSE), oracle.dft(x, (n,), topefft.INVERSE)) < TOL


def test_1d_mixed_four_step_large(oracle):
    # 10^6 = 1000 x 1000 and 5^9 = 3125 x 625: sampled bins against the
    # direct DFT, the round trip, and a planted exponential's closed form
    for n in (1000000, 1953125):
        x = seeded((n,), salt=3)
        k0 = 12345
        x += np.exp(2j * np.pi * ((k0 * np.arange(n)) % n) / n)
        fw = topefft.Framework()
        p = topefft.Plan(fw, (n,), x)
        names = p.pass_names()
        X = p.exec(topefft.FORWARD)
        back = p.exec(topefft.INVERSE)
        p.destroy()
        fw.destroy()
        assert names == ["mixed_radix_f64", "mixed_radix_f64"], names
        rng = np.random.default_rng(n)
        bins = np.concatenate([[0, 1, k0, n - 1], rng.integers(0, n, 40)])
        assert rel_l2(X[bins], oracle.dft_bins(x, (n,), bins, topefft.FORWARD)) < TOL, n
        assert rel_l2(back, x) < TOL, n
        assert abs(X[k0] - n) < 1e-2 * n


@pytest.mark.parametrize("n", [105, 1000, 2187])
def test_1d_mixed_radix_closed_forms(n):
    # delta -> all ones; exp(2 pi i k0 j / n) -> n at k0 only
    d = np.zeros(n, complex)
    d[0] = 1.0
    assert np.max(np.abs(topefft.transform(d, (n,), topefft.FORWARD) - 1.0)) < 1e-13
    k0 = 7
    e = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    want = np.zeros(n, complex)
    want[k0] = n
    assert np.max(np.abs(topefft.transform(e, (n,), topefft.FORWARD) - want)) < 1e-9 * n


MIXED_ND = [(48, 36), (100, 60), (1000, 7), (3, 729), (1024, 1000), (4000, 3), (12, 10, 14), (60, 9, 25),
            (7, 256, 5), (100, 11),
            # strided lines of 1025..2048 (two or three per block), odd line counts
            (5, 2000), (3, 1200), (4, 2, 1500), (7, 1125, 2)]


@pytest.mark.parametrize("dims", MIXED_ND, ids=lambda d: "x".join(map(str, d)))
def test_nd_mixed_radix(oracle, dims):
    # every 7-smooth non-power-of-two axis (rows, columns, the 3D Z axis)
    # runs the mixed-radix pass; 11 (not 7-smooth) runs Bluestein beside it
    x = seeded(dims)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x)
    names = p.pass_names()
    fwd = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    mixed_axes = sum(1 for n in dims if n & (n - 1) and smooth7(n))
    assert names.count("mixed_radix_f64") == mixed_axes, names
    assert ("chirp_pre" in names) == (11 in dims), names
    assert rel_l2(fwd, oracle.dft(x, dims, topefft.FORWARD)) < TOL, dims
    assert rel_l2(back, x) < TOL, dims


@pytest.mark.parametrize("n", [3, 5, 6, 7, 12, 17, 100, 243, 1000])
def test_1d_non_pow2_dft(gfw, oracle, n):
    x = seeded((n,))
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, (n,), x, d)
        assert rel_l2(got, oracle.dft(x, (n,), d)) < TOL, (n, d)


@pytest.mark.parametrize("n", [65, 4097, 10007, 12007, 30011])
def test_1d_bluestein(oracle, n):
    # non-power-of-two lengths above 64 run Bluestein's chirp-z convolution
    # over power-of-two passes (length M >= 2n - 1, four-step above 4096)
    x = seeded((n,))
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    names = p.pass_names()
    fwd = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert names[0] == "chirp_pre" and names[-1] == "chirp_post", names
    assert rel_l2(fwd, oracle.dft(x, (n,), topefft.FORWARD)) < TOL, n
    assert rel_l2(back, x) < TOL, n
    if n <= 12007:
        assert rel_l2(topefft.transform(x, (n,), topefft.INVERSE), oracle.dft(x, (n,), topefft.INVERSE)) < TOL


def test_1d_golden_fixtures(gfw, golden):
    import test_oracle as T
    for n in [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 3, 5, 7, 12, 17, 100]:
        x = workloads.complex_uniform(T.case_seed((n,)), n)
        assert rel_l2(run(gfw, (n,), x, 1), golden[f"fwd_{n}"]) < TOL
        assert rel_l2(run(gfw, (n,), x, 0), golden[f"inv_{n}"]) < TOL


ND_NON_POW2 = [(12, 1), (3, 5), (100, 64), (64, 100), (30, 50), (2048 - 1, 4), (4, 1000), (12, 10, 8),
               (16, 24, 5), (7, 9, 11), (8, 3, 64)]


@pytest.mark.parametrize("dims", ND_NON_POW2, ids=lambda d: "x".join(map(str, d)))
def test_nd_non_pow2(gfw, oracle, dims):
    # the reference exits on these sizes (src/fft2d.c:355-362); here 7-smooth
    # axes run the mixed-radix pass, the others Bluestein's chirp-z
    # convolution (2047, 11)
    x = seeded(dims)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, dims, x, d)
        assert rel_l2(got, oracle.dft(x, dims, d)) < TOL, (dims, d)


ND_CASES = [(2, 2), (4, 4), (8, 4), (4, 16), (16, 16), (64, 32), (1, 8), (8, 1), (256, 256), (1024, 16),
            (16, 1024), (4096, 4), (4, 4096), (2, 2, 2), (4, 4, 4), (8, 16, 4), (16, 8, 32), (32, 32, 32),
            (1, 4, 8), (64, 4, 8), (4, 64, 8), (8, 4, 64)]


@pytest.mark.parametrize("dims", ND_CASES, ids=lambda d: "x".join(map(str, d)))
def test_nd_against_naive(gfw, oracle, dims):
    x = seeded(dims)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, dims, x, d)
        assert rel_l2(got, oracle.dft(x, dims, d)) < TOL, (dims, d)


@pytest.mark.parametrize("dims", [(8192, 4), (4, 8192), (16384, 2), (2, 8, 8192), (4, 8192, 2),
                                  (8192, 2, 2)], ids=lambda d: "x".join(map(str, d)))
def test_nd_long_axes_four_step(gfw, oracle, dims):
    x = seeded(dims)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run(gfw, dims, x, d)
        assert rel_l2(got, oracle.dft(x, dims, d)) < TOL, (dims, d)


def test_nd_golden_fixtures(gfw, golden):
    import test_oracle as T
    for dims in T.CASES_ND:
        x = workloads.complex_uniform(T.case_seed(dims), int(np.prod(dims)))
        k = T.key(dims)
        assert rel_l2(run(gfw, dims, x, 1), golden[f"fwd_{k}"]) < TOL
        assert rel_l2(run(gfw, dims, x, 0), golden[f"inv_{k}"]) < TOL


def test_reference_harness_inputs(gfw):
    import test_oracle as T
    # ramp of test/1d.c:70 and flat sinusoid of test/2d.c:77 (closed forms)
    for n in (16, 4096, 1 << 16):
        x = np.arange(1, n + 1).astype(complex)
        assert rel_l2(run(gfw, (n,), x, 1), T.ramp_spectrum(n)) < TOL
    for nx, ny in ((64, 32), (1024, 512)):
        x = np.sin(2 * np.pi * np.arange(nx * ny) / (nx * ny)).astype(complex)
        assert rel_l2(run(gfw, (nx, ny), x, 1), T.sinusoid_spectrum(nx, ny)) < 1e-11


def test_forward_then_inverse_on_device():
    # the reference's Exec does not re-read d: FORWARD then INVERSE round
    # trips on the device-resident copy (test/2d.c:101-102)
    dims = (512, 256)
    x = seeded(dims)
    p = topefft.Plan(topefft.Framework(), dims, x)
    p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    assert rel_l2(back, x) < TOL


def full_size_check(_fw, oracle, dims, seed):
    fw = topefft.Framework()
    dims3 = tuple(list(dims) + [1] * (3 - len(dims)))
    x, peaks = workloads.planted_input(seed, dims3)
    n = x.size
    p = topefft.Plan(fw, dims, x)
    X = p.exec(topefft.FORWARD)
    # sampled bins + planted peaks against the direct DFT
    rng = np.random.default_rng(seed)
    pb = workloads.peak_bins(dims3, peaks)
    bins = np.concatenate([pb, rng.integers(0, n, 56), [0, n - 1]])
    want = oracle.dft_bins(x, dims3, bins, 1)
    assert rel_l2(X[bins], want) < TOL
    # each planted exponential shows up at its bin with magnitude ~ n
    for b, (kx, ky, kz, ph) in zip(pb, peaks):
        assert abs(X[b] - n * np.exp(1j * ph)) < 0.01 * n
    # Parseval
    ex, eX = np.vdot(x, x).real, np.vdot(X, X).real
    assert abs(eX - n * ex) / (n * ex) < TOL
    # inverse on the device-resident spectrum returns the input
    back = p.exec(topefft.INVERSE)
    p.destroy()
    assert rel_l2(back, x) < TOL
    return X


def test_full_2d_4096(gfw, oracle):
    full_size_check(gfw, oracle, (4096, 4096), workloads.SEED_2D)


def test_full_3d_256(gfw, oracle):
    full_size_check(gfw, oracle, (256, 256, 256), workloads.SEED_3D)


def test_full_1d_65536_seeded(gfw, oracle):
    x = workloads.complex_uniform(workloads.SEED_1D, 1 << 16)
    X = run(gfw, (1 << 16,), x, 1)
    assert rel_l2(X, oracle.dft(x, (1 << 16,), 1)) < TOL
    assert rel_l2(run(gfw, (1 << 16,), X, 0), x) < TOL


def test_twiddle_table_exact():
    # GPU FP64 twiddle generation vs extended-precision values: entry
    # off[s] + (m-1)*Ns + k = W_{Ns*r}^{k*m}
    import mpmath
    mpmath.mp.dps = 30
    n = 4096
    p = topefft.Plan(topefft.Framework(), (n,), seeded((n,)))
    tw = p.twiddles(0)
    p.destroy()
    # schedule of the n=4096 kernel: radix 16,16,16 (E = 16)
    want = []
    ns = 16
    for s in (1, 2):
        for m in range(1, 16):
            for k in range(ns):
                e, M = (k * m) % (ns * 16), ns * 16
                a = 2 * mpmath.pi * e / M
                want.append(complex(float(mpmath.cos(a)), -float(mpmath.sin(a))))
        ns *= 16
    want = np.array(want)
    assert tw.shape == want.shape
    assert np.max(np.abs(tw - want)) < 2.3e-16


def test_profile_counters_and_passes():
    gfw = topefft.Framework()
    dims = (4096, 4096)
    x = seeded((64, 64))  # small host buffer replicated below
    big = np.tile(x, 4096)
    p = topefft.Plan(gfw, dims, big)
    assert p.t.totalMemory > 0
    assert p.t.totalPreKernel > 0
    p.exec(topefft.FORWARD)
    assert p.t.totalKernel > 0
    assert gfw.lib.profileThis(gfw.f.event) > 0
    assert p.pass_count() == 2
    p.pass_timing(True)
    for _ in range(3):
        p.exec_device(topefft.FORWARD)
    ms = p.pass_times()
    assert len(ms) == 2 and all(m > 0 for m in ms)
    p.destroy()


def test_error_paths():
    gfw = topefft.Framework()
    # a 2D/3D axis longer than 2^29 that is not a power of two: its
    # Bluestein length would exceed 2^30; refused at planning, before the
    # input (a few bytes here) is read
    t = topefft.TopePlan2D()
    tiny = np.zeros(4)
    gfw.lib.tope2DPlanInit(C.byref(gfw.f), C.byref(t), (1 << 29) + 1, 1, topefft.C2C,
                           tiny.ctypes.data_as(C.POINTER(C.c_double)))
    assert gfw.f.error == topefft.TOPE_ERROR_INVALID_SIZE
    gfw.f.error = 0
    p = topefft.Plan(gfw, (16,), np.zeros(16, complex))
    with pytest.raises(topefft.TopeError) as e:
        p.exec(5)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    p.destroy()
    gfw.destroy()


# the reference exits on these (src/fft2d.c:355-362, src/fft3d.c:522-533):
# 7-smooth axes above 4096 run the mixed-radix four-step (two mixed-radix
# passes, 6000 = 80 x 75, 5000 = 100 x 50, 4900 = 70 x 70), other axes above
# 2048 Bluestein with an inner FFT of 8192 or 16384 points (four-step)
LONG_NON_POW2 = [(6000, 4), (4, 5000), (3001, 2), (4097, 3), (2, 4100, 3), (3, 2, 4099), (3, 6000, 2),
                 (2, 3, 4900), (5000, 2, 3), (7, 4116)]


@pytest.mark.parametrize("dims", LONG_NON_POW2, ids=lambda d: "x".join(map(str, d)))
def test_nd_long_non_pow2(oracle, dims):
    x = seeded(dims, salt=21)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x)
    names = p.pass_names()
    fwd = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    long_smooth = [n for n in dims if n > 4096 and smooth7(n)]
    short_mixed = [n for n in dims if n <= 4096 and n & (n - 1) and smooth7(n)]
    assert names.count("mixed_radix_f64") == 2 * len(long_smooth) + len(short_mixed), names
    assert ("chirp_pre" in names) == any(not smooth7(n) for n in dims), names
    assert rel_l2(fwd, oracle.dft(x, dims, topefft.FORWARD)) < TOL, dims
    assert rel_l2(back, x) < TOL, dims
    assert rel_l2(topefft.transform(x, dims, topefft.INVERSE), oracle.dft(x, dims, topefft.INVERSE)) < TOL


def test_destroy_last_plan_releases_framework():
    f = topefft.Framework()
    p = topefft.Plan(f, (8,), np.ones(8, complex))
    p.destroy()
    assert not f.f.impl  # reference src/fft1d.c:385-392 releases the context
    with pytest.raises(topefft.TopeError) as e:
        topefft.Plan(f, (8,), np.ones(8, complex))
    assert e.value.code == topefft.TOPE_ERROR_NOT_INITIALIZED


@pytest.mark.parametrize("dims", [(8, 4), (16, 64), (64, 32), (256, 256), (1024, 512), (2, 16), (4096, 16)],
                         ids=lambda d: "x".join(map(str, d)))
def test_2d_dif2_schedule_small(oracle, monkeypatch, dims):
    # the schedule 2D 4096^2 uses (X pass on row pairs with the radix-2 DIF
    # step of Y, then half-length Y columns written interleaved), forced on
    # small shapes where the oracle's naive DFT is cheap
    monkeypatch.setenv("TOPEFFT_DIF2_MIN", "4")
    x = seeded(dims, salt=3)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x)
    names = p.pass_names()
    assert "dif2" in names[0], names
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    assert rel_l2(X, oracle.dft(x, dims, 1)) < TOL
    assert rel_l2(back, x) < TOL


@pytest.mark.parametrize("dims", [(16, 16), (64, 32), (8, 16, 4), (32, 32, 32), (64, 4, 8),
                                  (12, 10), (100, 6), (15, 7, 9), (1000, 3), (3000, 2), (6000, 3), (2, 5000),
                                  (3, 2, 4900),
                                  (8192, 4), (4, 8192), (16384, 2), (2, 2, 8192), (5, 2000), (7, 1125, 2)],
                         ids=lambda d: "x".join(map(str, d)))
def test_nd_complex64(oracle, dims):
    # complex64 storage plans (tope2DPlanInitEx / tope3DPlanInitEx); 7-smooth
    # axes run the mixed-radix pass in FP32
    x = seeded(dims, salt=9).astype(np.complex64)
    x64 = x.astype(np.complex128)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32)
    X = p.exec(topefft.FORWARD).astype(np.complex128)
    back = p.exec(topefft.INVERSE).astype(np.complex128)
    p.destroy()
    assert rel_l2(X, oracle.dft(x64, dims, 1)) < 1e-6
    assert rel_l2(back, x64) < 1e-6


def test_3d_256_row_tiles_complex64_against_fp64():
    # 256^3 takes the 4-row tiles for its 256-point rows (shapes of at least
    # 8192 such blocks), in FP64 and complex64; the complex64 plan agrees
    # with the FP64 one and both round-trip
    dims = (256, 256, 256)
    x = workloads.complex_uniform(workloads.SEED_3D + 11, 256 ** 3)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x)
    q = topefft.Plan(fw, dims, x.astype(np.complex64), precision=topefft.PREC_F32)
    names64, names32 = p.pass_names(), q.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    X32 = q.exec(topefft.FORWARD).astype(np.complex128)
    q.destroy()
    p.destroy()
    assert names64[0] == "axis_double_n256_e16_tb4_row", names64
    assert names32[0] == "axis_float_n256_e16_tb4_row", names32
    assert rel_l2(back, x) < TOL
    assert rel_l2(X32, X) < 1e-6


def test_complex64_rejects_uncovered_axes():
    fw = topefft.Framework()
    # 11 is neither a power of two nor 7-smooth: FP64 only (Bluestein)
    with pytest.raises(topefft.TopeError) as e:
        topefft.Plan(fw, (11, 4), np.zeros(44, np.complex64), precision=topefft.PREC_F32)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    # the matrix-core modes take powers of two only
    with pytest.raises(topefft.TopeError) as e:
        topefft.Plan(fw, (12, 256), np.zeros(12 * 256, np.complex64), precision=topefft.PREC_F32,
                     mode=topefft.MODE_MFMA_FP16X3)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    fw.destroy()
