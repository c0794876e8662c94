"""Real-input transforms (topeReal*): R2C half spectrum against the FP64
oracle's full complex transform (sliced to kx <= nx/2), C2R by round trip and
against the oracle's inverse, rows above 8192 and columns above 4096
through the four-step (full comparison up to 32768 points per axis, sampled
bins at 2^22), rows of 2^25 complex points in three passes (sampled bins),
7-smooth axes through the mixed-radix pass (and its four-step above 4096), complex64
storage (1e-6), closed forms, and the size limits.  Bar:
rel-L2 <= 1e-12, as for the complex FP64 path."""
import numpy as np
import pytest

import topefft
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# x above 8192: the row transform runs as the four-step (two passes) and the
# half-spectrum step on its own
SHAPES = [(2,), (4,), (16,), (256,), (4096,), (8192,), (16384,), (32768,), (16384, 4), (16384, 2, 2),
          (16, 8192), (4, 2, 8192), (200,), (24, 10), (30, 6, 10), (2000, 3), (8, 4), (64, 32), (256, 256), (2, 8),
          (16, 8, 4), (32, 32, 32), (2, 4, 8), (128, 2, 16),
          # 7-smooth axes above 4096: the mixed-radix four-step (x/2 = 6000 =
          # 80 x 75 and 5000 = 100 x 50 on the rows, y = 6000 on the columns)
          (12000,), (10000, 3), (16, 6000), (6, 2, 4900)]


def real_input(dims, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, int(np.prod(dims)))


def half(spectrum, dims):
    nx = dims[0]
    ny = dims[1] if len(dims) > 1 else 1
    nz = dims[2] if len(dims) > 2 else 1
    return spectrum.reshape(nz, ny, nx)[:, :, : nx // 2 + 1].reshape(-1)


@pytest.mark.parametrize("dims", SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_r2c_c2r(oracle, dims):
    x = real_input(dims, 40 + int(np.prod(dims)) % 31)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, x)
    X = p.forward()
    back = p.inverse()
    names = p.pass_names()
    p.destroy()
    fw.destroy()
    full = oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD)
    want = half(full, dims)
    assert X.size == want.size
    assert rel_l2(X, want) < 1e-12, names
    assert rel_l2(back, x) < 1e-12


def test_c2r_against_oracle_inverse(oracle):
    # the C2R output equals the oracle's inverse of the full Hermitian spectrum
    dims = (64, 16, 8)
    x = real_input(dims, 3)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, x)
    p.exec_device(topefft.FORWARD)
    got = p.inverse()
    p.destroy()
    fw.destroy()
    full = oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD)
    want = oracle.dft(full, dims, topefft.INVERSE)
    assert np.max(np.abs(want.imag)) < 1e-12
    assert rel_l2(got, want.real) < 1e-12


def r2c(dims, x):
    # a fresh framework per plan: destroying the last plan releases it
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, x)
    try:
        return p.forward()
    finally:
        p.destroy()
        fw.destroy()


def test_r2c_closed_forms():
    n = 1024
    # delta at 0 -> all ones
    d = np.zeros(n)
    d[0] = 1.0
    assert np.allclose(r2c((n,), d), 1.0, atol=1e-14)
    # cos(2 pi k0 j / n) -> n/2 at k0, zero elsewhere
    k0 = 37
    c = np.cos(2 * np.pi * k0 * np.arange(n) / n)
    want = np.zeros(n // 2 + 1, complex)
    want[k0] = n / 2
    assert np.max(np.abs(r2c((n,), c) - want)) < 1e-11


def test_real_plan_errors():
    fw = topefft.Framework()
    with pytest.raises(topefft.TopeError) as e:
        topefft.RealPlan(fw, (7,), np.zeros(7))
    assert e.value.code == topefft.TOPE_ERROR_INVALID_SIZE
    # x/2 = 11: neither a power of two nor 7-smooth
    with pytest.raises(topefft.TopeError) as e:
        topefft.RealPlan(fw, (22,), np.zeros(22))
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    # y above 4096: the column four-step uses the real array's buffer as
    # scratch at the spectrum's size, so caller arrays are refused
    p = topefft.RealPlan(fw, (16, 8192), np.zeros(16 * 8192))
    with pytest.raises(topefft.TopeError) as e:
        p.exec_ptr(p.t.real, p.t.data, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    p.destroy()
    fw.destroy()


def test_r2c_large_four_step_bins(oracle):
    # x = 2^22: sampled bins of the half spectrum against the oracle's direct
    # DFT, and the C2R round trip
    n = 1 << 22
    x = real_input((n,), 77)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, (n,), x)
    names = p.pass_names()
    X = p.forward()
    back = p.inverse()
    p.destroy()
    fw.destroy()
    assert len(names) == 3 and names[-1] == "real_half_spectrum", names
    rng = np.random.default_rng(9)
    bins = np.concatenate([[0, 1, n // 2, n // 2 - 1, n // 4], rng.integers(0, n // 2 + 1, 24)])
    want = oracle.dft_bins(x.astype(np.complex128), (n,), bins, topefft.FORWARD)
    assert rel_l2(X[bins], want) < 1e-12
    assert rel_l2(back, x) < 1e-12


def test_r2c_three_pass_bins(oracle):
    # x = 2^26: the 2^25-point complex row transform in three passes
    # (512 x 256 x 256), the half-spectrum step on its own; sampled bins
    # against the oracle's direct DFT and the C2R round trip
    n = 1 << 26
    x = real_input((n,), 78)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, (n,), x)
    names = p.pass_names()
    X = p.forward()
    back = p.inverse()
    p.destroy()
    fw.destroy()
    assert len(names) == 4 and names[-1] == "real_half_spectrum", names
    rng = np.random.default_rng(10)
    bins = np.concatenate([[0, 1, n // 2, n // 2 - 1, n // 4], rng.integers(0, n // 2 + 1, 6)])
    want = oracle.dft_bins(x.astype(np.complex128), (n,), bins, topefft.FORWARD)
    assert rel_l2(X[bins], want) < 1e-12
    assert rel_l2(back, x) < 1e-12


C64_SHAPES = [(256,), (4096,), (8192,), (64, 32), (256, 256), (16, 8, 4), (2, 8), (200,), (24, 10), (30, 6, 10),
              (16384,), (32768,), (16, 8192), (12000,), (16, 6000)]


@pytest.mark.parametrize("dims", C64_SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_r2c_c2r_complex64(oracle, dims):
    # complex64 storage (topeRealPlanInitEx with TOPE_PREC_F32): FP32
    # arithmetic, tables generated in FP64 and rounded once; bar 1e-6
    # against the FP64 oracle
    x = real_input(dims, 90 + int(np.prod(dims)) % 17).astype(np.float32)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, x, precision=topefft.PREC_F32)
    assert p.t.precision == topefft.PREC_F32 and p.t.realSize == x.size * 4
    names = p.pass_names()
    X = p.forward()
    back = p.inverse()
    p.destroy()
    fw.destroy()
    assert X.dtype == np.complex64 and back.dtype == np.float32
    assert not any(nm.startswith("mixed_radix_f64") or "double" in nm for nm in names), names
    want = half(oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD), dims)
    assert rel_l2(X.astype(np.complex128), want) < 1e-6
    assert rel_l2(back.astype(np.float64), x.astype(np.float64)) < 1e-6


def test_real_complex64_limits():
    fw = topefft.Framework()
    with pytest.raises(topefft.TopeError) as e:
        topefft.RealPlan(fw, (22,), np.zeros(22, np.float32), precision=topefft.PREC_F32)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    with pytest.raises(topefft.TopeError) as e:
        topefft.RealPlan(fw, (16,), np.zeros(16, np.float32), precision=5)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    fw.destroy()


def test_r2c_complex64_packed_pair_tile(oracle, monkeypatch):
    # complex64 plans with column passes carry the DC and Nyquist columns
    # packed as one complex column (an even column count), so 4096-point
    # columns can run the paired-column tile: forced here onto 16 x 4096
    # (8 packed columns), forward against the FP64 oracle and the round trip
    monkeypatch.setenv("TOPEFFT_VARIANT_4096_col", "axis_float_n4096_e16_tb8_col_ct3")
    dims = (16, 4096)
    x = real_input(dims, 123).astype(np.float32)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, x, precision=topefft.PREC_F32)
    names = p.pass_names()
    X = p.forward()
    back = p.inverse()
    p.destroy()
    fw.destroy()
    # 8-point rows: the row pass, then the half-spectrum step on its own
    assert names[-3:] == ["real_half_spectrum", "axis_float_n4096_e16_tb8_col_ct3", "real_pack_dc_nyquist"], names
    want = half(oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD), dims)
    assert rel_l2(X.astype(np.complex128), want) < 1e-6
    assert rel_l2(back.astype(np.float64), x.astype(np.float64)) < 1e-6


def test_r2c_fp64_keeps_nyquist_column():
    # FP64 plans transform the m + 1 columns as they are (no pack pass)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, (64, 32), np.zeros(64 * 32))
    names = p.pass_names()
    p.destroy()
    fw.destroy()
    assert "real_pack_dc_nyquist" not in names, names
