"""The 3D 512^3 configuration at full size (2 GiB FP64, 1 GiB complex64).

FP64: the forward transform of the seeded uniform input with 8 planted unit
exponentials is compared with the direct DFT at every planted bin and at
sampled bins (rel-L2 <= 1e-12), each planted peak is checked by position and
phase, then Parseval and the device round trip.  This pins the two kernels
the plan runs (512-point rows in 8-row tiles, 512-point columns in 16-column
split tiles) in the forward direction at their real shape.

complex64: the VALU plan and the matrix-core plans (split-fp16 and bf16x6
operands, 512-point rows with a radix-2 VALU stage and 16-line column tiles
on Y and Z) against the FP64 spectrum at rel-L2 <= 1e-6, with the kernel
family of every pass asserted, and their round trips.
"""
import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu

DIMS = (512, 512, 512)
N = 512 ** 3


@pytest.fixture(scope="module")
def field():
    x, peaks = workloads.planted_input(workloads.SEED_3D_LARGE, DIMS)
    fw = topefft.Framework()
    p = topefft.Plan(fw, DIMS, x)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    return x, peaks, names, X, back


def test_512_fp64_kernels(field):
    _, _, names, _, _ = field
    assert names == ["axis_double_n512_e16_tb8_row", "axis_double_n512_e16_tb16_col_split",
                     "axis_double_n512_e16_tb16_col_split"], names


def test_512_fp64_bins_and_peaks(field, oracle):
    x, peaks, _, X, _ = field
    pb = workloads.peak_bins(DIMS, peaks)
    rng = np.random.default_rng(512)
    bins = np.concatenate([pb, rng.integers(0, N, 56), [0, 1, 512, 512 * 512, N - 1]])
    want = oracle.dft_bins(x, DIMS, bins, topefft.FORWARD)
    assert rel_l2(X[bins], want) < 1e-12
    for b, (kx, ky, kz, ph) in zip(pb, peaks):
        assert abs(X[b] - N * np.exp(1j * ph)) < 0.01 * N, (kx, ky, kz)


def test_512_fp64_parseval_and_round_trip(field):
    x, _, _, X, back = field
    ex, eX = np.vdot(x, x).real, np.vdot(X, X).real
    assert abs(eX - N * ex) / (N * ex) < 1e-12
    assert rel_l2(back, x) < 1e-12


MODES = [(None, ["axis_float_n512_e16_tb8_row", "axis_float_n512_e16_tb32_col_split",
                 "axis_float_n512_e16_tb32_col_split"]),
         (topefft.MODE_MFMA_FP16X3, ["mfma_rows_fp16x3", "mfma_cols_fp16x3", "mfma_cols_fp16x3"]),
         (topefft.MODE_MFMA_BF16X6, ["mfma_rows_bf16x6", "mfma_cols_bf16x6", "mfma_cols_bf16x6"])]


@pytest.mark.parametrize("mode,kernels", MODES, ids=["valu", "mfma_fp16x3", "mfma_bf16x6"])
def test_512_complex64_against_fp64(field, mode, kernels):
    x, _, _, X, _ = field
    x32 = x.astype(np.complex64)
    fw = topefft.Framework()
    q = topefft.Plan(fw, DIMS, x32, precision=topefft.PREC_F32, mode=mode)
    names = q.pass_names()
    X32 = q.exec(topefft.FORWARD)
    back = q.exec(topefft.INVERSE)
    q.destroy()
    fw.destroy()
    assert names == kernels, names
    err = rel_l2(X32.astype(np.complex128), X)
    rt = rel_l2(back.astype(np.complex128), x32.astype(np.complex128))
    print(f"512^3 complex64 {names[0]}: rel-L2 vs FP64 {err:.3e}, round trip {rt:.3e}")
    assert err < 1e-6
    assert rt < 1e-6


# 2D 4096^2 (the benchmark shape): the complex64 plans' 4096-point column
# pass runs the paired-column three-step tile (8 columns, two per thread)
# under every mode; their spectra against the FP64 plan's (pinned to the
# direct DFT by the full-size parity test of this shape), both directions.
DIMS2 = (4096, 4096)


@pytest.fixture(scope="module")
def field2d():
    x, _ = workloads.planted_input(workloads.SEED_2D, DIMS2 + (1,))
    fw = topefft.Framework()
    p = topefft.Plan(fw, DIMS2, x)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    Xi = p.exec(topefft.INVERSE)  # of the spectrum: x again
    p.destroy()
    fw.destroy()
    assert names == ["axis_double_n4096_e16_tb1_row_split", "axis_double_n4096_e16_tb4_col_ct3"], names
    return x, X, Xi


MODES2 = [(None, "axis_float_n4096_e16_tb1_row_split"), (topefft.MODE_MFMA_FP16X3, "mfma_rows_fp16x3"),
          (topefft.MODE_MFMA_BF16X6, "mfma_rows_bf16x6")]


@pytest.mark.parametrize("mode,rows", MODES2, ids=["valu", "mfma_fp16x3", "mfma_bf16x6"])
def test_4096_complex64_pair_tile(field2d, mode, rows):
    x, X, Xi = field2d
    x32 = x.astype(np.complex64)
    fw = topefft.Framework()
    q = topefft.Plan(fw, DIMS2, x32, precision=topefft.PREC_F32, mode=mode)
    names = q.pass_names()
    X32 = q.exec(topefft.FORWARD).astype(np.complex128)
    back = q.exec(topefft.INVERSE).astype(np.complex128)
    q.destroy()
    fw.destroy()
    assert names == [rows, "axis_float_n4096_e16_tb8_col_ct3"], names
    err = rel_l2(X32, X)
    rt = rel_l2(back, x32.astype(np.complex128))
    print(f"4096^2 complex64 {names}: rel-L2 vs FP64 {err:.3e}, round trip {rt:.3e}")
    assert err < 1e-6
    assert rt < 1e-6
    assert rel_l2(back, Xi) < 1e-6


def test_4096_real_complex64_against_fp64():
    # 2D 4096^2 R2C with complex64 storage: the row step packs the DC and
    # Nyquist columns, the 4096-point columns run the paired-column tile,
    # the last pass unpacks; the half spectrum against the FP64 real plan's
    # (whose kernels are pinned to the oracle) and the C2R round trip
    rng = np.random.default_rng(4096)
    x = rng.uniform(-1.0, 1.0, 4096 * 4096)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, DIMS2, x)
    X64 = p.forward()
    # the next plan before this one is destroyed (the last destroy releases
    # the framework)
    q = topefft.RealPlan(fw, DIMS2, x.astype(np.float32), precision=topefft.PREC_F32)
    p.destroy()
    names = q.pass_names()
    X32 = q.forward().astype(np.complex128)
    back = q.inverse().astype(np.float64)
    q.destroy()
    fw.destroy()
    assert names == ["real_rows_fused", "axis_float_n4096_e16_tb8_col_ct3", "real_pack_dc_nyquist"], names
    err = rel_l2(X32, X64)
    print(f"4096^2 real complex64 rel-L2 vs FP64 {err:.3e}")
    assert err < 1e-6
    assert rel_l2(back, x.astype(np.float32).astype(np.float64)) < 1e-6
