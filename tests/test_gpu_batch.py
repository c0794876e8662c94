"""Batched 1D transforms: FP64 and the FP32-storage variants (VALU and the
MFMA DFT-matrix modes) against the FP64 oracle.

Bars: FP64 rel-L2 <= 1e-12.  FP32 storage: plain FP32 arithmetic (VALU and
MFMA fp32 operands) and the 3-part bf16 split <= 1e-6; the 2-part bf16 split
(dropped terms ~2^-16) <= 1e-4.  The measured errors are printed.
"""
import ctypes as C

import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu

MODE_TOL = {topefft.MODE_VALU: 1e-6, topefft.MODE_MFMA_FP32: 1e-6, topefft.MODE_MFMA_BF16X6: 1e-6,
            topefft.MODE_MFMA_BF16X3: 1e-4, topefft.MODE_MFMA_FP16X3: 1e-6}


def batch_input(n, count, seed):
    return workloads.complex_gaussian(seed, n * count)


def run_batch(n, count, x, prec, mode, direction):
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, count, x, prec, mode)
    try:
        return p.exec(direction)
    finally:
        p.destroy()
        fw.destroy()


# powers of two; 7-smooth lengths (mixed-radix pass, four-step of two above
# 4096); others (Bluestein)
@pytest.mark.parametrize("n", [2, 16, 256, 1024, 4096, 8192, 3, 12, 100, 1000, 4000, 6000, 11, 1009, 2048 - 1])
def test_batch_fp64(oracle, n):
    count = 12
    x = batch_input(n, count, 100 + n)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run_batch(n, count, x, topefft.PREC_F64, topefft.MODE_VALU, d).reshape(count, n)
        for r in (0, 5, count - 1):
            assert rel_l2(got[r], oracle.dft(x[r * n:(r + 1) * n], (n,), d)) < 1e-12


@pytest.mark.parametrize("mode", [topefft.MODE_VALU, topefft.MODE_MFMA_BF16X3, topefft.MODE_MFMA_BF16X6,
                                  topefft.MODE_MFMA_FP32, topefft.MODE_MFMA_FP16X3],
                         ids=lambda m: topefft.MODE_NAMES[m])
@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
def test_batch_fp32_modes(oracle, n, mode):
    count = 40
    x = batch_input(n, count, 200 + n).astype(np.complex64)
    x64 = x.astype(np.complex128)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run_batch(n, count, x, topefft.PREC_F32, mode, d).reshape(count, n).astype(np.complex128)
        errs = []
        for r in (0, 1, 17, count - 1):
            want = oracle.dft(x64[r * n:(r + 1) * n], (n,), d)
            errs.append(rel_l2(got[r], want))
        print(f"n={n} mode={topefft.MODE_NAMES[mode]} dir={d} rel-L2 max {max(errs):.3e}")
        assert max(errs) < MODE_TOL[mode], errs


def test_batch_fp32_full_4096(oracle):
    # the benchmark batch: 4096 transforms of 4096, Gaussian input; every
    # MFMA mode against the FP64 GPU batch (itself pinned to the oracle)
    n = count = 4096
    x = workloads.complex_gaussian(workloads.SEED_BATCH, n * count)
    ref = run_batch(n, count, x, topefft.PREC_F64, topefft.MODE_VALU, topefft.FORWARD)
    xs = x.reshape(count, n)
    for r in (0, 2048, 4095):
        assert rel_l2(ref.reshape(count, n)[r], oracle.dft(xs[r], (n,), 1)) < 1e-12
    x32 = x.astype(np.complex64)
    for mode, tol in MODE_TOL.items():
        got = run_batch(n, count, x32, topefft.PREC_F32, mode, topefft.FORWARD).astype(np.complex128)
        err = rel_l2(got, ref)
        mx = np.max(np.abs(got - ref))
        print(f"4096x4096 mode={topefft.MODE_NAMES[mode]} rel-L2 vs FP64 {err:.3e} max-abs {mx:.3e}")
        assert err < tol


@pytest.mark.parametrize("n", [256, 512, 2048, 4096])
def test_batch_fp16x3_range(oracle, n):
    # fp16 operands: every transform is scaled by its own power of two, so
    # magnitudes far outside fp16's range, mixed within one batch, keep the
    # same relative accuracy; zero rows stay zero; one dominant element
    # beside tiny ones is exact up to FP32 rounding
    rows = []
    base = batch_input(n, 6, 300 + n).reshape(6, n)
    for r, s in enumerate([1e30, 1e-30, 1.0, 0.0, 3e4, 1e-6]):
        rows.append(base[r] * s)
    spike = np.full(n, 1e-12 + 1e-12j)
    spike[n // 3] = 1e20
    rows.append(spike)
    x = np.concatenate(rows).astype(np.complex64)
    count = len(rows)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run_batch(n, count, x, topefft.PREC_F32, topefft.MODE_MFMA_FP16X3, d).reshape(count, n)
        got = got.astype(np.complex128)
        for r in range(count):
            xr = x[r * n:(r + 1) * n].astype(np.complex128)
            want = oracle.dft(xr, (n,), d)
            if not np.any(xr):
                assert not np.any(got[r]), r
                continue
            err = rel_l2(got[r], want)
            print(f"n={n} dir={d} row {r} rel-L2 {err:.3e}")
            assert err < 1e-6, (r, err)
            assert np.all(np.isfinite(got[r]))


def test_batch_round_trip_mfma():
    n, count = 4096, 64
    x = batch_input(n, count, 7).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, count, x, topefft.PREC_F32, topefft.MODE_MFMA_BF16X6)
    p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert rel_l2(back.astype(np.complex128), x.astype(np.complex128)) < 2e-6


@pytest.mark.parametrize("n", [12, 100, 1000, 3000, 6000, 8192, 16384])
def test_batch_complex64_mixed_radix(oracle, n):
    # complex64 batch of 7-smooth lengths (the mixed-radix pass in FP32, the
    # mixed four-step above 4096) and of powers of two above 4096 (the
    # four-step in FP32)
    count = 10
    x = batch_input(n, count, 300 + n).astype(np.complex64)
    x64 = x.astype(np.complex128)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got = run_batch(n, count, x, topefft.PREC_F32, topefft.MODE_VALU, d).reshape(count, n).astype(np.complex128)
        for r in (0, 4, count - 1):
            assert rel_l2(got[r], oracle.dft(x64[r * n:(r + 1) * n], (n,), d)) < 1e-6, (n, d, r)


def test_batch_rejects_unsupported():
    fw = topefft.Framework()
    x = np.zeros(128 * 2, np.complex64)
    with pytest.raises(topefft.TopeError) as e:
        topefft.BatchPlan(fw, 128, 2, x, topefft.PREC_F32, topefft.MODE_MFMA_BF16X6)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    with pytest.raises(topefft.TopeError):
        topefft.BatchPlan(fw, 256, 2, np.zeros(512, complex), topefft.PREC_F64, topefft.MODE_MFMA_FP32)
    # complex64 batches: powers of two and 7-smooth lengths on the vector
    # ALUs only
    with pytest.raises(topefft.TopeError) as e:
        topefft.BatchPlan(fw, 11, 2, np.zeros(22, np.complex64), topefft.PREC_F32, topefft.MODE_VALU)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    with pytest.raises(topefft.TopeError) as e:
        topefft.BatchPlan(fw, 100, 2, np.zeros(200, np.complex64), topefft.PREC_F32, topefft.MODE_MFMA_FP16X3)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    # Bluestein rows of M = 2^k >= 2n - 1 up to 2^30 points: n = 2^29 + 1 is
    # refused at planning, before the input (a few bytes here) is read
    t = topefft.TopePlanBatch()
    tiny = np.zeros(4, complex)
    fw.lib.topeBatchPlanInit(C.byref(fw.f), C.byref(t), (1 << 29) + 1, 1, topefft.PREC_F64, topefft.MODE_VALU,
                             tiny.ctypes.data_as(C.c_void_p))
    assert fw.f.error == topefft.TOPE_ERROR_INVALID_SIZE
    fw.destroy()


def test_batch_long_bluestein(oracle):
    # FP64 batches of non-power-of-two lengths whose Bluestein rows are
    # longer than one pass (M = 8192: the inner FFTs run the four-step)
    n, count = 2053, 3
    rng = np.random.default_rng(2053)
    x = rng.standard_normal(n * count) + 1j * rng.standard_normal(n * count)
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, count, x, topefft.PREC_F64, topefft.MODE_VALU)
    X = p.exec(topefft.FORWARD)
    p.destroy()
    for i in range(count):
        seg = slice(i * n, (i + 1) * n)
        assert rel_l2(X[seg], oracle.dft(x[seg], (n,), topefft.FORWARD)) < 1e-12
