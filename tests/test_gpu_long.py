"""Axes of 4096^2 points and longer against the oracle.

Power-of-two lengths above 2^24 run three passes (n = A B C, each factor
at most 4096: a post-twiddled pass over the outer index, the four-step of
the inner A B points); non-power-of-two lengths whose Bluestein row
(M >= 2n - 1) is longer than 2^24 run their inner FFTs that way.  The
transforms are too long for the full naive DFT, so each is checked at
sampled bins against the oracle's direct DFT (FP64: rel-L2 <= 1e-12;
complex64 storage <= 1e-6), by Parseval and by the round trip on the
device; the pass lists are asserted.  The 2^27-point FP64 array (2 GiB)
takes the 64-bit addressing variants of the passes, and so do the column
tiles of 40000 x 4096 FP64 and 65536 x 4096 complex64.
"""
import ctypes as C

import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def seeded(n, salt):
    return workloads.complex_uniform(0x10A6 + salt * 7919 + n % 100003, n)


def sample_bins(n, seed, k=10):
    rng = np.random.default_rng(seed)
    return np.concatenate([[0, 1, n - 1, n // 2, n // 3], rng.integers(0, n, k)])


def check_1d(oracle, n, x, salt, tol=TOL, nbins=10):
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    bins = sample_bins(n, salt, nbins)
    assert rel_l2(X[bins], oracle.dft_bins(x, (n, 1, 1), bins, topefft.FORWARD)) < tol
    ex, eX = np.vdot(x, x).real, np.vdot(X, X).real
    assert abs(eX - n * ex) / (n * ex) < tol
    assert rel_l2(back, x) < tol
    return names


def test_1d_pow2_four_step_4096(oracle):
    # 2^24 = 4096 x 4096: the first step runs the three-step column tile
    # with the four-step post-twiddle (strided input and output), the
    # second the one-shot column tile (contiguous input, transposed output)
    n = 1 << 24
    names = check_1d(oracle, n, seeded(n, 6), 6)
    assert names == ["axis_double_n4096_e16_tb4_col_ct3", "axis_double_n4096_e16_tb4_col_split"], names


def test_1d_pow2_three_pass(oracle):
    n = 1 << 25  # 512 x 256 x 256
    names = check_1d(oracle, n, seeded(n, 1), 1)
    assert len(names) == 3 and "_n512_" in names[0] and "_n256_" in names[1] and "_n256_" in names[2], names


def test_1d_pow2_three_pass_64bit_addressing(oracle):
    # 2^27 points (2 GiB FP64): 512 x 512 x 512, spans beyond 2^31 bytes
    n = 1 << 27
    names = check_1d(oracle, n, seeded(n, 2), 2, nbins=4)
    assert len(names) == 3 and all("_n512_" in nm for nm in names), names


def test_batch_three_pass(oracle):
    n, count = 1 << 25, 2
    x = seeded(n * count, 3)
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, count, x, topefft.PREC_F64, topefft.MODE_VALU)
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    for r in range(count):
        seg = slice(r * n, (r + 1) * n)
        bins = sample_bins(n, 30 + r, 6)
        assert rel_l2(X[seg][bins], oracle.dft_bins(x[seg], (n, 1, 1), bins, topefft.FORWARD)) < TOL, r
    assert rel_l2(back, x) < TOL


def test_batch_three_pass_complex64(oracle):
    n = 1 << 25
    x = seeded(n, 4).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, 1, x, topefft.PREC_F32, topefft.MODE_VALU)
    X = p.exec(topefft.FORWARD).astype(np.complex128)
    back = p.exec(topefft.INVERSE).astype(np.complex128)
    p.destroy()
    fw.destroy()
    x64 = x.astype(np.complex128)
    bins = sample_bins(n, 40, 10)
    err = rel_l2(X[bins], oracle.dft_bins(x64, (n, 1, 1), bins, topefft.FORWARD))
    print(f"complex64 2^25 rel-L2 at sampled bins {err:.3e}")
    assert err < 1e-6
    assert rel_l2(back, x64) < 1e-6


def test_batch_four_step_4096_complex64(oracle):
    # complex64 2^24 = 4096 x 4096 (FP32 column tiles with the post-twiddle)
    n = 1 << 24
    x = seeded(n, 7).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.BatchPlan(fw, n, 1, x, topefft.PREC_F32, topefft.MODE_VALU)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD).astype(np.complex128)
    back = p.exec(topefft.INVERSE).astype(np.complex128)
    p.destroy()
    fw.destroy()
    assert len(names) == 2 and all("float_n4096" in nm for nm in names), names
    x64 = x.astype(np.complex128)
    bins = sample_bins(n, 70, 10)
    err = rel_l2(X[bins], oracle.dft_bins(x64, (n, 1, 1), bins, topefft.FORWARD))
    print(f"complex64 2^24 rel-L2 at sampled bins {err:.3e}")
    assert err < 1e-6
    assert rel_l2(back, x64) < 1e-6


def test_1d_bluestein_three_pass(oracle):
    # 16777219 = 1549 x 10831: Bluestein row M = 2^26 (inner FFTs in three passes)
    n = (1 << 24) + 3
    names = check_1d(oracle, n, seeded(n, 5), 5, nbins=6)
    assert "chirp_pre" in names and "chirp_post" in names, names


def test_2d_column_tile_64bit_addressing(oracle):
    # 2D 40000 x 4096 (2.4 GiB FP64; the rows run the mixed-radix
    # four-step, 200 x 200): the 4096-point column pass spans more than
    # 2^31 bytes per tile (rows 625 KiB apart), so it runs the three-step
    # tile's 64-bit addressing instantiation
    dims = (40000, 4096)
    n = dims[0] * dims[1]
    x = seeded(n, 9)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert names[-1] == "axis_double_n4096_e16_tb4_col_ct3", names
    bins = sample_bins(n, 90, 4)
    assert rel_l2(X[bins], oracle.dft_bins(x, (dims[0], dims[1], 1), bins, topefft.FORWARD)) < TOL
    assert rel_l2(back, x) < TOL


def test_exec_ptr_three_pass_matches_plan_exec():
    # topePlanExecPtr on a caller array (a torch tensor): the three passes
    # run in place on it and end in the plan's second array, then one copy
    # brings the result back; bit-identical to the plan's own exec
    torch = pytest.importorskip("torch")
    n = 1 << 25
    x = seeded(n, 8)
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    want = p.exec(topefft.FORWARD)
    q = topefft.Plan(fw, (n,), np.zeros(n, complex))
    t = torch.from_numpy(x.copy()).to("cuda")
    torch.cuda.synchronize()
    q.exec_ptr(t.data_ptr(), topefft.FORWARD)
    fw.synchronize()
    got = t.cpu().numpy()
    assert not np.any(q.download())  # the plan's own array is untouched
    q.destroy()
    p.destroy()
    fw.destroy()
    assert np.array_equal(got, want)


def test_refuses_beyond_the_three_pass_limit():
    # a Bluestein row of 2^31 points (n = 2^29 + 1) is refused at planning,
    # before anything is allocated or the input is read (the buffer passed
    # is a few bytes)
    fw = topefft.Framework()
    t = topefft.TopePlan1D()
    tiny = np.zeros(4)
    fw.lib.tope1DPlanInit(C.byref(fw.f), C.byref(t), (1 << 29) + 1, topefft.C2C,
                          tiny.ctypes.data_as(C.POINTER(C.c_double)))
    assert fw.f.error == topefft.TOPE_ERROR_INVALID_SIZE
    fw.destroy()


def test_2d_pair_tile_64bit_addressing_complex64(oracle):
    # complex64 65536 x 4096 (2 GiB): the 4096-point column pass spans more
    # than 2^31 bytes per tile (rows 512 KiB apart), so it runs the 64-bit
    # addressing instantiation of the paired-column three-step tile; the
    # rows run the four-step; sampled bins against the oracle's direct DFT
    # (complex64 storage: rel-L2 <= 1e-6) and the device round trip
    dims = (65536, 4096)
    n = dims[0] * dims[1]
    x = seeded(n, 11).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert names[-1] == "axis_float_n4096_e16_tb8_col_ct3", names
    bins = sample_bins(n, 110, 4)
    want = oracle.dft_bins(x[: n].astype(np.complex128), (dims[0], dims[1], 1), bins, topefft.FORWARD)
    err = rel_l2(X[bins].astype(np.complex128), want)
    print(f"65536 x 4096 complex64 rel-L2 at sampled bins {err:.3e}")
    assert err < 1e-6
    assert rel_l2(back.astype(np.complex128), x.astype(np.complex128)) < 1e-6
