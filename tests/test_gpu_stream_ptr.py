"""Caller streams and caller-owned device arrays (topeFFTSetStream,
topePlanExecPtr, topeRealExecPtr, topeManyExecPtr): torch CUDA tensors
transformed in place on torch's current stream, against the FP64 oracle
(rel-L2 <= 1e-12), for an in-place schedule (2D 64 x 32) and one whose last
step is out of place (1D 2^15, four-step), a complex64 plan, real-input
plans and packed uneven batches under every method."""
import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    # imported only when a GPU test runs (not at collection)
    return pytest.importorskip("torch")


@pytest.mark.parametrize("dims", [(64, 32), (1 << 15,), (16, 8, 4), (8192, 4), (4, 8, 8192)], ids=lambda d: "x".join(map(str, d)))
def test_exec_ptr_torch_stream(oracle, torch, dims):
    n = int(np.prod(dims))
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, np.zeros(n, complex))
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        t = torch.from_numpy(x.copy()).to("cuda")
        fw.set_stream(torch.cuda.current_stream().cuda_stream)
        p.exec_ptr(t.data_ptr(), topefft.FORWARD)
        X = t.cpu().numpy()  # ordered after the transform on the same stream
        p.exec_ptr(t.data_ptr(), topefft.INVERSE)
        back = t.cpu().numpy()
    fw.set_stream(None)
    # the plan's own array is untouched (still zeros)
    assert not np.any(p.download())
    p.destroy()
    fw.destroy()
    assert rel_l2(X, oracle.dft(x, dims, topefft.FORWARD)) < 1e-12
    assert rel_l2(back, x) < 1e-12


def test_exec_ptr_complex64(oracle, torch):
    dims = (256, 256)
    n = 256 * 256
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, np.zeros(n, np.complex64), precision=topefft.PREC_F32,
                     mode=topefft.MODE_MFMA_FP16X3)
    t = torch.from_numpy(x.copy()).to("cuda")
    torch.cuda.synchronize()
    p.exec_ptr(t.data_ptr(), topefft.FORWARD)
    fw.synchronize()
    X = t.cpu().numpy().astype(np.complex128)
    p.destroy()
    fw.destroy()
    assert rel_l2(X, oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD)) < 1e-6


def test_exec_ptr_complex64_column_pairs(oracle, torch, monkeypatch):
    # the complex64 three-step column tile moves two adjacent columns as one
    # 16-byte element: on a caller array it matches the plan's own exec bit
    # for bit, and an array that is only 8-byte aligned is refused before
    # anything is enqueued
    monkeypatch.setenv("TOPEFFT_VARIANT_4096_col", "axis_float_n4096_e16_tb8_col_ct3")
    dims = (16, 4096)
    n = 16 * 4096
    rng = np.random.default_rng(16)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32)
    assert p.pass_names()[-1] == "axis_float_n4096_e16_tb8_col_ct3", p.pass_names()
    want = p.exec(topefft.FORWARD)
    t = torch.from_numpy(np.concatenate([x, np.zeros(1, np.complex64)])).to("cuda")
    torch.cuda.synchronize()
    p.exec_ptr(t.data_ptr(), topefft.FORWARD)
    fw.synchronize()
    got = t.cpu().numpy()
    with pytest.raises(topefft.TopeError) as e:
        p.exec_ptr(t.data_ptr() + 8, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    fw.synchronize()
    assert np.array_equal(t.cpu().numpy(), got)  # nothing ran
    p.destroy()
    fw.destroy()
    assert np.array_equal(got[:n], want)
    assert got[n] == 0
    assert rel_l2(want.astype(np.complex128), oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD)) < 1e-6


def test_exec_ptr_errors():
    fw = topefft.Framework()
    p = topefft.Plan(fw, (16, 16), np.zeros(256, complex))
    with pytest.raises(topefft.TopeError):
        p.exec_ptr(0, topefft.FORWARD)
    fw.synchronize()
    # an array not aligned to the 16-byte element is refused before anything
    # is enqueued (the plan's own array, offset by 8 bytes, stays zero)
    with pytest.raises(topefft.TopeError) as e:
        p.exec_ptr(p.t.data + 8, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    fw.synchronize()
    assert not np.any(p.download())
    p.destroy()
    fw.destroy()


@pytest.mark.parametrize("dims", [(256, 64), (64, 8, 4), (8192,), (16, 16), (32768,)], ids=lambda d: "x".join(map(str, d)))
def test_real_exec_ptr(oracle, torch, dims):
    # topeRealExecPtr on torch arrays in the plan's device layout; the
    # output arrays start as NaN, so every element the result needs must be
    # written, and the plan's own arrays stay zero
    n = int(np.prod(dims))
    nx = dims[0]
    rows = n // nx
    x = np.random.default_rng(n + 7).uniform(-1.0, 1.0, n)
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, dims, np.zeros(n))
    pitch = p.t.pitch
    assert p.t.dataSize == pitch * rows * 16 and p.t.realSize == n * 8
    real = torch.from_numpy(x.copy()).to("cuda")
    spec = torch.full((pitch * rows,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    p.exec_ptr(real.data_ptr(), spec.data_ptr(), topefft.FORWARD)
    fw.synchronize()
    X = spec.cpu().numpy().reshape(rows, pitch)[:, : nx // 2 + 1].reshape(-1)
    back_t = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    p.exec_ptr(back_t.data_ptr(), spec.data_ptr(), topefft.INVERSE)
    fw.synchronize()
    back = back_t.cpu().numpy()
    own = p.forward()
    p.destroy()
    fw.destroy()
    full = oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD).reshape(rows, nx)
    assert rel_l2(X, full[:, : nx // 2 + 1].reshape(-1)) < 1e-12
    assert rel_l2(back, x) < 1e-12
    assert not np.any(own)


def test_many_exec_ptr(oracle, torch):
    # topeManyExecPtr with every method (the graph methods launch without
    # their graph): identical results, the oracle on two transforms per
    # length, and the plan's own array untouched
    sizes = workloads.uneven_sizes(0x5EED, 60, 4, 14)
    total = int(sizes.sum())
    x = workloads.complex_uniform(0x5EEE, total)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    outs = {}
    for m in (topefft.MANY_PERSISTENT, topefft.MANY_PERSISTENT_GRAPH, topefft.MANY_PER_TRANSFORM,
              topefft.MANY_PER_TRANSFORM_GRAPH):
        fw = topefft.Framework()
        q = topefft.ManyPlan(fw, sizes, np.zeros(total, complex))
        q.set_method(m)
        t = torch.from_numpy(x.copy()).to("cuda")
        torch.cuda.synchronize()
        q.exec_ptr(t.data_ptr(), topefft.FORWARD)
        fw.synchronize()
        fwd = t.cpu().numpy()
        q.exec_ptr(t.data_ptr(), topefft.INVERSE)
        fw.synchronize()
        outs[m] = (fwd, t.cpu().numpy())
        assert not np.any(q.exec(topefft.FORWARD)), topefft.MANY_METHOD_NAMES[m]
        q.destroy()
        fw.destroy()
    fwd, back = outs[topefft.MANY_PERSISTENT]
    seen = {}
    for n, o in zip(sizes, offs):
        if seen.get(n, 0) < 2:
            seen[n] = seen.get(n, 0) + 1
            assert rel_l2(fwd[o:o + n], oracle.dft(x[o:o + n], (int(n),), topefft.FORWARD)) < 1e-12, n
    assert rel_l2(back, x) < 1e-12
    for m, (f, b) in outs.items():
        assert np.array_equal(f, fwd) and np.array_equal(b, back), topefft.MANY_METHOD_NAMES[m]


def test_real_and_many_exec_ptr_errors():
    fw = topefft.Framework()
    p = topefft.RealPlan(fw, (16, 4), np.zeros(64))
    with pytest.raises(topefft.TopeError):
        p.exec_ptr(0, p.t.data, topefft.FORWARD)
    with pytest.raises(topefft.TopeError):
        p.exec_ptr(p.t.real, p.t.data, 7)
    # arrays not aligned to the complex element are refused
    with pytest.raises(topefft.TopeError) as e:
        p.exec_ptr(p.t.real + 8, p.t.data, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    with pytest.raises(topefft.TopeError) as e:
        p.exec_ptr(p.t.real, p.t.data + 8, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    q = topefft.ManyPlan(fw, np.array([16, 32], np.int32), np.zeros(48, complex))
    with pytest.raises(topefft.TopeError):
        q.exec_ptr(0, topefft.FORWARD)
    with pytest.raises(topefft.TopeError) as e:
        q.exec_ptr(q.t.data + 8, topefft.FORWARD)
    assert e.value.code == topefft.TOPE_ERROR_INVALID_ARG
    q.destroy()
    p.destroy()
    fw.destroy()
