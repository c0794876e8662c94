"""Caller streams and caller-owned device arrays (topeFFTSetStream,
topePlanExecPtr): torch CUDA tensors transformed in place on torch's
current stream, against the FP64 oracle (rel-L2 <= 1e-12), for an in-place
schedule (2D 64 x 32) and one whose last step is out of place (1D 2^15,
four-step), and a complex64 plan."""
import numpy as np
import pytest

import topefft
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


def test_exec_ptr_errors():
    fw = topefft.Framework()
    p = topefft.Plan(fw, (16, 16), np.zeros(256, complex))
    with pytest.raises(topefft.TopeError):
        p.exec_ptr(0, topefft.FORWARD)
    p.destroy()
    fw.destroy()
