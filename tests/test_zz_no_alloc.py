"""The allocation-free entry point (topeExec1DNoAlloc) against the existing
one it varies (topePlanExecPtr on a 1D plan of the same length, applied to
each transform of the array in turn), through the C-ABI, on caller-owned
device arrays (torch tensors): FP64, rel-L2 <= 1e-12, both directions.

Two calls in a row, the second with a smaller input (the static scratch
array keeps the first call's tables and intermediate array), a call at the
scratch limit, a rejected call above it (TOPE_ERROR_SCRATCH_LIMIT, nothing
launched) followed by a good one, and the oracle on a small case."""
import ctypes as C

import numpy as np
import pytest

import topefft
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    return pytest.importorskip("torch")


def _data(n, count, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n * count) + 1j * rng.standard_normal(n * count)


def _noalloc(torch, fw, x, n, count, direction):
    t = torch.from_numpy(x.copy()).to("cuda")
    torch.cuda.synchronize()
    fw.lib.topeExec1DNoAlloc(C.byref(fw.f), C.c_void_p(t.data_ptr()), n, count, direction)
    fw.check(f"topeExec1DNoAlloc({n}, {count})")
    fw.synchronize()
    return t.cpu().numpy()


def _plan_ptr(torch, x, n, count, direction):
    # the existing entry point: a 1D plan's topePlanExecPtr on each transform
    # (a framework of its own: destroying its last plan releases it)
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), np.zeros(n, complex))
    t = torch.from_numpy(x.copy()).to("cuda")
    torch.cuda.synchronize()
    for i in range(count):
        p.exec_ptr(t.data_ptr() + i * n * 16, direction)
    fw.synchronize()
    out = t.cpu().numpy()
    p.destroy()
    return out


@pytest.mark.parametrize("direction", [topefft.FORWARD, topefft.INVERSE], ids=["forward", "inverse"])
def test_noalloc_matches_plan_exec_ptr(torch, direction):
    fw = topefft.Framework()
    # first call: four-step length (intermediate array in the scratch); second:
    # a smaller one-pass length
    for n, count, seed in ((1 << 16, 8, 1), (1 << 12, 3, 2), (1 << 10, 5, 3)):
        x = _data(n, count, seed)
        got = _noalloc(torch, fw, x, n, count, direction)
        want = _plan_ptr(torch, x, n, count, direction)
        assert rel_l2(got, want) < 1e-12, (n, count)
    fw.destroy()


def test_noalloc_against_oracle(torch, oracle):
    fw = topefft.Framework()
    n, count = 4096, 3
    x = _data(n, count, 7)
    got = _noalloc(torch, fw, x, n, count, topefft.FORWARD)
    for i in range(count):
        seg = slice(i * n, (i + 1) * n)
        assert rel_l2(got[seg], oracle.dft(x[seg], (n,), topefft.FORWARD)) < 1e-12
    fw.destroy()


def test_noalloc_scratch_limit(torch):
    fw = topefft.Framework()
    lim = topefft.NOALLOC_MAX_ELEMENTS
    # at the limit: 2^20-point transforms, two of them
    n = 1 << 20
    x = _data(n, lim // n, 11)
    got = _noalloc(torch, fw, x, n, lim // n, topefft.FORWARD)
    want = _plan_ptr(torch, x, n, lim // n, topefft.FORWARD)
    assert rel_l2(got, want) < 1e-12
    # above it: refused before anything is launched, the array unchanged
    t = torch.zeros(2 * (lim + n), dtype=torch.float64, device="cuda")
    fw.lib.topeExec1DNoAlloc(C.byref(fw.f), C.c_void_p(t.data_ptr()), n, lim // n + 1, topefft.FORWARD)
    assert fw.f.error == topefft.TOPE_ERROR_SCRATCH_LIMIT
    assert fw.lib.topeErrorString(fw.f.error)
    fw.f.error = 0
    torch.cuda.synchronize()
    assert not bool(torch.any(t))
    # an array not aligned to the 16-byte element: refused, nothing launched
    fw.lib.topeExec1DNoAlloc(C.byref(fw.f), C.c_void_p(t.data_ptr() + 8), 1 << 13, 1, topefft.FORWARD)
    assert fw.f.error == topefft.TOPE_ERROR_INVALID_ARG
    fw.f.error = 0
    torch.cuda.synchronize()
    assert not bool(torch.any(t))
    # a good call after the refused ones
    m = 1 << 13
    y = _data(m, 4, 12)
    got = _noalloc(torch, fw, y, m, 4, topefft.INVERSE)
    want = _plan_ptr(torch, y, m, 4, topefft.INVERSE)
    assert rel_l2(got, want) < 1e-12
    fw.destroy()
