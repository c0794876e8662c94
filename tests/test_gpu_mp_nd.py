"""Complex64 2D/3D plans with the axis passes on the matrix cores
(tope2DPlanInitMode / tope3DPlanInitMode) against the FP64 oracle.

Contiguous rows of 256 or 4096 run the MFMA row kernel, unit-stride column
axes of 256, 512 or 1024 with a multiple of 16 lines run the 16-line column
tile kernel; other axes fall back to the FP32 VALU pass.  Bars as for the
batch modes: 3-part bf16, fp16 split and FP32 operands <= 1e-6 rel-L2 (FP32
storage), the 2-part bf16 split <= 1e-4.  Inputs: seeded Gaussian re/im.
"""
import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu

MODES = {topefft.MODE_MFMA_FP16X3: 1e-6, topefft.MODE_MFMA_BF16X6: 1e-6, topefft.MODE_MFMA_FP32: 1e-6,
         topefft.MODE_MFMA_BF16X3: 1e-4}

# (dims, expected kernel family per pass in order)
SHAPES = [
    ((256, 256), ["mfma_rows", "mfma_cols"]),
    ((64, 512), ["axis_float", "mfma_cols"]),
    ((32, 1024), ["axis_float", "mfma_cols"]),
    ((4096, 16), ["mfma_rows", "axis_float"]),
    ((8, 512), ["axis_float", "axis_float"]),          # 8 lines: no 16-line tile
    ((16, 16, 512), ["axis_float", "axis_float", "mfma_cols"]),
    ((32, 256, 16), ["axis_float", "mfma_cols", "axis_float"]),
    ((16, 16, 1024), ["axis_float", "axis_float", "mfma_cols"]),
    ((512, 32), ["mfma_rows", "axis_float"]),
    ((1024, 16), ["mfma_rows", "axis_float"]),
    ((2048, 16), ["mfma_rows", "axis_float"]),
    ((512, 16, 512), ["mfma_rows", "axis_float", "mfma_cols"]),
]

_want = {}


def oracle_cached(oracle, x64, dims, d):
    key = (dims, d)
    if key not in _want:
        _want[key] = oracle.dft(x64, dims, d)
    return _want[key]


def run(dims, x, mode, direction):
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32, mode=mode)
    try:
        return p.exec(direction), p.pass_names()
    finally:
        p.destroy()
        fw.destroy()


@pytest.mark.parametrize("mode", list(MODES), ids=lambda m: topefft.MODE_NAMES[m])
@pytest.mark.parametrize("dims,kinds", SHAPES, ids=lambda v: "x".join(map(str, v)) if isinstance(v[0], int) else "")
def test_mp_nd_modes(oracle, dims, kinds, mode):
    n = int(np.prod(dims))
    x = workloads.complex_gaussian(900 + n % 97, n).astype(np.complex64)
    x64 = x.astype(np.complex128)
    for d in (topefft.FORWARD, topefft.INVERSE):
        got, names = run(dims, x, mode, d)
        assert [nm.startswith(k) for nm, k in zip(names, kinds)] == [True] * len(kinds), names
        want = oracle_cached(oracle, x64, dims, d)
        err = rel_l2(got.astype(np.complex128), want)
        print(f"{dims} {topefft.MODE_NAMES[mode]} dir={d} passes={names} rel-L2 {err:.3e}")
        assert err < MODES[mode]


@pytest.mark.parametrize("scale", [1e30, 1e-30, 0.0])
def test_mp_nd_fp16x3_range(oracle, scale):
    # per-line power-of-two scaling keeps FP32-range inputs exact in range
    dims = (32, 512)
    n = int(np.prod(dims))
    x = (workloads.complex_gaussian(77, n) * scale).astype(np.complex64)
    got, names = run(dims, x, topefft.MODE_MFMA_FP16X3, topefft.FORWARD)
    assert names[1].startswith("mfma_cols")
    got = got.astype(np.complex128)
    assert np.all(np.isfinite(got))
    if scale == 0.0:
        assert not np.any(got)
        return
    want = oracle.dft(x.astype(np.complex128), dims, topefft.FORWARD)
    assert rel_l2(got, want) < 1e-6


def test_mp_nd_round_trip_3d():
    dims = (64, 64, 512)
    n = int(np.prod(dims))
    x = workloads.complex_gaussian(5, n).astype(np.complex64)
    fw = topefft.Framework()
    p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32, mode=topefft.MODE_MFMA_FP16X3)
    p.exec_device(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    names = p.pass_names()
    p.destroy()
    fw.destroy()
    assert names[2].startswith("mfma_cols")
    assert rel_l2(back.astype(np.complex128), x.astype(np.complex128)) < 2e-6


def test_mp_nd_rejects_fp64_modes():
    fw = topefft.Framework()
    with pytest.raises(topefft.TopeError) as e:
        topefft.Plan(fw, (256, 256), np.zeros(65536, np.complex128), precision=topefft.PREC_F64,
                     mode=topefft.MODE_MFMA_FP16X3)
    assert e.value.code == topefft.TOPE_ERROR_UNSUPPORTED
    fw.destroy()
