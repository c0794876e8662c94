"""Every compiled axis-pass kernel variant against the oracle.

The planner picks one variant per (length, layout) from the shape; most
shapes only ever reach a few of them.  Here each variant is forced through
TOPEFFT_VARIANT_<n>_<row|col>=<name> onto a small 2D shape whose axis of
length n runs it -- once with a line count that fills whole tiles and once
with a partial last tile -- and the result is compared with the naive DFT
(FP64: rel-L2 <= 1e-12; complex64 storage: <= 1e-6), plus the round trip
on the device.  The pass names are asserted, so a variant that the planner
silently replaced would fail.  Then the tile choices the planner makes by
itself for narrow shapes (single-column tiles) are pinned at their natural
shapes, and the long 1D four-step lengths that run them.
"""
import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _variants():
    try:
        return [v for v in topefft.axis_variants() if not v[7]]
    except OSError:
        return []


VARIANTS = _variants()


def _seeded(dims, salt):
    h = 0x5A17 + salt
    for d in dims:
        h = (h * 1000003 + d) & 0xFFFFFFFF
    return workloads.complex_uniform(h, int(np.prod(dims)))


def _line_counts(tb):
    # a whole number of tiles, and a partial last tile (the extra lines make
    # the other axis 7-smooth, so it runs the mixed-radix pass)
    full = max(2, 2 * tb)
    part = 3 if tb <= 2 else tb + tb // 2
    return full, part


@pytest.mark.parametrize("var", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_forced_variant(oracle, monkeypatch, var):
    name, prec, n, e, tb, col, split, _ = var
    monkeypatch.setenv(f"TOPEFFT_VARIANT_{n}_{'col' if col else 'row'}", name)
    tol = 1e-12 if prec == topefft.PREC_F64 else 1e-6
    for lines in _line_counts(tb):
        # columns: the Y axis (length n) has `lines` columns; rows: the X
        # axis (length n) has `lines` rows
        dims = (lines, n) if col else (n, lines)
        x = _seeded(dims, tb)
        fw = topefft.Framework()
        if prec == topefft.PREC_F64:
            p = topefft.Plan(fw, dims, x)
        else:
            x = x.astype(np.complex64)
            p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32)
        names = p.pass_names()
        X = p.exec(topefft.FORWARD).astype(np.complex128)
        back = p.exec(topefft.INVERSE).astype(np.complex128)
        p.destroy()
        fw.destroy()
        assert name in names, (dims, names)
        x64 = x.astype(np.complex128)
        assert rel_l2(X, oracle.dft(x64, dims, topefft.FORWARD)) < tol, (name, dims)
        assert rel_l2(back, x64) < tol, (name, dims)


def test_every_preferred_variant_is_compiled():
    # the planner's preference table names variants by string: each must
    # exist, or the preference silently falls through
    names = {v[0] for v in topefft.axis_variants()}
    for want in ("axis_double_n4096_e16_tb1_row_split", "axis_double_n4096_e16_tb4_col_split",
                 "axis_double_n4096_e16_tb4_col_ct3",
                 "axis_double_n2048_e16_tb4_col_split", "axis_double_n256_e16_tb32_col_split",
                 "axis_double_n512_e16_tb16_col_split", "axis_double_n256_e16_tb4_row",
                 "axis_float_n512_e16_tb32_col_split", "axis_float_n256_e16_tb32_col_split",
                 "axis_float_n4096_e16_tb1_row_split", "axis_float_n256_e16_tb4_row",
                 "axis_float_n4096_e16_tb8_col_ct3"):
        assert want in names, want


# narrow shapes take single-column tiles by themselves (the only tile width
# that gives every CU a block)
NARROW = [((16, 2048), topefft.PREC_F64, "axis_double_n2048_e16_tb1_col"),
          ((32, 512), topefft.PREC_F64, "axis_double_n512_e8_tb1_col"),
          ((64, 1024), topefft.PREC_F64, "axis_double_n1024_e16_tb1_col"),
          ((8, 2048), topefft.PREC_F32, "axis_float_n2048_e16_tb1_col"),
          ((64, 256), topefft.PREC_F64, "axis_double_n256_e4_tb1_col")]


@pytest.mark.parametrize("dims,prec,kernel", NARROW, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_narrow_single_column_tiles(oracle, dims, prec, kernel):
    x = _seeded(dims, 77)
    fw = topefft.Framework()
    if prec == topefft.PREC_F32:
        x = x.astype(np.complex64)
        p = topefft.Plan(fw, dims, x, precision=topefft.PREC_F32)
        tol = 1e-6
    else:
        p = topefft.Plan(fw, dims, x)
        tol = 1e-12
    names = p.pass_names()
    X = p.exec(topefft.FORWARD).astype(np.complex128)
    back = p.exec(topefft.INVERSE).astype(np.complex128)
    p.destroy()
    fw.destroy()
    assert kernel in names, names
    x64 = x.astype(np.complex128)
    assert rel_l2(X, oracle.dft(x64, dims, topefft.FORWARD)) < tol
    assert rel_l2(back, x64) < tol
    if prec == topefft.PREC_F64:
        Xi = topefft.transform(x64, dims, topefft.INVERSE)
        assert rel_l2(Xi, oracle.dft(x64, dims, topefft.INVERSE)) < tol


@pytest.mark.parametrize("logn", [17, 18])
def test_1d_four_step_long(oracle, logn):
    # 2^17 = 512 x 256 and 2^18 = 512 x 512: the four-step's column step runs
    # the 512-point single-column tile; sampled bins against the direct DFT,
    # both directions, and the round trip
    n = 1 << logn
    x = _seeded((n,), 5)
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), x)
    names = p.pass_names()
    X = p.exec(topefft.FORWARD)
    back = p.exec(topefft.INVERSE)
    p.destroy()
    fw.destroy()
    assert "axis_double_n512_e8_tb1_col" in names, names
    rng = np.random.default_rng(logn)
    bins = np.concatenate([[0, 1, n // 2, n - 1], rng.integers(0, n, 60)])
    assert rel_l2(X[bins], oracle.dft_bins(x, (n,), bins, topefft.FORWARD)) < 1e-12
    assert rel_l2(back, x) < 1e-12
    Xi = topefft.transform(x, (n,), topefft.INVERSE)
    assert rel_l2(Xi[bins], oracle.dft_bins(x, (n,), bins, topefft.INVERSE)) < 1e-12
