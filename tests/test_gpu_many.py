"""Packed batch of uneven power-of-two transforms (persistent kernel, graph
replay, per-transform launches, autotuned schedules) against the oracle.

Bars: FP64 rel-L2 <= 1e-12 per transform against the naive DFT; the four
execution methods bit-identical (same device code, same arithmetic order);
complex64 within 1e-6 of the FP64 result."""
import numpy as np
import pytest

import topefft
import workloads
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def make_batch(count, seed, lmin=4, lmax=14):
    sizes = workloads.uneven_sizes(seed, count, lmin, lmax)
    x = workloads.complex_uniform(seed + 1, int(sizes.sum()))
    return sizes, x


def check_against_oracle(oracle, sizes, x, out, direction, per_class=2, tol=1e-12):
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    seen = {}
    for i, (n, o) in enumerate(zip(sizes, offs)):
        if seen.get(n, 0) >= per_class:
            continue
        seen[n] = seen.get(n, 0) + 1
        want = oracle.dft(x[o:o + n], (int(n),), direction)
        assert rel_l2(out[o:o + n], want) < tol, (i, n)
    return seen


def test_many_fp64_all_methods(oracle):
    sizes, x = make_batch(400, 0xBA7C)
    outs = {}
    for m in (topefft.MANY_PERSISTENT, topefft.MANY_PERSISTENT_GRAPH, topefft.MANY_PER_TRANSFORM,
              topefft.MANY_PER_TRANSFORM_GRAPH):
        fw = topefft.Framework()
        q = topefft.ManyPlan(fw, sizes, x)
        assert np.array_equal(q.offsets(), np.concatenate([[0], np.cumsum(sizes)[:-1]]))
        q.set_method(m)
        outs[m] = (q.exec(topefft.FORWARD), q.exec(topefft.INVERSE))
        q.destroy()
    fwd, back = outs[topefft.MANY_PERSISTENT]
    seen = check_against_oracle(oracle, sizes, x, fwd, 1)
    assert set(seen) == set(int(v) for v in np.unique(sizes))
    assert rel_l2(back, x) < 1e-12
    for m, (f, b) in outs.items():
        assert np.array_equal(f, fwd), topefft.MANY_METHOD_NAMES[m]
        assert np.array_equal(b, back), topefft.MANY_METHOD_NAMES[m]


def poison_device_memory(mib=256):
    # a plan whose device arrays hold NaN, destroyed: its memory goes back to
    # the allocator, where the next plan's allocations are likely to find it
    n = mib * 2**20 // 16
    fw = topefft.Framework()
    p = topefft.Plan(fw, (n,), np.full(n, np.nan + 1j * np.nan))
    p.exec_device(topefft.FORWARD)
    fw.synchronize()
    p.destroy()
    fw.destroy()


@pytest.mark.parametrize("prec", [topefft.PREC_F64, topefft.PREC_F32], ids=["fp64", "c64"])
def test_many_independent_of_stale_device_memory(oracle, prec):
    # every output element is written from this plan's data alone: after
    # NaN-filled allocations were freed, all transforms (not only sampled
    # ones) match numpy's FFT and the round trip restores the input
    poison_device_memory()
    sizes, x = make_batch(600, 0x57A1E)
    if prec == topefft.PREC_F32:
        x = x.astype(np.complex64)
    tol = 1e-12 if prec == topefft.PREC_F64 else 1e-6
    fw = topefft.Framework()
    q = topefft.ManyPlan(fw, sizes, x, prec)
    q.set_method(topefft.MANY_PERSISTENT)
    fwd = q.exec(topefft.FORWARD).astype(np.complex128)
    back = q.exec(topefft.INVERSE).astype(np.complex128)
    q.destroy()
    fw.destroy()
    assert np.all(np.isfinite(fwd)) and np.all(np.isfinite(back))
    x64 = x.astype(np.complex128)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    worst = max(rel_l2(fwd[o:o + n], np.fft.fft(x64[o:o + n])) for n, o in zip(sizes, offs))
    assert worst < tol
    assert rel_l2(back, x64) < tol
    check_against_oracle(oracle, sizes, x64, fwd, 1, per_class=1, tol=tol)


def test_many_radix8_schedules_from_wisdom(oracle, tmp_path, monkeypatch):
    # force radix-8 stages for every class that allows them through a wisdom
    # file, as the autotuner would write it
    sizes, x = make_batch(200, 0x8888)
    fw = topefft.Framework()
    path = tmp_path / "wisdom.txt"
    p = topefft.ManyPlan(fw, sizes, x)
    p.tune(str(path), reps=2)
    line = path.read_text().split()
    assert line[1] == "f64" and line[2].startswith("grid_per_cu=")
    dev = line[0]
    path.write_text(f"{dev} f64 grid_per_cu=1 sched=" + ",".join(f"{L}:8" for L in range(4, 15)) + "\n")
    p.destroy()
    monkeypatch.setenv("TOPEFFT_WISDOM", str(path))
    fw = topefft.Framework()
    q = topefft.ManyPlan(fw, sizes, x)
    radix, grid = q.schedule()
    assert all(radix[L] == 8 for L in range(4, 14)) and radix[14] == 16  # 2^14/8 threads do not fit
    fwd = q.exec(topefft.FORWARD)
    back = q.exec(topefft.INVERSE)
    q.destroy()
    fw.destroy()
    check_against_oracle(oracle, sizes, x, fwd, 1)
    assert rel_l2(back, x) < 1e-12


def test_many_tune_keeps_data_and_results(oracle, tmp_path):
    sizes, x = make_batch(300, 0x7077)
    fw = topefft.Framework()
    p = topefft.ManyPlan(fw, sizes, x)
    p.tune(str(tmp_path / "w.txt"), reps=2)
    radix, grid = p.schedule()
    assert grid > 0 and all(r in (8, 16) for r in radix)
    out = p.exec(topefft.FORWARD)  # tuning restored the uploaded input
    p.destroy()
    fw.destroy()
    check_against_oracle(oracle, sizes, x, out, 1, per_class=1)


def test_many_complex64_vs_fp64():
    sizes, x = make_batch(500, 0xF32F)
    fw = topefft.Framework()
    p64 = topefft.ManyPlan(fw, sizes, x, topefft.PREC_F64)
    ref = p64.exec(topefft.FORWARD)
    p64.destroy()
    fw = topefft.Framework()
    p32 = topefft.ManyPlan(fw, sizes, x.astype(np.complex64), topefft.PREC_F32)
    got = p32.exec(topefft.FORWARD).astype(np.complex128)
    p32.destroy()
    fw.destroy()
    err = rel_l2(got, ref)
    print(f"uneven batch complex64 vs FP64 rel-L2 {err:.3e}")
    assert err < 1e-6


def test_many_full_20000(oracle):
    # the benchmark batch: 20,000 transforms, 2^6 .. 2^14 log-uniform
    sizes = workloads.uneven_sizes(workloads.SEED_MANY_SIZES, 20000)
    x = workloads.complex_uniform(workloads.SEED_MANY_DATA, int(sizes.sum()))
    fw = topefft.Framework()
    p = topefft.ManyPlan(fw, sizes, x)
    p.set_method(topefft.MANY_PERSISTENT_GRAPH)
    out = p.exec(topefft.FORWARD)
    p.destroy()
    fw.destroy()
    check_against_oracle(oracle, sizes, x, out, 1, per_class=1)


def test_many_full_20000_complex64(oracle):
    # the same batch in complex64 (FP32 storage and arithmetic): two
    # transforms of every length class against the naive DFT of the
    # complex64 input, and the whole batch against the FP64 plan
    sizes = workloads.uneven_sizes(workloads.SEED_MANY_SIZES, 20000)
    x = workloads.complex_uniform(workloads.SEED_MANY_DATA, int(sizes.sum())).astype(np.complex64)
    x64 = x.astype(np.complex128)
    fw = topefft.Framework()
    p = topefft.ManyPlan(fw, sizes, x, topefft.PREC_F32)
    p.set_method(topefft.MANY_PERSISTENT_GRAPH)
    out = p.exec(topefft.FORWARD).astype(np.complex128)
    p.destroy()
    fw.destroy()
    fw = topefft.Framework()  # the last plan's destroy released the first one
    p64 = topefft.ManyPlan(fw, sizes, x64)
    ref = p64.exec(topefft.FORWARD)
    p64.destroy()
    fw.destroy()
    seen = check_against_oracle(oracle, sizes, x64, out, 1, per_class=2, tol=1e-6)
    assert sorted(seen) == sorted(set(int(n) for n in sizes))
    err = rel_l2(out, ref)
    print(f"20000-transform complex64 batch vs FP64 rel-L2 {err:.3e}")
    assert err < 1e-6


def test_many_rejects_bad_sizes():
    fw = topefft.Framework()
    with pytest.raises(topefft.TopeError) as e:
        topefft.ManyPlan(fw, [64, 100], np.zeros(164, complex))
    assert e.value.code == topefft.TOPE_ERROR_INVALID_SIZE
    with pytest.raises(topefft.TopeError):
        topefft.ManyPlan(fw, [64, 1 << 15], np.zeros(64 + (1 << 15), complex))
    fw.destroy()
