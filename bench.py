"""Benchmark of the topeFFT transform path on one MI355X.

Workload (one "step"): one 2D C2C FP64 transform of a 4096 x 4096
complex128 array (256 MiB) already resident in HBM, executed through the
C-ABI's device-resident entry point (tope2DExecDevice); the steps alternate
FORWARD and INVERSE (the same kernels; INVERSE also applies the fused 1/N).
Input: seeded uniform [-1,1) re/im plus 8 planted unit exponentials.

value = algorithmic GB/s = 2 * 16 B * N * P / t, P = 2 (one HBM read and
one write per axis), N = 4096^2 -> 1 GiB per transform.  Also printed:
transforms/s, FP64 FLOP/s (5 N log2 N) against the 78.6 TFLOP/s vector
peak, the same transform in complex64 storage (FP32 VALU and matrix-core
plans) with its speedup and error against the FP64 result, side workloads,
and for the dominant kernel its HIP-event duration against the 8 TB/s HBM
peak.

cpu_baseline: the oracle's CPU restatement of the reference algorithm
(radix-8 DIT + bit-reversal swap + one full pass per stage, OpenMP) on the
same 2D input, timed on the host over a bounded number of transforms.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tope-fft_amd", "python"))

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def gbs(nbytes, seconds):
    return nbytes / seconds / 1e9


def measure(fw, plan, steps, warmup):
    """Timed region: `steps` transforms back to back on the plan's stream,
    bracketed by stream synchronisation, alternating FORWARD and INVERSE so
    the resident data stays bounded (the two directions run the same
    kernels; INVERSE also applies the fused 1/N).  Then each kernel alone:
    HIP events on the same stream around `steps` launches of that pass
    (events between passes would add ~8 us idle gaps to every step)."""
    import topefft
    dirs = [topefft.FORWARD, topefft.INVERSE]
    for i in range(warmup):
        plan.exec_device(dirs[i % 2])
    fw.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.exec_device(dirs[i % 2])
    fw.synchronize()
    t1 = time.perf_counter()
    pass_ms = np.array([plan.time_pass(i, topefft.FORWARD, steps) for i in range(plan.pass_count())])
    return (t1 - t0) / steps, pass_ms


def kernel_symbol(variant_name, sign=-1):
    """axis_double_n4096_e16_tb4_col_split -> the kernel's template name
    prefix as rocprofv3 prints it (forward, no post-twiddle); the
    even/odd-split column tiles (..._dit2) are dit2_kernel, the three-step
    column tile (..._ct3) ct3_kernel."""
    parts = variant_name.split("_")
    prec, n, e, tb = parts[1], parts[2][1:], parts[3][1:], parts[4][2:]
    col = "true" if parts[5] == "col" else "false"
    split = "true" if "split" in parts[6:] else "false"
    if "ct3" in parts[6:]:
        return f"tope::ct3_kernel<{prec}, {sign}, false"
    if "dit2" in parts[6:]:
        return f"tope::dit2_kernel<{prec}, {n}, {e}, {tb}, {split}, {sign}, false"
    return f"tope::axis_kernel<{prec}, {n}, {e}, {tb}, {col}, {split}, {sign}, false"


TRAFFIC_FILE = os.path.join("profiles", "pmc_traffic_2d4096_fp64.md")


def profiled_traffic(variant_name):
    """HBM bytes per launch of this kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a plain copy), or None."""
    path = os.path.join(REPO, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    want = kernel_symbol(variant_name)
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch ") and want in line:
            return int(line.split()[-1])
    return None


def measure_batch(fw, x, n, count, prec, mode, steps, warmup):
    """One batch plan: first FORWARD downloaded (accuracy), then timed."""
    import topefft
    p = topefft.BatchPlan(fw, n, count, x, prec, mode)
    out = p.exec(topefft.FORWARD)
    t, pass_ms = measure(fw, p, steps, warmup)
    name = p.pass_names()[0]
    p.destroy()
    return out, t, pass_ms, name


def mixed_precision_leg(fw, steps, warmup):
    """4096 transforms of length 4096: complex128 FP64 against complex64 with
    radix-16 stages on MFMA (and on the VALU), Gaussian input."""
    import topefft
    import workloads
    n = count = 4096
    x = workloads.complex_gaussian(workloads.SEED_BATCH, n * count)
    ref, t64, p64, name64 = measure_batch(fw, x, n, count, topefft.PREC_F64, topefft.MODE_VALU, steps, warmup)
    refn = np.linalg.norm(ref)
    res = {"workload": "batch 4096 x 1D C2C n=4096, Gaussian(0,1), forward and inverse alternating",
           "fp64": {"kernel": name64, "ms": round(t64 * 1e3, 4), "transforms_per_s": round(count / t64, 1),
                    "gbs": round(gbs(2 * 16 * n * count, t64), 1)},
           "complex64": {}}
    x32 = x.astype(np.complex64)
    for mode in (topefft.MODE_MFMA_FP16X3, topefft.MODE_MFMA_BF16X6, topefft.MODE_MFMA_FP32,
                 topefft.MODE_MFMA_BF16X3, topefft.MODE_VALU):
        out, t, pm, name = measure_batch(fw, x32, n, count, topefft.PREC_F32, mode, steps, warmup)
        d = out.astype(np.complex128) - ref
        res["complex64"][topefft.MODE_NAMES[mode]] = {
            "kernel": name, "ms": round(t * 1e3, 4), "transforms_per_s": round(count / t, 1),
            "gbs": round(gbs(2 * 8 * n * count, t), 1), "speedup_vs_fp64": round(t64 / t, 3),
            "rel_l2_vs_fp64": float(f"{np.linalg.norm(d) / refn:.3e}"),
            "max_abs_vs_fp64": float(f"{np.max(np.abs(d)):.3e}")}
    return res


def mp_2d_leg(fw, x, t64, steps, warmup):
    """The headline transform in complex64 storage: the FP32 VALU plan and
    the matrix-core plans (split-fp16 and bf16x6 operands; axes without an
    MFMA kernel run the FP32 VALU pass, named in `passes`), each timed like
    the FP64 headline and compared with the FP64 forward spectrum of the
    same input (rel-L2, max-abs)."""
    import topefft
    dims = (4096, 4096)
    n = dims[0] * dims[1]
    p = topefft.Plan(fw, dims, x)
    ref = p.exec(topefft.FORWARD)
    p.destroy()
    refn = np.linalg.norm(ref)
    x32 = x.astype(np.complex64)
    res = {"workload": "2D C2C 4096x4096 complex64 storage, forward and inverse alternating, device-resident",
           "fp64_ms": round(t64 * 1e3, 4)}
    for mode in (None, topefft.MODE_MFMA_FP16X3, topefft.MODE_MFMA_BF16X6):
        q = topefft.Plan(fw, dims, x32, precision=topefft.PREC_F32, mode=mode)
        out = q.exec(topefft.FORWARD).astype(np.complex128)
        q.exec(topefft.INVERSE)
        t, pm = measure(fw, q, steps, warmup)
        names = q.pass_names()
        q.destroy()
        d = out - ref
        res["valu" if mode is None else topefft.MODE_NAMES[mode]] = {
            "ms": round(t * 1e3, 4), "transforms_per_s": round(1.0 / t, 2),
            "gbs": round(gbs(2 * 8 * n * 2, t), 1), "speedup_vs_fp64": round(t64 / t, 3),
            "rel_l2_vs_fp64": float(f"{np.linalg.norm(d) / refn:.3e}"),
            "max_abs_vs_fp64": float(f"{np.max(np.abs(d)):.3e}"),
            "passes": [{"kernel": nm, "ms": round(float(m), 4)} for nm, m in zip(names, pm)]}
    return res


def uneven_batch_leg(fw, steps, warmup, graphs=True):
    """20,000 transforms, lengths 2^6..2^14 log-uniform, packed with an
    offset table: one persistent launch (direct and hipGraph replay, after
    autotuning) against one launch per transform (direct and graph)."""
    import topefft
    import workloads
    sizes = workloads.uneven_sizes(workloads.SEED_MANY_SIZES, 20000)
    total = int(sizes.sum())
    x = workloads.complex_uniform(workloads.SEED_MANY_DATA, total)
    res = {"workload": "20000 x 1D C2C, n = 2^6..2^14 log-uniform (seeded), packed, forward and inverse alternating",
           "elements": total}
    for prec, name, xs, sb in ((topefft.PREC_F64, "fp64", x, 16),
                               (topefft.PREC_F32, "complex64", x.astype(np.complex64), 8)):
        p = topefft.ManyPlan(fw, sizes, xs, prec)
        wisdom = os.path.join(tempfile.mkdtemp(prefix="topefft_"), "wisdom.txt")
        p.tune(wisdom, reps=3)
        radix, grid = p.schedule()
        entry = {"schedule_radix_by_log2n": {str(L): radix[L] for L in range(6, 15)}, "grid": grid}
        methods = (topefft.MANY_PERSISTENT, topefft.MANY_PERSISTENT_GRAPH, topefft.MANY_PER_TRANSFORM,
                   topefft.MANY_PER_TRANSFORM_GRAPH)
        for m in methods if graphs else methods[::2]:
            p.set_method(m)
            k = steps if m in (topefft.MANY_PERSISTENT, topefft.MANY_PERSISTENT_GRAPH) else 3
            dirs = [topefft.FORWARD, topefft.INVERSE]
            for i in range(max(1, warmup)):
                p.exec_device(dirs[i % 2])
            fw.synchronize()
            t0 = time.perf_counter()
            for i in range(k):
                p.exec_device(dirs[i % 2])
            fw.synchronize()
            t = (time.perf_counter() - t0) / k
            # device time per exec from HIP events on the plan's stream (the
            # graph methods: per replay of the captured graph)
            dev_ms = p.time_exec(k)
            entry[topefft.MANY_METHOD_NAMES[m]] = {"ms": round(t * 1e3, 4),
                                                    "gbs": round(gbs(2 * sb * total, t), 1),
                                                    "transforms_per_s": round(len(sizes) / t, 1),
                                                    "device_ms_hip_events": round(dev_ms, 4)}
        entry["per_class_alone"] = per_class_times(fw, sizes, xs, prec, sb, wisdom, max(5, steps))
        p.destroy()
        res[name] = entry
    return res


def per_class_times(fw, sizes, xs, prec, sb, wisdom, reps):
    """Each length class of the uneven batch alone (its transforms packed in
    their own plan, the tuned schedule from `wisdom`, persistent method):
    device ms per exec from HIP events and the HBM rate.  The 1024-thread
    classes (n >= 8192) own a CU per block; the others run four blocks per CU."""
    import topefft
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    out = {}
    old = os.environ.get("TOPEFFT_WISDOM")
    os.environ["TOPEFFT_WISDOM"] = wisdom
    try:
        for L in range(6, 15):
            sel = np.flatnonzero(sizes == (1 << L))
            if not len(sel):
                continue
            xc = np.concatenate([xs[offs[i]:offs[i] + sizes[i]] for i in sel])
            pc = topefft.ManyPlan(fw, sizes[sel], xc, prec)
            pc.set_method(topefft.MANY_PERSISTENT)
            pc.exec_device(topefft.FORWARD)
            ms = pc.time_exec(reps)
            pc.destroy()
            out[str(L)] = {"transforms": int(len(sel)), "device_ms_hip_events": round(ms, 4),
                           "gbs": round(gbs(2 * sb * float(sizes[sel].sum()), ms * 1e-3), 1)}
    finally:
        if old is None:
            os.environ.pop("TOPEFFT_WISDOM", None)
        else:
            os.environ["TOPEFFT_WISDOM"] = old
    return out


def large_3d_leg(fw, steps, warmup):
    """3D 512^3: Gaussian random field with |k|^-2 power spectrum (k-space
    synthesis, inverse transform by the FP64 plan) plus 1e-3 white noise;
    FP64 (2 GiB) and complex64 (1 GiB) plans -- FP32 VALU axis passes, and
    split-fp16 MFMA passes where a kernel covers the axis -- forward then
    inverse."""
    import topefft
    import workloads
    dims = (512, 512, 512)
    n = 512 ** 3
    spec = workloads.powerlaw_spectrum(workloads.SEED_3D_LARGE, dims)
    p = topefft.Plan(fw, dims, spec)
    field = p.exec(topefft.INVERSE)
    p.destroy()
    del spec
    field += 1e-3 * workloads.complex_gaussian(workloads.SEED_3D_LARGE + 1, n)
    res = {"workload": "3D C2C 512^3, |k|^-2 Gaussian random field + 1e-3 white noise, forward then inverse"}
    ref = None
    for prec, mode, name, sb in ((None, None, "fp64", 16), (topefft.PREC_F32, None, "complex64", 8),
                                 (topefft.PREC_F32, topefft.MODE_MFMA_FP16X3, "complex64_mfma_fp16x3", 8)):
        xs = field if prec is None else field.astype(np.complex64)
        q = topefft.Plan(fw, dims, xs, precision=prec, mode=mode)
        X = q.exec(topefft.FORWARD)
        back = q.exec(topefft.INVERSE)
        t, pass_ms = measure(fw, q, steps, warmup)
        names = q.pass_names()
        q.destroy()
        Xd = X.astype(np.complex128)
        entry = {"ms_per_transform": round(t * 1e3, 4), "gbs": round(gbs(2 * sb * n * 3, t), 1),
                 "passes": [{"kernel": nm, "ms": round(float(m), 4), "gbs": round(gbs(2 * sb * n, m * 1e-3), 1)}
                            for nm, m in zip(names, pass_ms)],
                 "round_trip_rel_l2": float(f"{np.linalg.norm(back.astype(np.complex128) - field) / np.linalg.norm(field):.3e}")}
        if ref is None:
            ref = Xd
        else:
            entry["rel_l2_vs_fp64"] = float(f"{np.linalg.norm(Xd - ref) / np.linalg.norm(ref):.3e}")
            entry["max_abs_vs_fp64"] = float(f"{np.max(np.abs(Xd - ref)):.3e}")
        del X, back, Xd, xs
        res[name] = entry
    return res


def real_2d_leg(fw, steps, warmup):
    """R2C/C2R 2D 4096^2 FP64 (seeded uniform real input): forward and
    inverse alternate as in the complex legs; algorithmic bytes per
    transform = 2 axes x (read + write of the (nx/2+1) x ny half spectrum)."""
    import topefft
    import workloads
    nx = ny = 4096
    x = workloads.uniform(workloads.SEED_2D, nx * ny)
    p = topefft.RealPlan(fw, (nx, ny), x)
    t, pass_ms = measure(fw, p, steps, warmup)
    names = p.pass_names()
    p.destroy()
    alg = 2 * 2 * 16 * (nx // 2 + 1) * ny
    out = {"workload": "2D R2C/C2R FP64 4096x4096 real, device-resident", "ms": round(t * 1e3, 4),
           "transforms_per_s": round(1.0 / t, 2), "gbs_algorithmic": round(gbs(alg, t), 1),
           "passes": [{"kernel": nm, "ms": round(float(m), 4)} for nm, m in zip(names, pass_ms)]}
    # complex64 storage on the same input: time and error against the FP64
    # half spectrum of one forward transform
    p64 = topefft.RealPlan(fw, (nx, ny), x)
    X64 = p64.forward()
    p64.destroy()
    q = topefft.RealPlan(fw, (nx, ny), x.astype(np.float32), precision=topefft.PREC_F32)
    X32 = q.forward().astype(np.complex128)
    t32, pass32 = measure(fw, q, steps, warmup)
    q.destroy()
    err = float(np.linalg.norm(X32 - X64) / np.linalg.norm(X64))
    out["complex64"] = {"ms": round(t32 * 1e3, 4), "transforms_per_s": round(1.0 / t32, 2),
                        "gbs_algorithmic": round(gbs(alg // 2, t32), 1), "speedup_vs_fp64": round(t / t32, 3),
                        "rel_l2_vs_fp64": float("%.3e" % err)}
    return out


def small_1d_leg(fw, steps, warmup):
    """1D 2^16 FP64 (latency-bound, two launches of the four-step schedule):
    device time per transform and transforms/s."""
    import topefft
    import workloads
    n = 1 << 16
    x = workloads.complex_uniform(workloads.SEED_1D, n)
    p = topefft.Plan(fw, (n,), x)
    t, pass_ms = measure(fw, p, max(steps, 50), warmup)
    names = p.pass_names()
    p.destroy()
    return {"workload": "1D C2C FP64 n=65536, device-resident", "us_per_transform": round(t * 1e6, 2),
            "transforms_per_s": round(1.0 / t, 1),
            "passes_us": [{"kernel": nm, "us": round(float(m) * 1e3, 2)} for nm, m in zip(names, pass_ms)]}


def long_1d_leg(fw, steps, warmup):
    """1D FP64 of 2^24 points (the four-step: two 4096-point passes) and
    2^25 points (three passes of 512, 256, 256): device time per transform,
    the rate of one HBM round trip of the array per pass, pass times."""
    import topefft
    import workloads
    out = {}
    for n in (1 << 24, 1 << 25):
        x = workloads.complex_uniform(workloads.SEED_1D + n % 1000, n)
        p = topefft.Plan(fw, (n,), x)
        t, pass_ms = measure(fw, p, steps, warmup)
        names = p.pass_names()
        p.destroy()
        del x
        out[str(n)] = {"us_per_transform": round(t * 1e6, 1), "passes": len(names),
                       "gbs_per_pass": round(gbs(2 * 16 * n * len(names), t), 1),
                       "passes_us": [{"kernel": nm, "us": round(float(m) * 1e3, 1)} for nm, m in zip(names, pass_ms)]}
    return out


def vendor_leg(steps):
    """rocFFT (ROCm's FFT library, run-only comparator) on the same
    device-resident transforms, timed by tope-fft_amd/lib/vendor_compare
    with HIP events on one stream beside libtopefft; also the rel-L2 between
    the two libraries' outputs."""
    import json
    import subprocess
    exe = os.path.join(REPO, "tope-fft_amd", "lib", "vendor_compare")
    if not os.path.exists(exe):
        return {"skipped": "vendor_compare not built"}
    r = subprocess.run([exe, str(steps), "2d4096", "3d256", "1d65536", "batch64", "batch32"], capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        return {"error": (r.stderr or r.stdout)[-300:]}
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def pattern_yardsticks(reps=20):
    """The access patterns of the two 2D 4096^2 passes as plain in-place
    copies of the same 256 MiB array (tope-fft_amd/lib/copy_calib: rows of
    4096 elements per 256-thread block; 4-column tiles per 1024-thread block
    laid out per XCD), timed with HIP events in this run: the floors the X
    and Y passes are set against."""
    import subprocess
    exe = os.path.join(REPO, "tope-fft_amd", "lib", "copy_calib")
    if not os.path.exists(exe):
        return {"skipped": "copy_calib not built"}
    out = {}
    for mode in ("rows", "cols"):
        r = subprocess.run([exe, "256", str(reps), mode], capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            return {"error": (r.stderr or r.stdout)[-300:]}
        us = float(r.stdout.split(":")[1].split()[0])
        out[mode] = {"us": us, "gbs": round(2 * 256 * 2**20 / (us * 1e-6) / 1e9, 1)}
    return out


def cpu_library_fft(x, dims, cores):
    """A tuned CPU FFT beside the restated reference: scipy.fft (pocketfft)
    fft2 on the same input with `cores` worker threads, median of 5."""
    try:
        import scipy.fft
    except ImportError:
        return {"skipped": "scipy absent"}
    a = np.asarray(x, dtype=np.complex128).reshape(dims[1], dims[0])
    scipy.fft.fft2(a, workers=cores)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        scipy.fft.fft2(a, workers=cores)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"library": f"scipy {scipy.__version__} fft2 (pocketfft)", "workers": cores,
            "ms_per_transform": round(t * 1e3, 2), "gbs": round(gbs(2 * 16 * dims[0] * dims[1] * 2, t), 2)}


def cpu_baseline(x, dims, budget_s=12.0):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py
    oracle_py.load()
    y = np.array(x, copy=True)
    oracle_py.refcpu(y, dims, oracle_py.FORWARD, inplace=True)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_start) < budget_s:
        y[:] = x
        t0 = time.perf_counter()
        oracle_py.refcpu(y, dims, oracle_py.FORWARD, inplace=True)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times), oracle_py.threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="uneven batch: skip the hipGraph-replay methods (for kernel-trace profiling)")
    ap.add_argument("--legs", default="all",
                    help="comma list of mp2d,1d,long,3d256,mp,many,3d512,r2c,vendor,yard (default all; 2d is always run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(one process, one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import topefft
    import workloads

    fw = topefft.Framework()

    # ---- 2D 4096^2 FP64 (the headline workload) --------------------------
    dims2 = (4096, 4096)
    n2 = dims2[0] * dims2[1]
    x2, _ = workloads.planted_input(workloads.SEED_2D, (4096, 4096, 1))
    p2 = topefft.Plan(fw, dims2, x2)
    t2, pass2 = measure(fw, p2, args.steps, args.warmup)
    names2 = p2.pass_names()
    alg2 = 2 * 16 * n2 * 2
    value = gbs(alg2, t2)
    dom = int(np.argmax(pass2))
    dom_bytes = 2 * 16 * n2  # one read + one write of the array per pass
    achieved = gbs(dom_bytes, pass2[dom] * 1e-3)

    # the side legs (reported beside the headline, not part of `value`) time
    # at most 20 steps each so the default run stays short
    side_steps = max(5, min(20, args.steps // 2))
    legs = set(args.legs.split(",")) if args.legs != "all" else {"mp2d", "1d", "long", "3d256", "mp", "many", "3d512",
                                                                  "r2c", "vendor", "yard"}
    extra = {}
    if "mp2d" in legs:
        extra["mixed_precision_2d_4096"] = mp_2d_leg(fw, x2, t2, args.steps, args.warmup)
    if "1d" in legs:
        extra["1d_65536"] = small_1d_leg(fw, args.steps, args.warmup)
    if "long" in legs:
        extra["1d_long"] = long_1d_leg(fw, min(side_steps, 10), min(args.warmup, 3))
    # ---- 3D 256^3 FP64 (same recipe, beside it) ---------------------------
    if "3d256" in legs:
        dims3 = (256, 256, 256)
        n3 = 256 ** 3
        x3, _ = workloads.planted_input(workloads.SEED_3D, dims3)
        p3 = topefft.Plan(fw, dims3, x3)
        t3, pass3 = measure(fw, p3, side_steps, args.warmup)
        alg3 = 2 * 16 * n3 * 3
        p3.destroy()
        del x3
        extra["3d_256cubed"] = {"gbs": round(gbs(alg3, t3), 2), "ms": round(t3 * 1e3, 4),
                                "transforms_per_s": round(1.0 / t3, 2),
                                "pass_ms": [round(float(m), 4) for m in pass3]}

    if "mp" in legs:
        extra["mixed_precision_batch"] = mixed_precision_leg(fw, side_steps, args.warmup)
    if "many" in legs:
        extra["uneven_batch"] = uneven_batch_leg(fw, min(side_steps, 10), min(args.warmup, 3), not args.no_graph)
    if "r2c" in legs:
        extra["r2c_2d_4096"] = real_2d_leg(fw, side_steps, args.warmup)
    if "3d512" in legs:
        extra["3d_512cubed"] = large_3d_leg(fw, 4, 1)
    if "vendor" in legs:
        extra["vendor_rocfft"] = vendor_leg(side_steps)
    yard = pattern_yardsticks() if "yard" in legs else None
    if yard is not None:
        extra["pattern_yardsticks"] = yard

    cpu = None
    if not args.no_cpu:
        ct, reps, cores = cpu_baseline(x2, dims2)
        cpu = {"value": round(gbs(alg2, ct), 3), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"median of {reps} forward 2D 4096x4096 C2C FP64 transforms by the oracle's "
                         f"CPU restatement of the reference algorithm (radix-8 DIT, bit-reversal swap, "
                         f"one pass per stage), same input, {cores} OpenMP threads, "
                         f"{ct * 1e3:.1f} ms per transform"}
        extra["cpu_library_fft"] = cpu_library_fft(x2, dims2, cores)

    flops2 = 5.0 * n2 * np.log2(n2)
    # pass i of the 2D plan is the row (0) or column (1) pattern
    pattern_us = None
    if isinstance(yard, dict) and "rows" in yard and len(names2) == 2:
        pattern_us = [yard["rows"]["us"], yard["cols"]["us"]]
    line = {
        "metric": "GB/s & transforms/s, FP64 vs mixed-prec MFMA at rel-L2<=1e-6, 2D 4096^2 C2C",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t2 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded uniform [-1,1) re/im + 8 planted unit exponentials",
        "config": {"workload": "2D C2C FP64 4096x4096 complex128, forward and inverse alternating, device-resident",
                   "transforms_per_s": round(1.0 / t2, 2),
                   "fp64_tflops_5nlogn": round(flops2 / t2 / 1e12, 3),
                   "fp64_flops_frac": round(flops2 / t2 / 1e12 / FP64_VECTOR_PEAK_TFLOPS, 4),
                   "algorithmic_bytes_per_transform": alg2,
                   "passes": {n: round(float(m), 4) for n, m in zip(names2, pass2)},
                   **extra},
        "roofline": {"bound": "hbm", "kernel": names2[dom], "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(names2[dom]),
                     "traffic_source": f"{TRAFFIC_FILE}: rocprofv3 FETCH_SIZE and WRITE_SIZE of this kernel, "
                                       "each scaled by what it reads for a plain copy (a separate profiled run)",
                     "per_launch_algorithmic_bytes": dom_bytes,
                     "avg_launch_ms": round(float(pass2[dom]), 5),
                     # the same kernel against an in-place copy in its own
                     # access pattern (column tiles), measured in this run
                     "frac_of_pattern_copy": (round(pattern_us[dom] / (float(pass2[dom]) * 1e3), 4)
                                              if pattern_us else None)},
        "cpu_baseline": cpu,
    }
    p2.destroy()
    fw.destroy()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
