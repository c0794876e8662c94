"""The C drop-in harnesses (the reference's test/{1,2,3}d.c call sequences
and the real-input one), built with gcc against include/topefft.h and run:
each checks a closed form and/or a round trip itself and exits 0 on pass."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "include")
LIBDIR = os.path.join(REPO, "tope-fft_amd", "lib")

RUNS = [("harness1d", ["65536"]), ("harness1d", ["1000"]), ("harness1d", ["4097"]), ("harness1d", ["1048576"]),
        ("harness2d", ["256", "64"]), ("harness2d", ["4", "65536"]), ("harness3d", ["32", "16", "8"]),
        ("harness3d", ["4", "4", "8192"]), ("harness_real", ["256", "16", "4"]), ("harness_real", ["4096", "1", "1"])]


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = tmp_path_factory.mktemp("harness")
    for name in sorted({r[0] for r in RUNS}):
        subprocess.run(["gcc", "-std=gnu99", "-O2", "-I", INC, os.path.join(REPO, "tope-fft_amd", "examples",
                        f"{name}.c"), "-o", str(d / name), "-L", LIBDIR, "-ltopefft", "-lm",
                        f"-Wl,-rpath,{LIBDIR}"], check=True)
    return d


@pytest.mark.parametrize("name,args", RUNS, ids=lambda v: v if isinstance(v, str) else "x".join(v))
def test_harness(built, name, args):
    r = subprocess.run([str(built / name)] + args, capture_output=True, text=True, timeout=120)
    print(name, " ".join(args), "->", r.stdout.strip())
    assert r.returncode == 0, (r.stdout, r.stderr)


VENDOR = os.path.join(LIBDIR, "vendor_compare")


@pytest.mark.skipif(not os.path.exists(VENDOR), reason="vendor_compare not built (needs rocFFT)")
def test_vendor_compare_agrees_with_rocfft():
    # the same forward transform through libtopefft and rocFFT, device-resident
    import json
    r = subprocess.run([VENDOR, "3", "2d4096", "1d65536", "batch32"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    assert [w["workload"] for w in rows] == ["2d4096", "1d65536", "batch32"]
    for w in rows:
        print(w)
        assert w["rel_l2"] < (1e-5 if w["precision"] == "complex64" else 1e-12), w
        assert w["topefft_us"] > 0 and w["rocfft_us"] > 0
