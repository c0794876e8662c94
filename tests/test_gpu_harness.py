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
