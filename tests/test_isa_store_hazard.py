"""Static check of the built library's gfx950 code: no buffer store of more
than 64 bits of data may be followed directly by a VALU instruction that
overwrites its data registers.

Such a store with an SGPR soffset gets no wait state from the compiler, and
on gfx950 the following VALU write has replaced the upper dwords of the
stored value in some lanes (the uneven-batch kernel returned one wrong
imaginary part per thread in 16 lanes).  The kernels therefore pass 128-bit
stores an inline-constant soffset, for which the compiler inserts the wait
state; this test disassembles every code object in libtopefft.so (our own
build) and fails on any unguarded store -> overwrite pair."""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "tope-fft_amd", "lib", "libtopefft.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

STORE = re.compile(r"^\s*(buffer|global|flat|scratch)_store_dwordx([34])\s+(?:\w+,\s*)?v\[(\d+):(\d+)\]")
VDST = re.compile(r"^\s*v_\w+\s+v(?:\[(\d+):(\d+)\]|(\d+))")


def code_objects(tmp):
    fat = os.path.join(tmp, "fatbin.bin")
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, fat], check=True)
    data = open(fat, "rb").read()
    starts = []
    i = data.find(MAGIC)
    while i >= 0:
        starts.append(i)
        i = data.find(MAGIC, i + 1)
    out = []
    for k, st in enumerate(starts):
        en = starts[k + 1] if k + 1 < len(starts) else len(data)
        b = os.path.join(tmp, f"b{k}.bin")
        with open(b, "wb") as f:
            f.write(data[st:en])
        co = os.path.join(tmp, f"co{k}.elf")
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o", f"--input={b}",
                        f"--targets={TARGET}", f"--output={co}"], check=True)
        if os.path.getsize(co) > 0:
            out.append(co)
    return out


def unguarded(lines):
    """(store, next) pairs where the next instruction is a VALU write that
    overlaps the store's data registers."""
    bad = []
    insts = [ln.split("//")[0].rstrip() for ln in lines]
    insts = [ln for ln in insts if ln.strip() and not ln.strip().endswith(":") and not ln.startswith(("0", "D"))]
    for a, b in zip(insts, insts[1:]):
        m = STORE.match(a)
        if not m or m.group(1) != "buffer":
            continue  # FLAT-encoded stores are always guarded by the compiler
        lo, hi = int(m.group(3)), int(m.group(4))
        d = VDST.match(b)
        if not d:
            continue
        dlo = int(d.group(1) or d.group(3))
        dhi = int(d.group(2) or d.group(3))
        if not (dhi < lo or dlo > hi):
            bad.append((a.strip(), b.strip()))
    return bad


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtopefft.so not built")
@pytest.mark.skipif(not (shutil.which("objcopy") and os.path.exists(f"{LLVM}/llvm-objdump")),
                    reason="objcopy / ROCm llvm tools absent")
def test_no_unguarded_wide_buffer_store(tmp_path):
    cos = code_objects(str(tmp_path))
    assert cos, "no gfx950 code object in libtopefft.so"
    stores = 0
    bad = []
    for co in cos:
        dis = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", co], check=True,
                             capture_output=True, text=True).stdout.splitlines()
        stores += sum(1 for ln in dis if STORE.match(ln))
        bad += unguarded(dis)
    assert stores > 100  # the FP64 kernels' 128-bit stores were found
    assert not bad, bad[:5]


def test_checker_flags_the_pattern():
    # the pattern that corrupted the batch kernel, and its guarded form
    bad = ["\tbuffer_store_dwordx4 v[12:15], v74, s[56:59], s11 offen   // 0000: E07C",
           "\tv_ldexp_f64 v[14:15], v[104:105], -14   // 0008: D2880000"]
    ok = ["\tbuffer_store_dwordx4 v[12:15], v74, s[56:59], 0 offen",
          "\ts_nop 0",
          "\tv_ldexp_f64 v[14:15], v[104:105], -14"]
    other = ["\tbuffer_store_dwordx4 v[12:15], v74, s[56:59], 0 offen",
             "\tv_add_f64 v[20:21], v[2:3], v[4:5]"]
    assert len(unguarded(bad)) == 1
    assert unguarded(ok) == []
    assert unguarded(other) == []
