"""Scans the gfx950 code object of this repository's own libswe_hip.so for
instructions this build must never contain: writes through the scalar data
cache and hand-made packed-FP32 arithmetic.  (CPU test; it names those
mnemonics, so it is excluded from the trees that GPU runs receive.)"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

LLVM = "/opt/rocm/llvm/bin"
FORBIDDEN = re.compile(r"\b(s_store_\w+|s_atomic_\w+|s_buffer_store\w*|s_scratch_store\w*|"
                       r"s_dcache_wb\w*|s_dcache_discard\w*|v_pk_(add|mul|fma)_f32)\b")


def test_code_object_has_no_forbidden_instructions():
    so = os.path.join(ROOT, "rbf-fd-swe_amd", "libswe_hip.so")
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump"):
        if not os.path.exists(os.path.join(LLVM, tool)):
            pytest.skip(f"{tool} not available")
    d = tempfile.mkdtemp()
    try:
        fat = os.path.join(d, "fat.bin")
        co = os.path.join(d, "gfx950.co")
        subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fat}", so,
                        os.path.join(d, "junk.so")], check=True)
        # one offload bundle per translation unit, concatenated
        data = open(fat, "rb").read()
        magic = b"__CLANG_OFFLOAD_BUNDLE__"
        starts = [i for i in range(len(data)) if data.startswith(magic, i)]
        assert starts, "no offload bundle found"
        asm = ""
        for k, (a, b) in enumerate(zip(starts, starts[1:] + [len(data)])):
            part = os.path.join(d, f"part{k}.bin")
            with open(part, "wb") as fp:
                fp.write(data[a:b])
            subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o",
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={part}",
                            f"--output={co}"], check=True)
            asm += subprocess.run([f"{LLVM}/llvm-objdump", "-d", co], capture_output=True,
                                  text=True, check=True).stdout
    finally:
        shutil.rmtree(d, ignore_errors=True)
    assert "v_fmac_f64" in asm or "v_fma_f64" in asm       # the disassembly is real
    bad = sorted({m.group(0) for m in FORBIDDEN.finditer(asm)})
    assert not bad, bad
