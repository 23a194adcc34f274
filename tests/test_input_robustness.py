"""The binary input reader (swe_code/io/input.c:189-414) and the host's
init pipeline (RCM, padding, layouts) on corrupted files, under
AddressSanitizer and UBSan (a host-only build): an error code or a file that
still reads, never an invalid access, an overflow or a crash."""
import os
import random
import shutil
import struct
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CASE = os.path.join(GOLDEN, "icos642_case.bin")

HARNESS = r"""
#include "swe_host.h"
#include <stdio.h>
#include <stdlib.h>
int main(int argc, char **argv)
{
    swe_gsmd g;
    int err = swe_get_gsmd_bin(argv[1], &g);
    if (!err)
        swe_gsmd_free(&g);
    double *H = NULL;
    int e2 = swe_get_state_bin_ex(argv[1], 1, 0, &H);
    free(H);
    swe_layout l = {1, 0, 4};
    swe_problem p;
    int e3 = swe_problem_load(argv[1], 1, &l, &p);
    if (!e3)
        swe_problem_free(&p);
    printf("%d %d %d\n", err, e2, e3);
    return 0;
}
"""


def test_malformed_binary_inputs_never_crash(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    host = os.path.join(ROOT, "rbf-fd-swe_amd", "host")
    src = tmp_path / "harness.c"
    src.write_text(HARNESS)
    exe = tmp_path / "bin_asan"
    srcs = [str(src)] + [os.path.join(host, f) for f in ("swe_io.c", "swe_ncio.c", "swe_rcm.c",
                                                         "swe_layout.c", "swe_profiling.c")]
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe)] + srcs + ["-lm"], check=True)
    base = open(CASE, "rb").read()
    n_nodes, nnbr = struct.unpack("<ii", base[:8])
    idx_off = 8 + 32 + 8 * (17 * n_nodes + 4 * n_nodes * nnbr)
    rng = random.Random(642)
    cases = [base]
    for k in range(120):
        b = bytearray(base)
        kind = k % 4
        if kind == 0:      # header counts
            off = rng.choice([0, 4])
            b[off:off + 4] = struct.pack("<i", rng.choice(
                [0, -1, 1, 2**31 - 1, -2**31, n_nodes + 1, nnbr + 1, 65]))
        elif kind == 1:    # stencil ids out of range
            off = idx_off + 4 * rng.randrange(n_nodes * nnbr)
            b[off:off + 4] = struct.pack("<i", rng.choice([0, -5, n_nodes + 1, 2**31 - 1]))
        elif kind == 2:    # truncated
            b = b[:rng.randrange(len(base))]
        else:              # a disconnected stencil graph: a node whose stencil is itself
            node = rng.randrange(n_nodes)
            off = idx_off + 4 * node * nnbr
            b[off:off + 4 * nnbr] = struct.pack(f"<{nnbr}i", *([node + 1] * nnbr))
        cases.append(bytes(b))
    env = dict(os.environ, ASAN_OPTIONS="exitcode=86:detect_leaks=0",
               UBSAN_OPTIONS="exitcode=87:print_stacktrace=1")
    for k, data in enumerate(cases):
        f = tmp_path / "m.bin"
        f.write_bytes(data)
        r = subprocess.run([str(exe), str(f)], env=env, capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0, (k, r.returncode, r.stderr[-2000:])
        if k == 0:
            assert r.stdout.split() == ["0", "0", "0"]
