"""NetCDF input (the reference's USE_NCIO path, swe_code/io/input.c:14-186)
and the nc -> bin converter (swe_code/io/nc2bin.c:10-84).

The reference's own .nc inputs are empty files and libnetcdf is absent, so
parity is pinned through the binary path, which the reference's reader
fixes: a NetCDF file written by scipy (an independent implementation of the
classic formats) with the binary case's raw values must give the same
prepared problem, array for array, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from nc_cases import write_case_nc, write_cdf5

CASE = os.path.join(GOLDEN, "icos642_case.bin")
ARRAYS = ("x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm", "Dx", "Dy", "Dz", "L",
          "idx", "H_init", "H_100")


@pytest.fixture(scope="module")
def raw(pkg):
    return pkg.mesh.read_bin(CASE)


def same_problem(a, b):
    assert (a.n_nodes, a.nnbr, a.padded) == (b.n_nodes, b.nnbr, b.padded)
    assert (a.a, a.gh0, a.dt, a.gamma) == (b.a, b.gh0, b.dt, b.gamma)
    for nm in ARRAYS:
        assert np.array_equal(getattr(a, nm), getattr(b, nm)), nm
    if a.mapping is not None or b.mapping is not None:
        assert np.array_equal(a.mapping, b.mapping)


@pytest.mark.parametrize("version", [1, 2, 5])
@pytest.mark.parametrize("layout,rcm", [("cfdl", True), ("sfdl", False)])
def test_netcdf_equals_binary(pkg, raw, tmp_path, version, layout, rcm):
    nc = tmp_path / f"case_v{version}.nc"
    if version == 5:
        write_cdf5(raw, nc)
    else:
        write_case_nc(raw, nc, version=version)
    a = pkg.host.Problem(CASE, use_rcm=rcm, layout=layout, simd=4)
    b = pkg.host.Problem(str(nc), use_rcm=rcm, layout=layout, simd=4, netcdf=True)
    same_problem(a, b)


def test_netcdf_fp32_reading_equals_binary(pkg, raw, tmp_path):
    nc = tmp_path / "case.nc"
    write_case_nc(raw, nc, version=2)
    same_problem(pkg.host.Problem(CASE, fp32=True),
                 pkg.host.Problem(str(nc), fp32=True, netcdf=True))


def test_netcdf_types_are_converted(pkg, raw, tmp_path):
    """Variables stored as float, short and double-typed ids are read as
    nc_get_var_double / nc_get_var_int convert them."""
    nc = tmp_path / "typed.nc"
    types = {"x": np.float32, "ghm": np.float32, "DPx": np.float32, "idx": np.int16,
             "Nnodes": np.float64, "H": np.float32}
    write_case_nc(raw, nc, version=1, types=types)
    p = pkg.host.Problem(str(nc), use_rcm=False, layout="cfdl", netcdf=True)
    N, n = raw.n_nodes, raw.nnbr
    assert np.array_equal(p.x[:N], raw.X[:, 0].astype(np.float32).astype(np.float64))
    assert np.array_equal(p.ghm[:N], raw.ghm.astype(np.float32).astype(np.float64))
    Dx = p.Dx.reshape(p.padded, p.padded_nnbr)[:N, :n]
    assert np.array_equal(Dx, raw.Dx.astype(np.float32).astype(np.float64) / raw.a)
    idx = p.idx.reshape(p.padded, p.padded_nnbr)[:N, :n]
    assert np.array_equal(idx, raw.idx)
    H = p.to_node_major(p.H_init)[:N]
    assert np.array_equal(H, raw.H_init.astype(np.float32).astype(np.float64))


def test_netcdf_record_variables(pkg, raw, tmp_path):
    """H and H_100 as record variables of one record (the unlimited
    dimension first): the whole variable is the state."""
    nc = tmp_path / "rec.nc"
    write_case_nc(raw, nc, version=1, record_state=True)
    same_problem(pkg.host.Problem(CASE), pkg.host.Problem(str(nc), netcdf=True))


def test_netcdf_interleaved_records_read_as_scipy_reads_them(pkg, raw, tmp_path):
    """One record per node: the records of H and H_100 interleave in the
    file (record size 64 B).  The reader must return what scipy's reader
    returns for the same bytes (scipy's writer does not reproduce the input
    for this layout, so the file itself is the reference here)."""
    from scipy.io import netcdf_file
    nc = tmp_path / "rec_nodes.nc"
    write_case_nc(raw, nc, version=1, record_state="nodes")
    with netcdf_file(str(nc), "r", mmap=False) as f:
        assert f.variables["H"].shape == (raw.n_nodes, 4)
        want0, want1 = f.variables["H"][:].copy(), f.variables["H_100"][:].copy()
    p = pkg.host.Problem(str(nc), use_rcm=False, layout="cfdl", netcdf=True)
    N = raw.n_nodes
    assert np.array_equal(p.to_node_major(p.H_init)[:N], want0)
    assert np.array_equal(p.to_node_major(p.H_100)[:N], want1)


def test_nc2bin_round_trip(pkg, raw, tmp_path):
    """nc2bin.c: weights scaled by 1/a and gamma on reading, unscaled again on
    writing, ids back to 1-based; the converter program and the library call
    write the same file."""
    nc, b1, b2 = tmp_path / "c.nc", tmp_path / "lib.bin", tmp_path / "tool.bin"
    write_case_nc(raw, nc, version=2)
    pkg.host.nc2bin(str(nc), str(b1))
    env = dict(os.environ, NC2BIN_INPUT_FILE=str(nc), NC2BIN_OUTPUT_FILE=str(b2))
    r = subprocess.run([os.path.join(ROOT, "rbf-fd-swe_amd", "swe_nc2bin")], env=env,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Nnodes = {raw.n_nodes}" in r.stdout
    assert b1.read_bytes() == b2.read_bytes()
    back = pkg.mesh.read_bin(str(b1))
    for nm in ("X", "f", "ghm", "p_u", "p_v", "p_w", "gradghm", "idx", "H_init", "H_final"):
        assert np.array_equal(getattr(back, nm), getattr(raw, nm)), nm
    assert (back.a, back.gh0, back.dt, back.gamma) == (raw.a, raw.gh0, raw.dt, raw.gamma)
    for nm in ("Dx", "Dy", "Dz"):
        assert np.array_equal(getattr(back, nm), (getattr(raw, nm) / raw.a) * raw.a), nm
    assert np.array_equal(back.L, (raw.L * raw.gamma) / raw.gamma)
    # the converted file prepares the same problem as the NetCDF one
    same_problem(pkg.host.Problem(str(b1)), pkg.host.Problem(str(nc), netcdf=True))


def _code(fn):
    with pytest.raises(RuntimeError) as e:
        fn()
    return e.value.code


def test_netcdf_errors(pkg, raw, tmp_path):
    P = pkg.host.Problem
    miss = tmp_path / "miss.nc"
    write_case_nc(raw, miss, drop=("gh0",))
    assert _code(lambda: P(str(miss), netcdf=True)) == -7            # ENOVAR
    noidx = tmp_path / "noidx.nc"
    write_case_nc(raw, noidx, drop=("idx",))
    assert _code(lambda: P(str(noidx), netcdf=True)) == -7
    hdf = tmp_path / "nc4.nc"
    hdf.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    assert _code(lambda: P(str(hdf), netcdf=True)) == -6             # EFORMAT
    assert _code(lambda: P(CASE, netcdf=True)) == -6                  # a binary file
    good = tmp_path / "good.nc"
    write_case_nc(raw, good, version=1)
    cut = tmp_path / "cut.nc"
    cut.write_bytes(good.read_bytes()[:-4000])
    assert _code(lambda: P(str(cut), netcdf=True)) == -2             # EREAD
    assert _code(lambda: P(str(tmp_path / "none.nc"), netcdf=True)) == -1   # EOPEN
    big = tmp_path / "big.nc"
    write_case_nc(raw, big, types={"idx": np.float64})
    _set_var(big, "idx", lambda a: np.where(a == a.max(), 3.0e9, a))
    assert _code(lambda: P(str(big), netcdf=True)) == -8             # ERANGE
    txt = tmp_path / "txt.nc"
    write_case_nc(raw, txt, types={"x": "S1"})
    assert _code(lambda: P(str(txt), netcdf=True)) == -6             # text where numbers are read
    short = tmp_path / "short.nc"
    write_case_nc(raw, short)
    _set_var(short, "Nnodes", lambda a: a - 1)
    assert _code(lambda: P(str(short), netcdf=True)) == -2           # sizes disagree


def _set_var(path, name, fn):
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "a") as f:
        v = f.variables[name]
        v[...] = fn(v[...].copy())


def test_malformed_files_never_crash_the_reader(pkg, raw, tmp_path):
    """The reader under AddressSanitizer and UBSan (a host-only build of the
    converter program): headers with flipped bytes, extreme counts, sizes and
    offsets, and truncated files give an error code (or a file that still
    reads), never an invalid access, overflow or crash."""
    import random
    import shutil
    import struct
    if shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    host = os.path.join(ROOT, "rbf-fd-swe_amd", "host")
    exe = tmp_path / "nc2bin_asan"
    srcs = [os.path.join(host, f) for f in ("swe_nc2bin.c", "swe_ncio.c", "swe_io.c",
                                             "swe_rcm.c", "swe_layout.c", "swe_profiling.c")]
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe)] + srcs + ["-lm"], check=True)
    good = tmp_path / "good.nc"
    write_case_nc(raw, good, version=2)
    base = bytearray(good.read_bytes())
    hdr = 1200   # the header of this file (dimensions, variables, offsets)
    rng = random.Random(20261203)
    extremes = [0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0xFFFFFFFE, 0x40000000, 3]
    cases = []
    for n in range(160):
        b = bytearray(base)
        kind = n % 4
        if kind == 0:      # flipped bytes
            for _ in range(rng.randint(1, 6)):
                b[rng.randrange(hdr)] ^= 1 << rng.randrange(8)
        elif kind == 1:    # an extreme 32-bit field at a word boundary
            off = 4 * rng.randrange(hdr // 4)
            b[off:off + 4] = struct.pack(">I", rng.choice(extremes))
        elif kind == 2:    # an extreme 64-bit field (offsets of CDF-2)
            off = 4 * rng.randrange(hdr // 4 - 1)
            b[off:off + 8] = struct.pack(">Q", rng.choice(
                [0, 2**62, 2**63 - 1, 2**63, 2**64 - 1, len(base) - 8]))
        else:              # truncated
            b = b[:rng.randrange(len(base))]
        cases.append(bytes(b))
    env = dict(os.environ, ASAN_OPTIONS="exitcode=86:detect_leaks=0",
               UBSAN_OPTIONS="exitcode=87:print_stacktrace=1")
    for k, data in enumerate(cases):
        f = tmp_path / "m.nc"
        f.write_bytes(data)
        env.update(NC2BIN_INPUT_FILE=str(f), NC2BIN_OUTPUT_FILE=str(tmp_path / "m.bin"))
        r = subprocess.run([str(exe)], env=env, capture_output=True, text=True, timeout=60)
        assert r.returncode in (0, 1, 2), (k, r.returncode, r.stderr[-2000:])
