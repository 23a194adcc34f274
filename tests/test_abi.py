"""The C-ABI libraries load and export every declared entry point; argument
errors are reported without touching a GPU; the product never links the
oracle."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swe_[a-z0-9_]+)\s*\(", text)))


def _exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_hip_library_exports_header(pkg):
    so = pkg.runtime.library_path()
    names = _declared("swe_hip.h")
    assert len(names) >= 16
    missing = [n for n in names if n not in _exported(so)]
    assert not missing
    assert set(pkg.runtime.EXPORTS) == set(names)
    C.CDLL(so)   # loads (HIP runtime present, no device needed)


def test_host_library_exports_header(pkg):
    so = os.path.join(ROOT, "rbf-fd-swe_amd", "libswe_host.so")
    names = _declared("swe_host.h")
    missing = [n for n in names if n not in _exported(so)]
    assert not missing


def test_invalid_arguments_rejected_before_device(pkg):
    L = pkg.runtime.lib()
    out = C.c_void_p()
    assert L.swe_hip_create(None, 1, 0, 0, C.byref(out)) == -1
    assert out.value is None
    assert L.swe_hip_set_state(None, 0, None) == -1
    assert L.swe_hip_rk_substep(None, 0) == -1
    assert L.swe_hip_rk_steps(None, 1) == -1
    assert L.swe_hip_strerror(-4) == b"stencil index outside the patch"
    assert b"NULL" in L.swe_hip_last_error() or L.swe_hip_last_error()


def test_bad_static_data_rejected(pkg):
    """Multi-patch layouts and members out of range fail with EINVAL."""
    path = pkg.mesh.case_path(642, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    s, keep = pkg.runtime.static_from_problem(prob)
    L = pkg.runtime.lib()
    out = C.c_void_p()
    s.compute_pid_s = 4
    assert L.swe_hip_create(C.byref(s), 1, 0, 0, C.byref(out)) == -1
    s.compute_pid_s = 0
    assert L.swe_hip_create(C.byref(s), 0, 0, 0, C.byref(out)) == -1
    s.Nnbr = 65
    assert L.swe_hip_create(C.byref(s), 1, 0, 0, C.byref(out)) == -1


def test_bad_options_rejected_before_device(pkg):
    """Out-of-range options fail with EINVAL (no device needed to find out)."""
    path = pkg.mesh.case_path(642, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    s, keep = pkg.runtime.static_from_problem(prob)
    L = pkg.runtime.lib()
    base = dict(members=1, flags=0, device=0, member_block=0, share=None, kernel=0,
                prefetch=0, node_order=0, tiled=0, slot_block=0)
    for field, bad in [("kernel", 8), ("kernel", -1), ("prefetch", 2), ("node_order", 3),
                       ("slot_block", 8), ("members", 0), ("member_block", 3),
                       ("tiled", 2), ("tiled", -1), ("tile_nodes", 100), ("tile_nodes", -2),
                       ("tile_rows", 8), ("tile_rows", 71), ("kernel", 1), ("kernel", 2)]:
        o = pkg.runtime.SweHipOptions(**dict(base, **{field: bad}))
        if field == "member_block":
            o.members = 8
        if field == "tiled" and bad == 2:   # and a tiled block other than 32
            o.tiled, o.members, o.member_block = 1, 64, 8
        out = C.c_void_p()
        assert L.swe_hip_create_ex(C.byref(s), C.byref(o), C.byref(out)) == -1, field
        assert out.value is None


def test_fp32_kernel_options_rejected_before_device(pkg):
    """Exact FP32 contexts sum every stencil in order (one lane per node, or
    LDS tiles); OCT is FP32-only; QUADW is an FP64 mapping; tile sizes are
    64, 128 or 256."""
    path = pkg.mesh.case_path(642, seed=1)
    p32 = pkg.host.Problem(path, use_rcm=True, layout="cfdl", fp32=True)
    p64 = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    bad = [(p32, dict(fp32=True, exact=True, kernel="pair")),
           (p32, dict(fp32=True, exact=True, kernel="quad")),
           (p32, dict(fp32=True, kernel="quadw")), (p32, dict(fp32=True, kernel="tile",
                                                            tile_nodes=100)),
           (p64, dict(kernel="oct")),
           # the resident kernel is single-state
           (p64, dict(kernel="resident", members=2))]
    for prob, kw in bad:
        with pytest.raises(pkg.runtime.SweHipError) as ei:
            pkg.runtime.HipContext(prob, **kw)
        assert ei.value.code == -1, kw


def test_no_device_is_a_loud_error(pkg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    path = pkg.mesh.case_path(642, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    with pytest.raises(pkg.runtime.SweHipError) as ei:
        pkg.runtime.HipContext(prob)
    assert ei.value.code == -3


def test_product_does_not_use_the_oracle(pkg):
    for so in ("libswe_hip.so", "libswe_host.so"):
        deps = subprocess.run(["ldd", os.path.join(ROOT, "rbf-fd-swe_amd", so)],
                              capture_output=True, text=True).stdout
        assert "oracle" not in deps and "swe_ref" not in deps
    pkg_dir = os.path.join(ROOT, "rbf-fd-swe_amd")
    for dirpath, _, files in os.walk(pkg_dir):
        for fn in files:
            if fn.endswith((".py", ".c", ".h", ".hip", ".cpp")) or fn == "Makefile":
                text = open(os.path.join(dirpath, fn)).read()
                assert "pyoracle" not in text and "liboracle" not in text, fn
                assert "_ref/" not in text, fn


def test_driver_errors(pkg, tmp_path):
    exe = os.path.join(ROOT, "rbf-fd-swe_amd", "swe_run")
    env = {k: v for k, v in os.environ.items() if not k.startswith("SWE_")}
    r = subprocess.run([exe], env=env, capture_output=True, text=True)
    assert r.returncode == 1 and "SWE_INPUT_FILE" in r.stderr
    env.update(SWE_INPUT_FILE=str(tmp_path / "nope.bin"), SWE_USE_NETCDF="0",
               SWE_NUM_COMPUTE_UNITS="1", SWE_NSTEPS="1", SWE_NATTEMPTS="1", SWE_USE_RCM="1",
               SWE_USE_OCL="0")
    r = subprocess.run([exe], env=env, capture_output=True, text=True)
    assert r.returncode == 2 and "Error opening file" in r.stdout
    env["SWE_USE_NETCDF"] = "1"   # the NetCDF reader: unopenable, then not NetCDF
    r = subprocess.run([exe], env=env, capture_output=True, text=True)
    assert r.returncode == 2 and "Error opening file" in r.stdout
    env["SWE_INPUT_FILE"] = os.path.join(ROOT, "tests", "golden", "icos642_case.bin")
    r = subprocess.run([exe], env=env, capture_output=True, text=True)
    assert r.returncode == 2 and "Error in input.c" in r.stdout


def test_serve_library_exports_header(pkg):
    so = pkg.serve.library_path()
    names = _declared("swe_serve.h")
    assert set(names) == set(pkg.serve.EXPORTS)
    missing = [n for n in names if n not in _exported(so)]
    assert not missing
    L = pkg.serve.lib()
    out = C.c_void_p()
    assert L.swe_srv_create(None, C.byref(out)) == -1
    cfg = pkg.serve.SrvConfig()
    L.swe_srv_default_config(C.byref(cfg))
    assert cfg.hbm_cap_bytes == 200 << 30 and cfg.slots_per_batch == 64 and cfg.tiled == 0
    cfg.slots_per_batch = 3
    assert L.swe_srv_create(C.byref(cfg), C.byref(out)) == -1
    # tiled batches need member block 32 and a multiple of 32 slots
    L.swe_srv_default_config(C.byref(cfg))
    cfg.tiled = 1
    assert L.swe_srv_create(C.byref(cfg), C.byref(out)) == -1   # member_block 16
    cfg.member_block = 32
    cfg.slots_per_batch = 48
    assert L.swe_srv_create(C.byref(cfg), C.byref(out)) == -1
    cfg.tiled = 2
    cfg.slots_per_batch = 64
    assert L.swe_srv_create(C.byref(cfg), C.byref(out)) == -1
    cfg.tiled = 1
    assert L.swe_srv_create(C.byref(cfg), C.byref(out)) == 0
    L.swe_srv_destroy(out)


def test_server_queues_without_device(pkg):
    """Submitting and pumping before any device work: admission fails with a
    device error only when a batch is actually needed."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    prob = pkg.host.Problem(pkg.mesh.case_path(642, seed=1), use_rcm=True, layout="cfdl")
    srv = pkg.serve.Server(slots=4, member_block=2)
    mid = srv.add_mesh(prob)
    srv.submit(pkg.serve.Request(mesh=mid, inits=[prob.H_init], nsteps=3, tag=1))
    assert srv.stats()["requests_pending"] == 1
    with pytest.raises(pkg.runtime.SweHipError):
        srv.pump()
    srv.close()
