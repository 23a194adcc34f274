"""The drop-in proved against the reference's own driver: swe_code/main/main.c
with the HIP back-end inserted (oracle/dropin_main.py), built from the
reference's sources by oracle/build_dropin.sh in FP64 and in the shipped FP32
type, linked to libswe_hip.so.  Its reader, RCM, layouts, patch set-up,
verify_output (main.c:439-477) and profiling report are the reference's;
only RK_substep (or the step loop) runs on the GPU.  Input: the committed
642-node case whose known-answer slot holds the reference's 100-step FP64
state."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "icos642_case.bin")


def _exe(name):
    path = os.path.join(ROOT, "oracle", "_ref", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (needs the reference tree at build time)")
    return path


def _run(exe, hip_mode, exact, attempts=3, case=CASE):
    env = dict(os.environ, SWE_INPUT_FILE=case, SWE_USE_NETCDF="0", SWE_NUM_COMPUTE_UNITS="1",
               SWE_NSTEPS="100", SWE_NATTEMPTS=str(attempts), SWE_USE_RCM="1",
               SWE_USE_OCL="0", SWE_USE_HIP=str(hip_mode), SWE_HIP_EXACT=str(int(exact)),
               OMP_NUM_THREADS="4")
    print(f"case: {os.path.basename(exe)} SWE_USE_HIP={hip_mode} exact={exact}", flush=True)
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("layout,mode,exact", [("cfdl", 1, True), ("cfdl", 2, False),
                                               ("sfdl", 1, False), ("default", 2, True),
                                               ("cfdl_sfdl", 1, True)])
def test_reference_driver_fp64(layout, mode, exact):
    out = _run(_exe(f"swe_ref_hip_{layout}.exe"), mode, exact)
    assert "HIP back-end" in out
    assert out.count("VERIFICATION SUCCEEDED") == 3, out
    assert "Eval_Rhs" in out


@pytest.mark.parametrize("layout,mode,exact", [("cfdl", 1, True), ("sfdl", 2, False)])
def test_reference_driver_fp32(layout, mode, exact):
    """The reference as shipped (RUN_32BIT): its FP32 tolerance 4e-2 against
    the FP64 known answer."""
    out = _run(_exe(f"swe_ref32_hip_{layout}.exe"), mode, exact)
    assert out.count("VERIFICATION SUCCEEDED") == 3, out


@pytest.fixture(scope="module")
def case10242(pkg, pyoracle, tmp_path_factory):
    """A 10242-node case whose known-answer slot holds the oracle's exact
    100-step state (the reference CPU build's, bit for bit), in file order."""
    path = pkg.mesh.case_path(10242, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl", simd=4)
    H = pyoracle.Oracle(prob).run(prob.H_init, 100)
    Hn = prob.to_node_major(H)[: prob.n_nodes]
    answer = np.empty_like(Hn)
    answer[prob.mapping] = Hn   # back to the file's node order
    raw = pkg.mesh.read_bin(path)
    out = str(tmp_path_factory.mktemp("dropin") / "icos10242_answer.bin")
    pkg.mesh.write_bin(out, dataclasses.replace(raw, H_final=answer))
    return out


@pytest.mark.parametrize("exe,exact", [("swe_ref_hip_cfdl.exe", True),
                                       ("swe_ref_hip_sfdl.exe", False),
                                       ("swe_ref32_hip_cfdl.exe", True)])
def test_reference_driver_resident_default(case10242, exe, exact):
    """At N = 10242 the default kernel of the fused step loop is the resident
    whole-step kernel (FP64 pair mapping, FP32 lane mapping): the
    reference's own verify_output accepts its 100 steps in every attempt."""
    out = _run(_exe(exe), 2, exact, case=case10242)
    assert out.count("VERIFICATION SUCCEEDED") == 3, out
