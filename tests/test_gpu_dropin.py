"""The drop-in proved against the reference's own driver: swe_code/main/main.c
with the HIP back-end inserted (oracle/dropin_main.py), built from the
reference's sources by oracle/build_dropin.sh in FP64 and in the shipped FP32
type, linked to libswe_hip.so.  Its reader, RCM, layouts, patch set-up,
verify_output (main.c:439-477) and profiling report are the reference's;
only RK_substep (or the step loop) runs on the GPU.  Input: the committed
642-node case whose known-answer slot holds the reference's 100-step FP64
state."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "icos642_case.bin")


def _exe(name):
    path = os.path.join(ROOT, "oracle", "_ref", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (needs the reference tree at build time)")
    return path


def _run(exe, hip_mode, exact, attempts=3):
    env = dict(os.environ, SWE_INPUT_FILE=CASE, SWE_USE_NETCDF="0", SWE_NUM_COMPUTE_UNITS="1",
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
