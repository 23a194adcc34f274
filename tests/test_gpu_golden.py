"""HIP path against vectors made by the reference build itself
(tests/golden, see make_golden.py): exact mode must match bit for bit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_by_var

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "icos642_case.bin")
REFV = os.path.join(GOLDEN, "icos642_ref.npz")


def natural(prob, H):
    Hn = prob.to_node_major(H)[: prob.n_nodes]
    out = np.empty_like(Hn)
    out[prob.mapping] = Hn
    return out


@pytest.fixture(scope="module")
def g():
    return np.load(REFV, allow_pickle=False)


@pytest.mark.parametrize("layout", ["default", "cfdl", "sfdl", "cfdl_sfdl"])
def test_rhs_and_steps_bitwise(pkg, g, layout):
    prob = pkg.host.Problem(CASE, use_rcm=True, layout=layout, simd=4)
    print(f"case: golden 642 layout={layout}", flush=True)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        F = ctx.eval_rhs(prob.H_init)
        assert np.array_equal(natural(prob, F), g["F0"])
        ctx.set_state(prob.H_init)
        for s in range(4):
            ctx.rk_substep(s)
        assert np.array_equal(natural(prob, ctx.get_state()), g["H1"])
        ctx.set_state(prob.H_init)
        ctx.rk_steps(10)
        assert np.array_equal(natural(prob, ctx.get_state()), g["H10"])


def test_known_answer_100_steps(pkg):
    """main.c's verify_output on the reference-written H_100 slot."""
    prob = pkg.host.Problem(CASE, use_rcm=True, layout="cfdl", simd=4)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(100)
        H = ctx.get_state()
    assert prob.verify(H, prob.H_100, 0.0) == 0
    with pkg.runtime.HipContext(prob, exact=False) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(100)
        Hf = ctx.get_state()
    assert prob.verify(Hf, prob.H_100, 1e-9) == 0
    assert (rel_by_var(Hf, prob.H_100, prob) <= 1e-12).all()


@pytest.mark.parametrize("mode", ["substep", "steps"])
def test_driver_history_and_known_answer(pkg, tmp_path, mode):
    """swe_run (the reference driver on the HIP path): VERIFICATION SUCCEEDED
    on the reference-made known answer, and its history snapshots equal the
    reference build's states (natural node order) bit for bit."""
    import subprocess
    from conftest import ROOT
    hist = tmp_path / "hist.bin"
    env = dict(os.environ, SWE_INPUT_FILE=CASE, SWE_USE_NETCDF="0", SWE_NUM_COMPUTE_UNITS="1",
               SWE_NSTEPS="100", SWE_NATTEMPTS="3", SWE_USE_RCM="1", SWE_USE_OCL="0",
               SWE_LAYOUT="sfdl", SWE_HIP_MODE=mode, SWE_HIP_EXACT="1",
               SWE_OUTPUT_FILE=str(hist), SWE_HIST_INT="10")
    print(f"case: swe_run {mode}", flush=True)
    r = subprocess.run([os.path.join(ROOT, "rbf-fd-swe_amd", "swe_run")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("VERIFICATION SUCCEEDED") == 3
    assert "Eval_Rhs" in r.stdout
    steps, times, states = pkg.mesh.read_history(str(hist))
    assert steps == list(range(10, 101, 10))
    g = np.load(REFV, allow_pickle=False)
    assert np.array_equal(states[0], g["H10"])
    raw = pkg.mesh.read_bin(CASE)
    assert np.array_equal(states[-1], raw.H_final)


def test_driver_netcdf_input_known_answer(pkg, tmp_path):
    """swe_run with SWE_USE_NETCDF=1 (the reference's USE_NCIO input path) on
    the golden case written as a NetCDF classic file: VERIFICATION SUCCEEDED
    on every attempt and the same history as from the binary file."""
    import subprocess
    from conftest import ROOT
    from nc_cases import write_case_nc
    nc = tmp_path / "case.nc"
    write_case_nc(pkg.mesh.read_bin(CASE), nc, version=2)
    exe = os.path.join(ROOT, "rbf-fd-swe_amd", "swe_run")
    base = dict(os.environ, SWE_NUM_COMPUTE_UNITS="1", SWE_NSTEPS="100", SWE_NATTEMPTS="2",
                SWE_USE_RCM="1", SWE_USE_OCL="0", SWE_LAYOUT="cfdl", SWE_HIP_MODE="steps",
                SWE_HIP_EXACT="1", SWE_HIST_INT="25")
    print("case: swe_run NetCDF input", flush=True)
    hists = {}
    for fmt, path in (("1", nc), ("0", CASE)):
        hists[fmt] = tmp_path / f"hist{fmt}.bin"
        r = subprocess.run([exe], env=dict(base, SWE_INPUT_FILE=str(path), SWE_USE_NETCDF=fmt,
                                           SWE_OUTPUT_FILE=str(hists[fmt])),
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.count("VERIFICATION SUCCEEDED") == 2
        assert ("NETCDF" if fmt == "1" else "BINARY") in r.stdout
    assert hists["1"].read_bytes() == hists["0"].read_bytes()


@pytest.mark.parametrize("mode", ["substep", "steps"])
def test_driver_restart_continues_bitwise(pkg, tmp_path, mode):
    """Restart from a history snapshot (SWE_RESTART_FILE/STEP): the continued
    run's snapshots equal the uninterrupted run's bit for bit, and its final
    state still passes the known-answer check."""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "rbf-fd-swe_amd", "swe_run")
    base = dict(os.environ, SWE_INPUT_FILE=CASE, SWE_USE_NETCDF="0", SWE_NUM_COMPUTE_UNITS="1",
                SWE_NSTEPS="100", SWE_NATTEMPTS="1", SWE_USE_RCM="1", SWE_USE_OCL="0",
                SWE_LAYOUT="cfdl", SWE_HIP_MODE=mode, SWE_HIP_EXACT="1", SWE_HIST_INT="10")
    full, cont = tmp_path / "full.bin", tmp_path / "cont.bin"
    print(f"case: swe_run restart {mode}", flush=True)
    r = subprocess.run([exe], env=dict(base, SWE_OUTPUT_FILE=str(full)), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([exe], env=dict(base, SWE_OUTPUT_FILE=str(cont),
                                       SWE_RESTART_FILE=str(full), SWE_RESTART_STEP="50",
                                       SWE_NATTEMPTS="2"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Restart from step 50" in r.stdout
    assert r.stdout.count("VERIFICATION SUCCEEDED") == 2
    s_full, _, h_full = pkg.mesh.read_history(str(full))
    s_cont, _, h_cont = pkg.mesh.read_history(str(cont))
    assert s_cont == list(range(60, 101, 10))
    assert np.array_equal(h_cont, h_full[5:])
    bad = subprocess.run([exe], env=dict(base, SWE_RESTART_FILE=str(full),
                                         SWE_RESTART_STEP="55"),
                         capture_output=True, text=True, timeout=300)
    assert bad.returncode == 2 and "restart snapshot" in bad.stdout
