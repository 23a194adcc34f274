"""The C host and the CPU oracle against the reference build (oracle/_ref,
the reference's own sources) on the reference's shipped 2562 input.
Skipped where the reference tree is absent."""
import numpy as np
import pytest

from conftest import REFERENCE_BIN, have_reference

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference tree not present")

LAYOUTS = ["default", "cfdl", "sfdl", "cfdl_sfdl"]
ARRAYS = ["x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm",
          "Dx", "Dy", "Dz", "L", "idx", "H_init", "H_100"]


@pytest.fixture(scope="module", params=[(lay, rcm) for lay in LAYOUTS for rcm in (True, False)],
                ids=lambda p: f"{p[0]}-rcm{int(p[1])}")
def pair(request, pkg, pyoracle):
    lay, rcm = request.param
    if not pyoracle.ref_available(lay):
        pytest.skip("reference build missing")
    ref = pyoracle.RefBuild(REFERENCE_BIN, use_rcm=rcm, layout=lay)
    prob = pkg.host.Problem(REFERENCE_BIN, use_rcm=rcm, layout=lay, simd=4)
    return ref, prob


def test_host_prep_bitwise(pair):
    ref, prob = pair
    assert (prob.compute_size, prob.patch_size, prob.padded_nnbr) == \
        (ref.compute_size, ref.patch_size, ref.padded_nnbr)
    for nm in ARRAYS:
        assert np.array_equal(getattr(prob, nm), getattr(ref, nm)), nm
    if ref.mapping is None:
        assert prob.mapping is None
    else:
        assert np.array_equal(prob.mapping, ref.mapping)


def test_oracle_rhs_bitwise(pair, pyoracle):
    ref, prob = pair
    F_ref = ref.eval_rhs(ref.H_init)
    F = pyoracle.Oracle(prob).eval_rhs(prob.H_init)
    assert np.array_equal(F, F_ref)


def test_oracle_100_steps_bitwise_and_known_answer(pair, pyoracle):
    ref, prob = pair
    H_ref = ref.run(ref.H_init, 100)
    orc = pyoracle.Oracle(prob)
    H = orc.run(prob.H_init, 100)
    assert np.array_equal(H, H_ref)
    # main.c:439-477 known-answer check at the FP64 tolerance (swe_config.h:88)
    assert orc.verify(H, prob.H_100, 1e-9) == 0
    assert prob.verify(H, prob.H_100, 1e-9) == 0


def test_known_answer_error_profile(pkg, pyoracle):
    """FP64 vs the file's H_100 (computed elsewhere, probably MATLAB): max
    error in (1e-11, 1e-10]; without RCM two more values exceed 1e-11 -- the
    two padded copies of the last node, which RCM moves away from the end."""
    counts = {}
    for rcm in (True, False):
        prob = pkg.host.Problem(REFERENCE_BIN, use_rcm=rcm, layout="cfdl", simd=4)
        orc = pyoracle.Oracle(prob)
        H = orc.run(prob.H_init, 100)
        assert orc.verify(H, prob.H_100, 1e-10) == 0
        counts[rcm] = orc.verify(H, prob.H_100, 1e-11)
    assert counts[True] > 0
    assert counts[False] == counts[True] + 2


def test_fast_oracle_close_to_strict(pkg, pyoracle):
    prob = pkg.host.Problem(REFERENCE_BIN, use_rcm=True, layout="sfdl", simd=4)
    a = pyoracle.Oracle(prob, "strict").run(prob.H_init, 100)
    b = pyoracle.Oracle(prob, "fast").run(prob.H_init, 100)
    from conftest import rel_by_var
    assert (rel_by_var(b, a, prob) < 1e-12).all()


def _report_rows(text):
    """[phase][avg, min, max, stddev] from a profiling report's five
    '(seconds/timestep)' lines, in report order."""
    import re
    rows = []
    for line in text.splitlines():
        if "(seconds/timestep)" in line:
            rows.append([float(v) for v in re.findall(r"[-+]?\d\.\d+e[-+]\d+", line)])
    return rows


@pytest.mark.parametrize("nattempts", [1, 2, 3, 4, 7])
def test_profiling_report_matches_reference(pkg, pyoracle, nattempts):
    """swe_process_profiling against the reference's process_profiling_results
    (profiling.c:54-172) on the same per-attempt timers: the warm-up rule
    (Eval_Rhs skips attempts 0 and 1, all zeros with <= 2 attempts) and the
    STDDEV = sqrt(sum (avg - x)^2) / nattempts formula.  The timer values are
    chosen so that including the warm-up attempts, or dividing by sqrt(n),
    changes the printed digits."""
    if not pyoracle.ref_available("cfdl"):
        pytest.skip("reference build missing")
    rng = np.random.default_rng(nattempts)
    times = {k: rng.uniform(0.5, 2.0, nattempts) for k in pkg.host.PHASES}
    times["t_eval_rhs"][:2] *= 10.0   # warm-up attempts are slow
    nsteps = 37
    ours = pkg.host.process_profiling(times, nsteps)
    ref = _report_rows(pyoracle.ref_profiling(times, nsteps))
    assert len(ref) == 5
    for k in range(5):
        assert [f"{v:e}" for v in ours[k]] == [f"{v:e}" for v in ref[k]], (k, ours[k], ref[k])
    if nattempts <= 2:
        assert (ours[1, :3] == 0.0).all() and ours[1, 3] > 0.0


@pytest.fixture(scope="module", params=[(lay, rcm) for lay in LAYOUTS for rcm in (True, False)],
                ids=lambda p: f"{p[0]}-rcm{int(p[1])}")
def pair32(request, pkg, pyoracle):
    """The reference's shipped FP32 build (RUN_32BIT, swe_config.h:75) and
    the C host's FP32 reading of the same input."""
    lay, rcm = request.param
    if not pyoracle.ref_available(lay, fp32=True):
        pytest.skip("FP32 reference build missing")
    ref = pyoracle.RefBuild(REFERENCE_BIN, use_rcm=rcm, layout=lay, fp32=True)
    prob = pkg.host.Problem(REFERENCE_BIN, use_rcm=rcm, layout=lay, simd=4, fp32=True)
    return ref, prob


def test_fp32_host_prep_bitwise(pair32):
    """Every prepared array of the FP32 reading (rounding of input.c:211-319:
    values to float, weights scaled in double and rounded again) equals the
    FP32 reference build's, as floats."""
    ref, prob = pair32
    for nm in ARRAYS:
        got = getattr(prob, nm)
        want = getattr(ref, nm)
        if nm != "idx":
            assert np.array_equal(got.astype(np.float32).astype(np.float64), got), nm
            got = got.astype(np.float32)
        assert np.array_equal(got, want), nm
    assert np.float32(prob.gh0) == np.float32(ref.gh0) and np.float32(prob.dt) == np.float32(ref.dt)


def test_fp32_oracle_bitwise_and_known_answer(pair32, pyoracle):
    """The FP32 oracle equals the FP32 reference build bit for bit (RHS and 100
    steps) and passes the reference's FP32 known-answer check (4e-2,
    swe_config.h:80)."""
    ref, prob = pair32
    orc = pyoracle.Oracle(prob, "strict32")
    assert np.array_equal(orc.eval_rhs(prob.H_init), ref.eval_rhs(ref.H_init))
    H = orc.run(prob.H_init, 100)
    assert H.dtype == np.float32
    assert np.array_equal(H, ref.run(ref.H_init, 100))
    assert orc.verify(H, prob.H_100, 4e-2) == 0


def test_fp32_update_h_is_double_rounded_once(pkg, pyoracle):
    """Planted-mistake check of the FP32 stage-3 update: computing
    H += (dt/6)*D in float instead of the reference's double (the C promotion
    of rk4_rbffd_swe.c:115) changes the trajectory, so the bitwise test above
    would catch it."""
    prob = pkg.host.Problem(REFERENCE_BIN, use_rcm=True, layout="cfdl", fp32=True)
    orc = pyoracle.Oracle(prob, "strict32")
    H = prob.H_init.astype(np.float32)
    for s in range(3):
        orc.rk_substep(H, s)
    D = orc.D.copy()
    orc.rk_substep(H, 3)
    dt = np.float32(prob.dt)
    H0 = prob.H_init.astype(np.float32)
    Df = D + orc.F   # update_D with coeff 1 (float)
    H_double = (H0.astype(np.float64) + (np.float64(dt) / 6.0) * Df.astype(np.float64)).astype(np.float32)
    H_float = H0 + np.float32(np.float64(dt) / 6.0) * Df
    assert np.array_equal(H, H_double)
    assert not np.array_equal(H_double, H_float)
