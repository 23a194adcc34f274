"""swe_hip_rk_steps_noalloc (no device allocation, RK scratch in a
library-owned device array) against swe_hip_rk_steps on the same inputs:
bitwise equal, on two calls in a row, the second on a smaller mesh (so the
second call runs over scratch the first one left behind)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_pair(pkg, n, steps, exact):
    prob = pkg.host.Problem(pkg.mesh.case_path(n, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: no-alloc steps N={n} steps={steps} exact={exact}", flush=True)
    with pkg.runtime.HipContext(prob, exact=exact) as ref, \
            pkg.runtime.HipContext(prob, exact=exact) as ctx, \
            pkg.runtime.HipContext(prob, exact=exact) as holder:
        ref.set_state(prob.H_init)
        ref.rk_steps(steps)
        want = ref.get_state()
        # the state lives in memory ctx does not own (another context's
        # state buffer, allocated before the call)
        holder.set_state(prob.H_init)
        ctx.rk_steps_noalloc(holder.state_device_ptr(), steps)
        ctx.synchronize()
        got = holder.get_state()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("exact", [False, True])
def test_noalloc_matches_rk_steps_two_calls(pkg, exact):
    _run_pair(pkg, 10242, 7, exact)
    _run_pair(pkg, 2562, 5, exact)


def test_noalloc_rejects_ensembles_and_bad_args(pkg):
    prob = pkg.host.Problem(pkg.mesh.case_path(642, seed=1), use_rcm=True, layout="cfdl")
    with pkg.runtime.HipContext(prob, members=4, member_block=2) as ens:
        with pytest.raises(pkg.runtime.SweHipError, match="single-state"):
            ens.rk_steps_noalloc(ens.state_device_ptr(), 1)
    with pkg.runtime.HipContext(prob) as ctx:
        with pytest.raises(pkg.runtime.SweHipError):
            ctx.rk_steps_noalloc(0, 1)
