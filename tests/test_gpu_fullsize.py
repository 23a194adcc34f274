"""Full-size parity at the benchmark configurations (1e-12 relative
max-norm per variable against the oracle on the same inputs)."""
import numpy as np
import pytest

from conftest import rel_by_var

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,steps,kernel", [(40962, 1000, "auto"), (655362, 200, "auto"),
                                            (655362, 200, "tile")])
def test_single_member_full_size(pkg, pyoracle, n, steps, kernel):
    print(f"case: full size N={n} steps={steps} kernel={kernel}", flush=True)
    prob = pkg.host.Problem(pkg.mesh.case_path(n, seed=1), use_rcm=True, layout="cfdl")
    want = pyoracle.Oracle(prob, "strict").run(prob.H_init, steps)
    with pkg.runtime.HipContext(prob, kernel=kernel, order="hilbert") as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(steps)
        got = ctx.get_state()
    err = rel_by_var(got, want, prob)
    print("rel err per var", err, flush=True)
    assert np.isfinite(got).all()
    assert (err <= 1e-12).all()


def test_c2_exact_bitwise_1000_steps(pkg, pyoracle):
    print("case: exact N=40962 1000 steps", flush=True)
    prob = pkg.host.Problem(pkg.mesh.case_path(40962, seed=1), use_rcm=True, layout="sfdl")
    want = pyoracle.Oracle(prob, "strict").run(prob.H_init, 1000)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(1000)
        got = ctx.get_state()
    assert np.array_equal(got, want)
