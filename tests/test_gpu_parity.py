"""HIP path vs the CPU oracle on the same inputs (GPU).

Exact mode (SWE_HIP_EXACT) must be bitwise equal to the oracle, which is
itself bitwise equal to the reference build; the fast mode (FMA) must agree
to 1e-12 relative max-norm per variable."""
import numpy as np
import pytest

from conftest import rel_by_var

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _prob(pkg, n, layout="cfdl", seed=1, rcm=True, simd=4):
    path = pkg.mesh.case_path(n, seed=seed)
    return pkg.host.Problem(path, use_rcm=rcm, layout=layout, simd=simd)


@pytest.mark.parametrize("n", [642, 2562])
def test_exact_rhs_bitwise(pkg, pyoracle, n):
    prob = _prob(pkg, n)
    print(f"case: exact RHS N={n}", flush=True)
    want = pyoracle.Oracle(prob).eval_rhs(prob.H_init)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        got = ctx.eval_rhs(prob.H_init)
    assert np.array_equal(got, want)


def test_exact_substeps_bitwise(pkg, pyoracle):
    prob = _prob(pkg, 2562)
    print("case: exact drop-in substeps N=2562 x10", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 10)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        ctx.set_state(prob.H_init)
        for _ in range(10):
            for s in range(4):
                ctx.rk_substep(s)
        got = ctx.get_state()
    assert np.array_equal(got, want)


def test_exact_fused_bitwise(pkg, pyoracle):
    prob = _prob(pkg, 2562)
    print("case: exact fused steps N=2562 x11", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 11)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(11)
        got = ctx.get_state()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("layout", ["default", "cfdl", "sfdl", "cfdl_sfdl"])
def test_layouts_exact(pkg, pyoracle, layout):
    prob = _prob(pkg, 2562, layout=layout)
    print(f"case: layout {layout} exact 3 steps", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 3)
    with pkg.runtime.HipContext(prob, exact=True) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(3)
        got = ctx.get_state()
    assert np.array_equal(got, want)


def test_fast_100_steps(pkg, pyoracle):
    prob = _prob(pkg, 2562)
    print("case: fast fused N=2562 x100", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 100)
    with pkg.runtime.HipContext(prob, exact=False) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(100)
        got = ctx.get_state()
    err = rel_by_var(got, want, prob)
    print("rel err per var", err)
    assert (err <= TOL).all()
    assert prob.verify(got, want, 1e-9) == 0
