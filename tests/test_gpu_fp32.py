"""The FP32 path (SWE_HIP_FP32 contexts): the reference's shipped model
precision (RUN_32BIT, fType = float, swe_config.h:75-80).

Exact mode against the FP32 oracle (itself bitwise equal to the reference's
FP32 build) and against the vectors that build wrote (tests/golden,
icos642_ref32.npz): bit for bit.  Fast mode (fused multiply-adds): the
reference's own FP32 acceptance test, |H - H_100| <= 4e-2 at every value
(swe_config.h:80, main.c:439-477), plus a relative max-norm per variable
of at most 1e-3 against the exact FP32 trajectory."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_by_var

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "icos642_case.bin")
REFV32 = os.path.join(GOLDEN, "icos642_ref32.npz")


def natural(prob, H):
    Hn = prob.to_node_major(H)[: prob.n_nodes]
    out = np.empty_like(Hn)
    out[prob.mapping] = Hn
    return out


@pytest.mark.parametrize("layout", ["default", "cfdl", "sfdl", "cfdl_sfdl"])
def test_fp32_exact_golden_bitwise(pkg, layout):
    g = np.load(REFV32, allow_pickle=False)
    prob = pkg.host.Problem(CASE, use_rcm=True, layout=layout, simd=4, fp32=True)
    print(f"case: FP32 exact golden 642 layout={layout}", flush=True)
    with pkg.runtime.HipContext(prob, exact=True, fp32=True) as ctx:
        assert np.array_equal(natural(prob, ctx.eval_rhs(prob.H_init)), g["F0"])
        ctx.set_state(prob.H_init)
        for s in range(4):
            ctx.rk_substep(s)
        assert np.array_equal(natural(prob, ctx.get_state()), g["H1"])
        ctx.set_state(prob.H_init)
        ctx.rk_steps(10)
        assert np.array_equal(natural(prob, ctx.get_state()), g["H10"])
        ctx.rk_steps(90)
        H = ctx.get_state()
        assert np.array_equal(natural(prob, H), g["H100"])
        assert prob.verify(H, prob.H_100, 4e-2) == 0


@pytest.mark.parametrize("n,layout", [(2562, "cfdl"), (10242, "sfdl")])
def test_fp32_exact_vs_oracle(pkg, pyoracle, n, layout):
    prob = pkg.host.Problem(pkg.mesh.case_path(n, seed=1), use_rcm=True, layout=layout,
                            fp32=True)
    print(f"case: FP32 exact N={n} {layout} vs oracle", flush=True)
    orc = pyoracle.Oracle(prob, "strict32")
    with pkg.runtime.HipContext(prob, exact=True, fp32=True) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(20)
        assert np.array_equal(ctx.get_state(), orc.run(prob.H_init, 20))


@pytest.mark.parametrize("n,steps,kernel", [(642, 100, "auto"), (642, 100, "lane"),
                                            (642, 100, "quad"), (642, 100, "oct"),
                                            (40962, 1000, "auto"), (40962, 1000, "lane")])
def test_fp32_fast_known_answer(pkg, pyoracle, n, steps, kernel):
    """Fast FP32 kernels (one lane per node; 2 (auto below 262144 nodes), 4 or
    8 lanes per node splitting the stencil slots) within 1e-3 of the strict
    FP32 oracle."""
    path = CASE if n == 642 else pkg.mesh.case_path(n, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl", fp32=True)
    print(f"case: FP32 fast N={n} {steps} steps kernel {kernel}", flush=True)
    want = pyoracle.Oracle(prob, "strict32").run(prob.H_init, steps)
    with pkg.runtime.HipContext(prob, fp32=True, kernel=kernel) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(steps)
        got = ctx.get_state()
    err = rel_by_var(got.astype(np.float64), want.astype(np.float64), prob)
    print("FP32 fast vs exact rel err per var", err, flush=True)
    assert np.isfinite(got).all()
    assert (err <= 1e-3).all()
    if n == 642:   # the file's known answer (the FP64 reference's 100-step state)
        assert prob.verify(got, prob.H_100, 4e-2) == 0


def test_fp32_entry_points_reject_the_other_precision(pkg):
    prob = pkg.host.Problem(CASE, use_rcm=True, layout="cfdl", fp32=True)
    with pkg.runtime.HipContext(prob, fp32=True) as c32, pkg.runtime.HipContext(prob) as c64:
        L = pkg.runtime.lib()
        H = np.zeros(4 * prob.patch_size)
        assert L.swe_hip_set_state(c32.handle, 0, H.ctypes.data_as(pkg.runtime._dp)) == -1
        H32 = np.zeros(4 * prob.patch_size, dtype=np.float32)
        assert L.swe_hip_set_state32(c64.handle, 0, H32.ctypes.data_as(pkg.runtime._fp)) == -1
        with pytest.raises(pkg.runtime.SweHipError):
            c32.rk_steps_noalloc(c32.state_device_ptr(), 1)


def test_fp32_tuner_and_split_kernels_agree(pkg):
    """swe_hip_tune_ex over the FP32 mappings (and the resident kernel) returns a
    fast-mode kernel, and
    every split mapping gives the same state after 10 steps (the group sums
    are formed alike in every lane; the mappings differ only in the order of
    the partial sums, so within 1e-4 of one another, FP32 rounding)."""
    prob = pkg.host.Problem(pkg.mesh.case_path(10242, seed=1), use_rcm=True, layout="cfdl",
                            fp32=True)
    print("case: FP32 tuner N=10242", flush=True)
    with pkg.runtime.HipContext(prob, fp32=True, kernel="lane") as base:
        opts, ms = base.tune(members=1, steps=3)
        assert opts["kernel"] in ("lane", "pair", "quad", "oct", "tile", "resident") and ms > 0
        out = {}
        for k in ("lane", "pair", "quad", "oct"):
            with pkg.runtime.HipContext(prob, fp32=True, share=base, kernel=k) as ctx:
                ctx.set_state(prob.H_init)
                ctx.rk_steps(10)
                out[k] = ctx.get_state().astype(np.float64)
    for k in ("pair", "quad", "oct"):
        err = rel_by_var(out[k], out["lane"], prob)
        assert (err <= 1e-4).all(), (k, err)


@pytest.mark.parametrize("tile_nodes", [64, 128, 256])
def test_fp32_tile_kernel_matches_lane_kernel_bitwise(pkg, pyoracle, tile_nodes):
    """The FP32 LDS-tile kernel sums every stencil in the lane kernel's order:
    exact mode bitwise equal to the strict FP32 oracle, fast mode bitwise
    equal to the one-lane fast kernel (Hilbert-ordered device copy)."""
    prob = pkg.host.Problem(pkg.mesh.case_path(10242, seed=1), use_rcm=True, layout="cfdl",
                            fp32=True)
    print(f"case: FP32 tile kernel N=10242 tile_nodes={tile_nodes}", flush=True)
    want = pyoracle.Oracle(prob, "strict32").run(prob.H_init, 5)
    with pkg.runtime.HipContext(prob, fp32=True, exact=True, kernel="tile", order="hilbert",
                                tile_nodes=tile_nodes) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(5)
        assert np.array_equal(ctx.get_state(), want)
    out = {}
    for k in ("lane", "tile"):
        with pkg.runtime.HipContext(prob, fp32=True, kernel=k, order="hilbert",
                                    tile_nodes=tile_nodes if k == "tile" else 0) as ctx:
            ctx.set_state(prob.H_init)
            ctx.rk_steps(20)
            out[k] = ctx.get_state()
    assert np.array_equal(out["tile"], out["lane"])
