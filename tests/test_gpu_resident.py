"""The resident whole-step kernel (SWE_HIP_KERNEL_RESIDENT): all RK steps of
a call in one cooperative launch, grid barriers between the stages.

It keeps the arithmetic of the pair mapping (rk4_rbffd_swe.c:30-72 stages,
:285-353 sums and RHS), so exact mode is bitwise equal to the oracle and
fast mode bitwise equal to the stage kernels replayed from graphs, at every
size, node order and stencil layout, over repeated launches.  A stale read
of a neighbour's state across the XCDs' caches would change the bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _prob(pkg, n, seed=1, rcm=True, layout="cfdl"):
    return pkg.host.Problem(pkg.mesh.case_path(n, seed=seed), use_rcm=rcm, layout=layout, simd=4)


@pytest.mark.parametrize("n,order,slot_block", [(642, "host", 0), (2562, "hilbert", 0),
                                                (2562, "host", 16)])
def test_exact_bitwise_vs_oracle(pkg, pyoracle, n, order, slot_block):
    prob = _prob(pkg, n)
    print(f"case: resident exact N={n} order={order} slot_block={slot_block} 4+7 steps",
          flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 11)
    with pkg.runtime.HipContext(prob, exact=True, kernel="resident", order=order,
                                slot_block=slot_block) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(4)
        ctx.rk_steps(0)
        ctx.rk_steps(7)
        got = ctx.get_state()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,order,steps", [(2562, "hilbert", 300), (10242, "morton", 200),
                                           (40962, "hilbert", 1000), (40962, "host", 100)])
def test_fast_bitwise_vs_stage_kernels(pkg, n, order, steps):
    prob = _prob(pkg, n)
    print(f"case: resident fast N={n} order={order} {steps} steps", flush=True)
    with pkg.runtime.HipContext(prob, kernel="pair", order=order) as ref:
        ref.set_state(prob.H_init)
        ref.rk_steps(steps)
        want = ref.get_state()
    with pkg.runtime.HipContext(prob, kernel="resident", order=order) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(steps // 2)
        ctx.rk_steps(steps - steps // 2)
        got = ctx.get_state()
    assert np.isfinite(got).all()
    assert np.array_equal(got, want)


def test_auto_mixes_graph_and_resident_calls(pkg):
    """AUTO picks the resident kernel at N=40962 and replays the stage graphs
    for calls shorter than 4 steps; any mix of calls gives the same bits."""
    prob = _prob(pkg, 40962)
    print("case: auto at N=40962, calls of 1, 2, 3, 10 and 4 steps", flush=True)
    with pkg.runtime.HipContext(prob, kernel="pair") as ref:
        ref.set_state(prob.H_init)
        ref.rk_steps(20)
        want = ref.get_state()
    with pkg.runtime.HipContext(prob) as ctx:
        ctx.set_state(prob.H_init)
        for k in (1, 2, 3, 10, 4):
            ctx.rk_steps(k)
        got = ctx.get_state()
    assert np.array_equal(got, want)


def test_substeps_and_rhs_of_a_resident_context(pkg, pyoracle):
    """Drop-in substeps and eval_rhs of a resident context run the pair
    mapping; mixing them with resident steps keeps the bits."""
    prob = _prob(pkg, 2562)
    print("case: resident context, substeps + steps + RHS, exact", flush=True)
    orc = pyoracle.Oracle(prob)
    want = orc.run(prob.H_init, 5)
    with pkg.runtime.HipContext(prob, exact=True, kernel="resident") as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(2)
        for s in range(4):
            ctx.rk_substep(s)
        ctx.rk_steps(2)
        got = ctx.get_state()
        rhs = ctx.eval_rhs(got)
    assert np.array_equal(got, want)
    assert np.array_equal(rhs, orc.eval_rhs(want))


@pytest.mark.parametrize("n,exact,order,slot_block,steps", [
    (2562, True, "host", 0, 11), (2562, True, "hilbert", 16, 11), (40962, False, "hilbert", 0, 300),
    (40962, True, "host", 0, 20)])
def test_fp32_bitwise_vs_lane_kernel(pkg, n, exact, order, slot_block, steps):
    """FP32: the resident kernel keeps the lane kernel's float sums, RHS and
    stage update (and its double update_H), bit for bit in both modes."""
    prob = pkg.host.Problem(pkg.mesh.case_path(n, seed=1), use_rcm=True, layout="cfdl", simd=4,
                            fp32=True)
    print(f"case: FP32 resident N={n} exact={exact} order={order} slot_block={slot_block} "
          f"{steps} steps", flush=True)
    kw = dict(fp32=True, exact=exact, order=order, slot_block=slot_block)
    with pkg.runtime.HipContext(prob, kernel="lane", **kw) as ref:
        ref.set_state(prob.H_init)
        ref.rk_steps(steps)
        want = ref.get_state()
    with pkg.runtime.HipContext(prob, kernel="resident", **kw) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(steps // 3)
        ctx.rk_steps(steps - steps // 3)
        got = ctx.get_state()
    assert np.isfinite(got).all()
    assert np.array_equal(got, want)


def test_fp32_auto_is_resident_and_mixes_calls(pkg):
    prob = pkg.host.Problem(pkg.mesh.case_path(40962, seed=1), use_rcm=True, layout="cfdl",
                            simd=4, fp32=True)
    print("case: FP32 auto at N=40962, calls of 1, 3, 8 steps", flush=True)
    with pkg.runtime.HipContext(prob, fp32=True, kernel="lane") as ref:
        ref.set_state(prob.H_init)
        ref.rk_steps(12)
        want = ref.get_state()
    with pkg.runtime.HipContext(prob, fp32=True) as ctx:
        ctx.set_state(prob.H_init)
        for k in (1, 3, 8):
            ctx.rk_steps(k)
        got = ctx.get_state()
    assert np.array_equal(got, want)


def test_two_contexts_in_flight_at_once(pkg):
    """Two resident contexts (two meshes, two streams) with their launches
    queued back to back before either is waited for: both give the bits of
    their stage-kernel runs (the runtime serialises or co-schedules the
    cooperative launches; every wait is bounded either way)."""
    pa, pb = _prob(pkg, 10242), _prob(pkg, 40962)
    print("case: resident N=10242 and N=40962 launched back to back on two streams",
          flush=True)
    want = []
    for p in (pa, pb):
        with pkg.runtime.HipContext(p, kernel="pair", order="hilbert") as ref:
            ref.set_state(p.H_init)
            ref.rk_steps(200)
            want.append(ref.get_state())
    with pkg.runtime.HipContext(pa, kernel="resident", order="hilbert") as ca, \
            pkg.runtime.HipContext(pb, kernel="resident", order="hilbert") as cb:
        ca.set_state(pa.H_init)
        cb.set_state(pb.H_init)
        for _ in range(4):
            ca.rk_steps(50)
            cb.rk_steps(50)
        ca.synchronize()
        cb.synchronize()
        assert np.array_equal(ca.get_state(), want[0])
        assert np.array_equal(cb.get_state(), want[1])


def test_wait_time_out_is_reported_and_recovered(pkg, monkeypatch):
    """A wait that exceeds its poll limit aborts the launch in every tile and
    is reported at the next synchronising call; the next launch starts from
    cleared counters.  The time-out is forced from the host with a limit of
    one poll on the kernel's own bounded wait (every wave still exits)."""
    prob = _prob(pkg, 10242)
    print("case: resident N=10242, wait poll limit 1 (forced time-out), then a normal run",
          flush=True)
    with pkg.runtime.HipContext(prob, kernel="pair", order="hilbert") as ref:
        ref.set_state(prob.H_init)
        ref.rk_steps(30)
        want = ref.get_state()
    monkeypatch.setenv("SWE_HIP_RESIDENT_SPIN_LIMIT", "1")
    with pkg.runtime.HipContext(prob, kernel="resident", order="hilbert") as ctx:
        monkeypatch.delenv("SWE_HIP_RESIDENT_SPIN_LIMIT")
        ctx.set_state(prob.H_init)
        errors = 0
        for _ in range(20):   # a time-out is likely in every launch, certain in some
            ctx.rk_steps(30)
            try:
                ctx.synchronize()
            except pkg.runtime.SweHipError as e:
                assert e.code == -3 and "limit" in str(e)
                errors += 1
        assert errors > 0
    with pkg.runtime.HipContext(prob, kernel="resident", order="hilbert") as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(30)
        assert np.array_equal(ctx.get_state(), want)


def test_too_many_nodes_per_compute_unit_is_rejected(pkg):
    prob = _prob(pkg, 163842)
    print("case: resident kernel at N=163842 (above 256 nodes per CU)", flush=True)
    with pytest.raises(pkg.runtime.SweHipError) as ei:
        pkg.runtime.HipContext(prob, kernel="resident")
    assert ei.value.code == -1


def test_scattered_order_rejects_resident_and_tuner_skips_it(pkg):
    """In the generator's own node order (no RCM, host order) the tiles of
    consecutive nodes are scattered over the sphere (more than 64
    neighbouring tiles): a resident context is refused with SWE_HIP_EINVAL,
    and the autotuner passes over that candidate instead of failing."""
    prob = _prob(pkg, 10242, rcm=False)
    print("case: resident kernel and tuner on a scattered host order, N=10242", flush=True)
    with pytest.raises(pkg.runtime.SweHipError) as ei:
        pkg.runtime.HipContext(prob, kernel="resident", order="host")
    assert ei.value.code == -1
    with pkg.runtime.HipContext(prob, kernel="pair", order="host") as base:
        opts, ms = base.tune(1, steps=2)
    assert opts["kernel"] != "resident" and ms > 0
