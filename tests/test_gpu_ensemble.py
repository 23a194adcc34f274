"""Ensembles: every member of a batched launch equals its solo run."""
import numpy as np
import pytest

from conftest import rel_by_var

pytestmark = pytest.mark.gpu


def _members(pkg, prob, B, seed=11):
    case = pkg.mesh.read_bin(pkg.mesh.case_path(prob.n_nodes, seed=1))
    E = pkg.mesh.perturb_members(case, B, seed=seed)
    out = []
    for m in range(B):
        Hn = E[m][prob.mapping] if prob.mapping is not None else E[m]
        full = np.concatenate([Hn, np.repeat(Hn[-1:], prob.padded - prob.n_nodes, 0)])
        out.append(prob.from_node_major(full))
    return out


@pytest.mark.parametrize("B", [2, 4, 48, 64])
def test_exact_members_equal_oracle(pkg, pyoracle, B):
    prob = pkg.host.Problem(pkg.mesh.case_path(642, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: exact ensemble B={B} N=642", flush=True)
    H0 = _members(pkg, prob, B)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, members=B, exact=True) as ctx:
        for m in range(B):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(5)
        for m in sorted({0, 1, B // 2, B - 1}):
            assert np.array_equal(ctx.get_state(m), orc.run(H0[m], 5)), m


def test_fast_batch_composition_invariance(pkg):
    """A member's state after fused fast steps is bitwise the same whether it
    runs alone (single-state kernel) or inside a 64-member batch."""
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print("case: composition invariance B=64 N=2562", flush=True)
    H0 = _members(pkg, prob, 64)
    with pkg.runtime.HipContext(prob, members=64) as ctx:
        for m in range(64):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(20)
        batch = [ctx.get_state(m) for m in (0, 37)]
    for k, m in enumerate((0, 37)):
        with pkg.runtime.HipContext(prob, members=1) as solo:
            solo.set_state(H0[m])
            solo.rk_steps(20)
            assert np.array_equal(solo.get_state(), batch[k])


def test_fast_members_vs_oracle(pkg, pyoracle):
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="sfdl")
    print("case: fast ensemble B=16 N=2562 sfdl 50 steps", flush=True)
    H0 = _members(pkg, prob, 16)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, members=16) as ctx:
        for m in range(16):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(50)
        for m in (3, 15):
            err = rel_by_var(ctx.get_state(m), orc.run(H0[m], 50), prob)
            assert (err <= 1e-12).all(), (m, err)


def test_autotuner_picks_admissible_block(pkg):
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print("case: member-block autotuner B=64 N=2562", flush=True)
    with pkg.runtime.HipContext(prob) as base:
        mb, ms = base.tune_member_block(64, steps=3)
        assert mb in (2, 4, 8, 16, 32, 64) and ms > 0.0
        with pkg.runtime.HipContext(prob, members=64, member_block=mb, share=base) as ctx:
            assert ctx.static_bytes() == base.static_bytes()   # shared, not copied
            n = prob.compute_size
            # H, Ka, Kb, D (the drop-in F only once a substep runs) + staging
            assert ctx.device_bytes() == 4 * 4 * n * 64 * 8 + 4 * n * 8 + 4
            ctx.set_state(prob.H_init, 5)
            ctx.rk_steps(2)
            base.set_state(prob.H_init)
            base.rk_steps(2)
            assert np.array_equal(ctx.get_state(5), base.get_state())


def test_ensemble_tuner_covers_tiled_kernels(pkg):
    """The ensemble tuner times the tiled kernels too and its choice runs:
    the tuned context equals the single-state path bitwise (fast mode)."""
    prob = pkg.host.Problem(pkg.mesh.case_path(10242, seed=1), use_rcm=True, layout="cfdl")
    print("case: ensemble tuner B=64 N=10242 hilbert", flush=True)
    with pkg.runtime.HipContext(prob, order="hilbert") as base:
        mb, tiled, ms = base.tune_ensemble(64, steps=3)
        assert mb in (2, 4, 8, 16, 32, 64) and ms > 0.0
        assert not tiled or mb == 32
        with pkg.runtime.HipContext(prob, members=64, member_block=mb, tiled=tiled,
                                    share=base) as ctx:
            ctx.set_state(prob.H_init, 9)
            ctx.rk_steps(3)
            base.set_state(prob.H_init)
            base.rk_steps(3)
            assert np.array_equal(ctx.get_state(9), base.get_state())


def test_set_active_and_move_member(pkg):
    """Slots beyond the active count are skipped; a moved member continues
    exactly as if it had not moved."""
    prob = pkg.host.Problem(pkg.mesh.case_path(642, seed=1), use_rcm=True, layout="cfdl")
    H0 = _members(pkg, prob, 4)
    with pkg.runtime.HipContext(prob, members=8, member_block=2) as ctx:
        for m in range(4):
            ctx.set_state(H0[m], m)
        ctx.set_state(H0[3], 7)
        ctx.set_active(4)                 # member block of slots 6,7 idle
        ctx.rk_steps(3)
        frozen = ctx.get_state(7)
        assert np.array_equal(frozen, H0[3])
        ctx.move_member(3, 1)             # slot 1 <- slot 3
        ctx.set_active(2)
        ctx.rk_steps(2)
        got = ctx.get_state(1)
    with pkg.runtime.HipContext(prob) as solo:
        solo.set_state(H0[3])
        solo.rk_steps(5)
        assert np.array_equal(solo.get_state(), got)


@pytest.mark.parametrize("B,mb,order", [(32, 32, "morton"), (64, 32, "morton"),
                                        (96, 32, "morton"), (64, 32, "hilbert"),
                                        (32, 32, "host")])
def test_tiled_exact_members_equal_oracle(pkg, pyoracle, B, mb, order):
    """The member-lane LDS-tile kernel over any device node order: bitwise
    equal to the oracle (node order and staging never change a node's
    arithmetic)."""
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="sfdl")
    print(f"case: tiled exact B={B} mb={mb} order={order} N=2562", flush=True)
    H0 = _members(pkg, prob, B)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, members=B, member_block=mb, exact=True,
                                order=order, tiled=True) as ctx:
        for m in range(B):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(4)
        for m in sorted({0, B // 3, B - 1}):
            assert np.array_equal(ctx.get_state(m), orc.run(H0[m], 4)), m


def test_member_lane_tiles_host_order_substeps_exact(pkg, pyoracle):
    """Member-lane LDS tiles on the caller's (RCM) node order: the drop-in
    substep path (RHS kernel + separate axpys) and fused steps both give the
    oracle's bits."""
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print("case: member-lane tiles B=32 host order, substeps + fused, exact", flush=True)
    H0 = _members(pkg, prob, 32)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, members=32, member_block=32, exact=True,
                                tiled=True) as ctx:
        for m in range(32):
            ctx.set_state(H0[m], m)
        for s in range(4):
            ctx.rk_substep(s)
        ctx.rk_steps(2)
        for m in (0, 13, 31):
            assert np.array_equal(ctx.get_state(m), orc.run(H0[m], 3)), m


@pytest.mark.parametrize("mb,order", [(32, "morton"), (32, "hilbert")])
def test_tiled_fast_composition_and_active(pkg, mb, order):
    prob = pkg.host.Problem(pkg.mesh.case_path(10242, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: tiled fast B=64 mb={mb} order={order} N=10242 vs solo", flush=True)
    H0 = _members(pkg, prob, 64)
    with pkg.runtime.HipContext(prob, members=64, member_block=mb, order=order,
                                tiled=True) as ctx:
        for m in range(64):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(12)
        got = {m: ctx.get_state(m) for m in (0, 17, 63)}
        ctx.set_active(16)
        ctx.rk_steps(3)
        got15 = ctx.get_state(15)
    with pkg.runtime.HipContext(prob) as solo:
        for m in (0, 17, 63):
            solo.set_state(H0[m])
            solo.rk_steps(12)
            assert np.array_equal(solo.get_state(), got[m]), m
        solo.set_state(H0[15])
        solo.rk_steps(15)
        assert np.array_equal(solo.get_state(), got15)


@pytest.mark.parametrize("kernel,pf,sblk,order", [("quadw", 0, 0, "morton"),
                                                 ("quadw", 8, 0, "host"),
                                                 ("quadw", 4, 16, "morton"),
                                                 ("pair", 0, 0, "morton"), ("pair", 0, 16, "morton"),
                                                 ("pair", 0, 0, "hilbert"), ("tile", 0, 0, "hilbert"),
                                                 ("tile", 0, 16, "morton"), ("tile", 0, 0, "host")])
def test_single_state_variants_exact(pkg, pyoracle, kernel, pf, sblk, order):
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: single-state {kernel} prefetch={pf} slot_block={sblk} {order} exact", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 3)
    with pkg.runtime.HipContext(prob, exact=True, kernel=kernel, prefetch=pf,
                                order=order, slot_block=sblk) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(3)
        assert np.array_equal(ctx.get_state(), want)


@pytest.mark.parametrize("kernel,order,sblk", [("pair", "host", 0), ("pair", "morton", 16),
                                               ("pair", "hilbert", 0), ("quadw", "morton", 16),
                                               ("quadw", "host", 0), ("tile", "hilbert", 0),
                                               ("tile", "host", 16)])
def test_single_state_variants_fast(pkg, pyoracle, kernel, order, sblk):
    """Fast single-state kernels within 1e-12 of the oracle after 50 steps."""
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: single-state fast {kernel} {order} slot_block={sblk} 50 steps", flush=True)
    want = pyoracle.Oracle(prob).run(prob.H_init, 50)
    with pkg.runtime.HipContext(prob, kernel=kernel, order=order, slot_block=sblk) as ctx:
        ctx.set_state(prob.H_init)
        ctx.rk_steps(50)
        err = rel_by_var(ctx.get_state(), want, prob)
        assert (err <= 1e-12).all(), err


@pytest.mark.parametrize("tiled,mb", [(False, 64), (False, 8), (True, 32)])
def test_blocked_stencil_layout_members_exact(pkg, pyoracle, tiled, mb):
    """Node-blocked stencil arrays (slot_block=16, N=642 is not a multiple
    of 16) give the oracle's bits for the ensemble kernels too."""
    prob = pkg.host.Problem(pkg.mesh.case_path(642, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: slot_block=16 members tiled={tiled} mb={mb}", flush=True)
    H0 = _members(pkg, prob, 64)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, members=64, member_block=mb, exact=True, tiled=tiled,
                                order="morton" if tiled else "host", slot_block=16) as ctx:
        for m in range(64):
            ctx.set_state(H0[m], m)
        ctx.rk_steps(3)
        for m in (0, 33, 63):
            assert np.array_equal(ctx.get_state(m), orc.run(H0[m], 3)), m


@pytest.mark.parametrize("n", [642, 10242])
def test_tile_kernel_fast_equals_pair_and_drop_in(pkg, pyoracle, n):
    """The LDS-tile single-state kernel runs the pair kernel's arithmetic:
    fused fast steps bitwise equal to the pair kernel's, and the drop-in
    substeps and a plain RHS bitwise equal to the oracle in exact mode."""
    prob = pkg.host.Problem(pkg.mesh.case_path(n, seed=1), use_rcm=True, layout="sfdl")
    print(f"case: tile kernel N={n} vs pair (fast) and oracle (exact substeps)", flush=True)
    with pkg.runtime.HipContext(prob, kernel="tile", order="hilbert") as t, \
            pkg.runtime.HipContext(prob, kernel="pair", order="hilbert") as p:
        for c in (t, p):
            c.set_state(prob.H_init)
            c.rk_steps(9)
        assert np.array_equal(t.get_state(), p.get_state())
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, kernel="tile", order="hilbert", exact=True) as t:
        assert np.array_equal(t.eval_rhs(prob.H_init), orc.eval_rhs(prob.H_init))
        t.set_state(prob.H_init)
        for s in range(4):
            t.rk_substep(s)
        t.rk_steps(1)
        assert np.array_equal(t.get_state(), orc.run(prob.H_init, 2))


@pytest.mark.parametrize("tn,pf", [(64, 0), (64, 4), (128, 8), (256, 4), (256, 8)])
def test_tile_kernel_shapes_and_staging_exact(pkg, pyoracle, tn, pf):
    """Every tile size and row-staging schedule of the LDS-tile single-state
    kernel gives the oracle's bits (fused steps and the drop-in RHS)."""
    prob = pkg.host.Problem(pkg.mesh.case_path(2562, seed=1), use_rcm=True, layout="cfdl")
    print(f"case: tile kernel tile_nodes={tn} prefetch={pf} exact", flush=True)
    orc = pyoracle.Oracle(prob)
    with pkg.runtime.HipContext(prob, exact=True, kernel="tile", order="hilbert",
                                tile_nodes=tn, prefetch=pf) as ctx:
        assert np.array_equal(ctx.eval_rhs(prob.H_init), orc.eval_rhs(prob.H_init))
        ctx.set_state(prob.H_init)
        ctx.rk_steps(3)
        assert np.array_equal(ctx.get_state(), orc.run(prob.H_init, 3))
