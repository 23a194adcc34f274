"""Parity at the benchmarked batch configurations.

* 64-member ensemble at N=163842 on the member-lane LDS-tile kernel
  (member block 32, Hilbert-ordered device copy, persistent items), built
  the way bench.py builds it: fast steps within 1e-12 of the strict oracle
  and bitwise equal to the single-state path; exact steps bitwise equal to
  the oracle (rk4_rbffd_swe.c:30-72, 189-362).
* The serving mix over every benchmarked mesh (N = 2562 ... 163842), with
  up to 32 members and at least 50 steps per request: every result equals
  its solo run bitwise.
* The server under an HBM cap that holds one batch of either of two meshes:
  idle batches are evicted and every request finishes.
"""
import numpy as np
import pytest

from conftest import rel_by_var

pytestmark = pytest.mark.gpu

MESHES = (2562, 10242, 40962, 163842)


def _problem(pkg, n):
    path = pkg.mesh.case_path(n, seed=1)
    p = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    p.path = path
    return p


@pytest.fixture(scope="module")
def c4(pkg):
    prob = _problem(pkg, 163842)
    inits = pkg.serve.member_pool(pkg.mesh, prob, 64, seed=4)
    return prob, inits


def test_c4_fast_tiles_vs_oracle_and_single(pkg, pyoracle, c4):
    prob, inits = c4
    print("case: C4 N=163842 B=64 member block 32 tiled hilbert, 10 fast steps", flush=True)
    picks = (0, 31, 32, 63)
    with pkg.runtime.HipContext(prob, order="hilbert") as base:
        with pkg.runtime.HipContext(prob, members=64, member_block=32, tiled=True,
                                    share=base) as ens:
            for m in range(64):
                ens.set_state(inits[m], m)
            ens.rk_steps(10)
            got = {m: ens.get_state(m) for m in picks}
        for m in picks:
            base.set_state(inits[m])
            base.rk_steps(10)
            assert np.array_equal(base.get_state(), got[m]), m
    orc = pyoracle.Oracle(prob, "strict")
    for m in (0, 31, 63):
        err = rel_by_var(got[m], orc.run(inits[m], 10), prob)
        print(f"member {m} rel err per var {err}", flush=True)
        assert (err <= 1e-12).all(), (m, err)


def test_c4_exact_tiles_bitwise_oracle(pkg, pyoracle, c4):
    prob, inits = c4
    print("case: C4 N=163842 B=64 member block 32 tiled hilbert, 2 exact steps", flush=True)
    with pkg.runtime.HipContext(prob, members=64, member_block=32, tiled=True,
                                order="hilbert", exact=True) as ens:
        for m in range(64):
            ens.set_state(inits[m], m)
        ens.rk_steps(2)
        got = {m: ens.get_state(m) for m in (0, 31, 63)}
    orc = pyoracle.Oracle(prob, "strict")
    for m, H in got.items():
        assert np.array_equal(H, orc.run(inits[m], 2)), m


def test_c5_serving_all_meshes_equals_solo(pkg, pyoracle):
    probs = [_problem(pkg, n) for n in MESHES]
    pools = [pkg.serve.member_pool(pkg.mesh, p, 32, seed=p.n_nodes) for p in probs]
    # (mesh, members, steps, arrival s): full 32-member requests on every mesh,
    # joins while other batches run, partial batches, >= 50 steps each
    plan = [(3, 32, 60, 0.00), (2, 17, 55, 0.00), (1, 32, 50, 0.01), (0, 5, 80, 0.01),
            (3, 3, 52, 0.02), (2, 32, 50, 0.03), (0, 32, 64, 0.03), (3, 9, 70, 0.05)]
    reqs = []
    for tag, (mid, B, steps, t) in enumerate(plan):
        inits = [pools[mid][(tag + m) % 32] for m in range(B)]
        reqs.append(pkg.serve.Request(mesh=mid, inits=inits, nsteps=steps, tag=tag,
                                      arrival_s=t))
    print(f"case: serving mix over N={MESHES}, {len(reqs)} requests, tiled 64-slot batches",
          flush=True)
    with pkg.serve.Server(slots=64, member_block=32, chunk_steps=16, tiled=True) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs)
    print("stats", stats, flush=True)
    assert stats["requests_done"] == len(reqs)
    assert stats["member_steps_done"] >= sum(len(r.inits) * r.nsteps for r in reqs)
    ctxs = [pkg.runtime.HipContext(p) for p in probs]
    try:
        for r in reqs:
            for m, H0 in enumerate(r.inits):
                ctx = ctxs[r.mesh]
                ctx.set_state(H0)
                ctx.rk_steps(r.nsteps)
                assert np.array_equal(ctx.get_state(), r.results[m]), (r.tag, m)
    finally:
        for c in ctxs:
            c.close()
    # one member of the largest mesh against the strict oracle
    r = reqs[0]
    err = rel_by_var(r.results[7], pyoracle.Oracle(probs[3], "strict").run(r.inits[7], r.nsteps),
                     probs[3])
    assert (err <= 1e-12).all(), err


def test_hbm_cap_evicts_idle_batches_of_other_meshes(pkg):
    probs = [_problem(pkg, 642), _problem(pkg, 2562)]
    big = probs[1].compute_size
    # static data + one 4-slot batch (H, Ka, Kb, D per slot, staging, counter)
    one_batch = big * (31 * 36 + 17 * 8) + 4 * 8 * 4 * big * 4 + 4 * 8 * big + 64
    rng = np.random.default_rng(3)
    reqs = []
    for tag in range(8):
        mid = tag % 2
        pool = pkg.serve.member_pool(pkg.mesh, probs[mid], 3, seed=tag)
        reqs.append(pkg.serve.Request(mesh=mid, inits=pool[:int(rng.integers(1, 4))],
                                      nsteps=int(rng.integers(3, 12)), tag=tag,
                                      arrival_s=0.0))
    print("case: HBM cap of one batch, two meshes alternating", flush=True)
    with pkg.serve.Server(hbm_cap_bytes=int(one_batch * 1.05), slots=4, member_block=2,
                          chunk_steps=4) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs)
    print("stats", stats, flush=True)
    assert stats["requests_done"] == len(reqs)
    assert stats["evictions"] >= 1
    assert stats["hbm_peak"] <= one_batch * 1.05
    ctxs = [pkg.runtime.HipContext(p) for p in probs]
    try:
        for r in reqs:
            for m, H0 in enumerate(r.inits):
                ctxs[r.mesh].set_state(H0)
                ctxs[r.mesh].rk_steps(r.nsteps)
                assert np.array_equal(ctxs[r.mesh].get_state(), r.results[m]), (r.tag, m)
    finally:
        for c in ctxs:
            c.close()


def test_c4_matrix_core_sums_vs_oracle_and_composition(pkg, pyoracle, c4):
    """Stencil sums on the FP64 matrix cores (member-lane tiles, fast mode):
    within 1e-12 of the strict oracle after 10 steps at N=163842 x 64, and a
    member's bits do not depend on the other members of its batch (the same
    member in a batch of different companions, fewer active slots)."""
    prob, inits = c4
    print("case: C4 MFMA sums, 10 fast steps", flush=True)
    kw = dict(members=64, member_block=32, tiled=True, prefetch=8)
    with pkg.runtime.HipContext(prob, order="hilbert") as base:
        with pkg.runtime.HipContext(prob, share=base, **kw) as ens:
            for m in range(64):
                ens.set_state(inits[m], m)
            ens.rk_steps(10)
            got = {m: ens.get_state(m) for m in (0, 31, 33, 63)}
        with pkg.runtime.HipContext(prob, share=base, **kw) as other:
            # member 33 alone in slot 5 of a batch with one other active block
            for m in range(32):
                other.set_state(inits[(m + 7) % 64], m)
            other.set_state(inits[33], 5)
            other.set_active(32)
            other.rk_steps(10)
            assert np.array_equal(other.get_state(5), got[33])
    orc = pyoracle.Oracle(prob, "strict")
    for m in (0, 31, 63):
        err = rel_by_var(got[m], orc.run(inits[m], 10), prob)
        print(f"member {m} rel err per var {err}", flush=True)
        assert (err <= 1e-12).all(), (m, err)
