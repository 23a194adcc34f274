"""Continuous batching: every request's result is the same alone and in the
mix (bitwise: members never interact and chunking does not change any
member's arithmetic), under joins/leaves mid-flight and an HBM cap."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(pkg, probs, nreq, seed, max_members=5):
    rng = np.random.default_rng(seed)
    reqs = []
    for t in range(nreq):
        mid = int(rng.integers(len(probs)))
        prob = probs[mid]
        case = pkg.mesh.read_bin(prob.path)
        E = pkg.mesh.perturb_members(case, int(rng.integers(1, max_members + 1)),
                                     seed=1000 + t)
        inits = []
        for Hn in E:
            Hn = Hn[prob.mapping]
            full = np.concatenate([Hn, np.repeat(Hn[-1:], prob.padded - prob.n_nodes, 0)])
            inits.append(prob.from_node_major(full))
        reqs.append(pkg.serve.Request(mesh=mid, inits=inits, nsteps=int(rng.integers(3, 40)),
                                      tag=t, arrival_s=float(rng.uniform(0.0, 0.05))))
    return reqs


def _probs(pkg):
    out = []
    for n in (642, 2562):
        path = pkg.mesh.case_path(n, seed=1)
        p = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
        p.path = path
        out.append(p)
    return out


@pytest.mark.parametrize("exact", [False, True])
def test_mix_equals_solo(pkg, exact):
    probs = _probs(pkg)
    reqs = _mix(pkg, probs, 14, seed=5)
    print(f"case: serving mix exact={exact} requests={len(reqs)}", flush=True)
    with pkg.serve.Server(slots=8, member_block=2, chunk_steps=4, max_batches=2,
                          exact=exact) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs)
    assert stats["requests_done"] == len(reqs)
    assert stats["member_steps_done"] >= sum(len(r.inits) * r.nsteps for r in reqs)
    ctxs = [pkg.runtime.HipContext(p, exact=exact) for p in probs]
    try:
        for r in reqs:
            for m, H0 in enumerate(r.inits):
                ctx = ctxs[r.mesh]
                ctx.set_state(H0)
                ctx.rk_steps(r.nsteps)
                assert np.array_equal(ctx.get_state(), r.results[m]), (r.tag, m)
            assert r.latency_s >= r.queue_s >= 0.0
    finally:
        for c in ctxs:
            c.close()


def test_hbm_cap_queues_instead_of_failing(pkg):
    probs = _probs(pkg)[:1]
    reqs = _mix(pkg, probs, 6, seed=9, max_members=4)
    prob = probs[0]
    n = prob.compute_size
    one_batch = n * (31 * 36 + 17 * 8) + 4 * 8 * 4 * n * 4 + 4 * 8 * n + 64
    with pkg.serve.Server(hbm_cap_bytes=int(one_batch * 1.05), slots=4, member_block=2,
                          chunk_steps=3) as srv:
        srv.add_mesh(prob)
        stats = srv.run(reqs)
    print("stats", stats, flush=True)
    assert stats["requests_done"] == len(reqs)
    assert stats["hbm_peak"] <= one_batch * 1.05
    assert stats["batches"] == 1


def test_tiled_batches_equal_solo(pkg):
    """Batches on the member-lane LDS-tile kernel (Hilbert-ordered copies,
    member block 32) give every request its solo fast-mode bits."""
    probs = _probs(pkg)
    reqs = _mix(pkg, probs, 10, seed=13, max_members=6)
    print(f"case: serving mix tiled requests={len(reqs)}", flush=True)
    with pkg.serve.Server(slots=32, member_block=32, chunk_steps=4, max_batches=2,
                          tiled=True) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs)
    assert stats["requests_done"] == len(reqs)
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


def test_result_files_hold_each_request(pkg, tmp_path):
    """Per-request result files (swe_srv_submit_ex): the history format of
    swe_run, one record per member in the file node order, equal to the
    results returned in memory."""
    probs = _probs(pkg)
    reqs = _mix(pkg, probs, 6, seed=21, max_members=3)
    for r in reqs:
        r.result_path = str(tmp_path / f"req{r.tag}.bin")
    print("case: serving result files", flush=True)
    with pkg.serve.Server(slots=4, member_block=2, chunk_steps=5) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs)
    assert stats["requests_done"] == len(reqs)
    for r in reqs:
        prob = probs[r.mesh]
        steps, times, states = pkg.mesh.read_history(r.result_path)
        assert steps == [r.nsteps] * len(r.inits)
        for m, H in enumerate(r.results):
            Hn = prob.to_node_major(H)[: prob.n_nodes]
            nat = np.empty_like(Hn)
            nat[prob.mapping] = Hn
            assert np.array_equal(states[m], nat), (r.tag, m)
