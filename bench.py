#!/usr/bin/env python3
"""Benchmark of the RBF-FD shallow-water RK4 hot path on one MI355X.

Workload (default): the N=40962 icosahedral mesh, 31-point stencils, FP64,
one state, RK4 with the fused graph-captured stage kernels.  One "step" is
one RK4 timestep = 4 eval_RHS passes over every node.  Inputs are generated
(seeded) in the reference's binary format, prepared by the C host (RCM,
layout) and resident in HBM before the timed region.

Prints one JSON line:
  value      eval_RHS throughput, GNode-RHS/s (N * 4 * steps / time)
  roofline   the dominant kernel (the fused stage kernel) against HBM peak:
             algorithmic bytes per launch / average launch duration, the
             duration taken from HIP events recorded on the context's own
             stream around the timed region
  cpu_baseline  the reference's own CPU path (oracle/_ref/swe_ref_cfdl.exe,
             built from the reference sources with -O3 -fopenmp
             -DRHS_INNER_SIMD) on the same input, timed by its own
             profiling report, on a bounded sample of steps

Flags: --gpus N (only 1 is supported: every GPU box here has one GPU),
--steps K, --warmup W, --nodes N, --members B, --exact, --no-cpu.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NNBR = 31
# Algorithmic HBM bytes per node per fused stage launch, single state:
# static data 31*(4 B id + 4*8 B weights) + 17*8 B geometry = 1252 B, the
# node's own state read once (32 B) and the stage epilogue, averaged over the
# four stages (D,K written; D,H read + D,K written x2; D,H read + H written:
# (64 + 128 + 128 + 96) / 4 = 104 B).  Per member beyond the first: 32 + 104.
STATIC_B = NNBR * (4 + 4 * 8) + 17 * 8
STATE_B = 32 + 104
# The same for the FP32 model type: 31*(4 + 4*4) + 17*4 = 688 B static,
# 16 B own state, (32 + 64 + 64 + 48) / 4 = 52 B stage epilogue.
STATIC_B32 = NNBR * (4 + 4 * 4) + 17 * 4
STATE_B32 = 16 + 52


def stage_bytes(n_nodes: int, members: int, fp32: bool = False) -> float:
    if fp32:
        return float(n_nodes) * (STATIC_B32 + STATE_B32 * members)
    return float(n_nodes) * (STATIC_B + STATE_B * members)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--nodes", type=int, default=40962)
    ap.add_argument("--members", type=int, default=1)
    ap.add_argument("--member-block", type=int, default=0)
    ap.add_argument("--kernel", default="auto", choices=["auto", "quad", "lane", "quadw", "pair", "tile"])
    ap.add_argument("--prefetch", type=int, default=0)
    ap.add_argument("--order", default="hilbert", choices=["host", "morton", "hilbert"],
                    help="node order of the device copy (results are identical in any order; "
                         "Hilbert is the fastest measured at N=40962)")
    ap.add_argument("--tiled", action="store_true")
    ap.add_argument("--slot-block", type=int, default=0, choices=[0, 16])
    ap.add_argument("--tile-nodes", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--layout", default="cfdl")
    ap.add_argument("--exact", action="store_true")
    ap.add_argument("--fp32", action="store_true",
                    help="the reference's FP32 model type (SWE_HIP_FP32 context) for the main leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 4 device copies of 1 GiB (HBM counter calibration) and exit")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble leg (N=163842, 64 members, 500 steps)")
    ap.add_argument("--serve", action="store_true",
                    help="add the serving leg (Poisson mix of forecast requests)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=25.0)
    ap.add_argument("--serve-kernel", default="tiles", choices=["tiles", "pairs"],
                    help="batch kernel of the serving leg: member-lane LDS tiles (member "
                         "block 32) or member pairs (member block 64)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() else n


def run_reference_cpu(path: str, n_nodes: int, target_s: float):
    """Times the reference build's own main() (Eval_Rhs seconds/timestep
    from its profiling report, attempts >= 2) on a bounded step sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "swe_ref_cfdl.exe")
    if not os.path.exists(exe):
        return None
    cores = cpu_cores()
    env = dict(os.environ, SWE_INPUT_FILE=path, SWE_USE_NETCDF="0", SWE_NUM_COMPUTE_UNITS="1",
               SWE_USE_RCM="1", SWE_USE_OCL="0", OMP_NUM_THREADS=str(cores),
               OMP_SCHEDULE="static", OMP_PROC_BIND="close")

    def run(steps, attempts):
        e = dict(env, SWE_NSTEPS=str(steps), SWE_NATTEMPTS=str(attempts))
        t0 = time.time()
        out = subprocess.run([exe], env=e, capture_output=True, text=True, timeout=600).stdout
        m = re.search(r"Eval_Rhs\s+\(seconds/timestep\) -> \s*Average: \s*([0-9.eE+-]+)", out)
        return (float(m.group(1)) if m else None), time.time() - t0, out
    per_step, wall, _ = run(10, 3)
    if not per_step:
        return None
    attempts = 4
    steps = int(max(10, min(1000, target_s / (attempts * max(per_step, 1e-9)))))
    per_step, wall, out = run(steps, attempts)
    if not per_step:
        return None
    return {"value": n_nodes * 4 / per_step / 1e9, "unit": "GNode-RHS/s", "cores": cores,
            "kind": "reference",
            "sample": (f"reference main() (-O3 -fopenmp -DRHS_INNER_SIMD, FP64, cfdl, RCM) on the "
                       f"same N={n_nodes} input: {attempts} attempts x {steps} steps, Eval_Rhs "
                       f"s/step averaged over attempts >= 2 = {per_step:.4e} s "
                       f"({wall:.1f} s wall)"),
            "verified": "VERIFICATION" in out}


def dominant_kernel(args):
    """Name of the stage kernel the options select (the one timed)."""
    if args.members == 1:
        if args.kernel == "auto":
            return "k_stage_pair"
        return {"quadw": "k_stage_quadw", "quad": "k_stage_quad",
                "lane": "k_stage_lane", "pair": "k_stage_pair",
                "tile": "k_stage_ptile"}[args.kernel]
    if args.tiled:
        return "k_stage_tspec" if args.member_block == 32 else "k_stage_tiled"
    return "k_stage_mpair" if args.member_block in (0, 64) and args.members % 64 == 0 \
        else "k_stage_members"


def ensemble_leg(pkg, torch, steps=500):
    """64-member ensemble at N=163842 against the single-state path on the
    same mesh.  The ensemble kernel is picked by timing the candidates (member
    pairs at two pipeline depths, LDS tiles) on this mesh."""
    import numpy as np
    prob = pkg.host.Problem(pkg.mesh.case_path(163842, seed=1), use_rcm=True, layout="cfdl")
    prob.path = pkg.mesh.case_path(163842, seed=1)
    inits = pkg.serve.member_pool(pkg.mesh, prob, 64, seed=4)

    def timed(ctx, k):
        st = torch.cuda.ExternalStream(ctx.stream())
        ctx.rk_steps(2)
        ctx.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        ctx.rk_steps(k)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / k
    single = pkg.runtime.HipContext(prob)
    single.set_state(inits[0])
    ms1 = timed(single, steps)
    base = pkg.runtime.HipContext(prob, order="hilbert")
    candidates = [("member pairs (16-byte gathers), pipeline depth 3",
                   dict(member_block=64, prefetch=0)),
                  ("LDS tiles with member lanes, member block 32, loader and compute waves "
                   "over two LDS images, persistent",
                   dict(member_block=32, tiled=True))]
    best = None
    for name, kw in candidates:
        ctx = pkg.runtime.HipContext(prob, members=64, share=base, **kw)
        for m in range(64):
            ctx.set_state(inits[m], m)
        ms = timed(ctx, 20)
        if best is None or ms < best[2]:
            best = (name, kw, ms)
        ctx.close()
    ens = pkg.runtime.HipContext(prob, members=64, share=base, **best[1])
    for m in range(64):
        ens.set_state(inits[m], m)
    ms = timed(ens, steps)
    finite = bool(np.isfinite(ens.get_state(63)).all())
    ens.close()
    base.close()
    single.close()
    n = prob.n_nodes
    return {"workload": f"N={n}, 64 members (seeded 1e-3 perturbations), {steps} RK4 steps, "
                        f"{best[0]}, Hilbert-ordered device copy",
            "member_steps_per_s": 64 / (ms * 1e-3), "ms_per_step": ms,
            "gnode_rhs_per_s": n * 4 * 64 / (ms * 1e-3) / 1e9,
            "fp64_tflops": n * 64 * 4 * 1058 / (ms * 1e-3) / 1e12,
            "single_member_steps_per_s": 1 / (ms1 * 1e-3),
            "ratio_vs_single": (64 / ms) / (1 / ms1), "state_finite": finite}


def serve_leg(pkg, n_requests, rate, kernel="tiles", seed=2026):
    """Continuous-batching mix: Poisson arrivals, N in {2562, 10242, 40962,
    163842}, 1-32 members, 50-500 steps, HBM cap 200 GB."""
    import numpy as np
    probs, pools = [], []
    for n in (2562, 10242, 40962, 163842):
        path = pkg.mesh.case_path(n, seed=1)
        p = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
        p.path = path
        probs.append(p)
        pools.append(pkg.serve.member_pool(pkg.mesh, p, 32, seed=n))
    reqs = pkg.serve.poisson_mix(pools, n_requests, rate, seed)
    lat, que, done = [], [], [0]

    def on_done(r):
        lat.append(r.latency_s)
        que.append(r.queue_s)
        if not all(np.isfinite(x[:8]).all() for x in r.results):
            raise RuntimeError(f"request {r.tag} returned non-finite state")
        r.results = None
        r.inits = None
        done[0] += 1
        if done[0] % 200 == 0:
            print(f"bench: serve {done[0]}/{len(reqs)} done", file=sys.stderr, flush=True)
    work = sum(len(r.inits) * r.nsteps for r in reqs)
    node_rhs = sum(len(r.inits) * r.nsteps * 4 * probs[r.mesh].n_nodes for r in reqs)
    tiles = kernel == "tiles"
    with pkg.serve.Server(hbm_cap_bytes=200 << 30, slots=64, member_block=32 if tiles else 64,
                          chunk_steps=16, tiled=tiles) as srv:
        for p in probs:
            srv.add_mesh(p)
        stats = srv.run(reqs, on_done=on_done)
    lat = np.array(lat)
    return {"workload": f"{n_requests} requests, Poisson {rate}/s, N in {{2562,10242,40962,163842}}, "
                        "1-32 members, 50-500 steps, HBM cap 200 GB, batches of 64 slots on "
                        + ("member-lane LDS tiles" if tiles else "member pairs"),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "p50_queue_s": float(np.percentile(np.array(que), 50)),
            "member_steps_per_s": work / stats["makespan_s"],
            "gnode_rhs_per_s": node_rhs / stats["makespan_s"] / 1e9,
            "makespan_s": stats["makespan_s"], "hbm_peak_gb": stats["hbm_peak"] / 2**30,
            "batches": stats["batches"], "chunks": stats["chunks_launched"]}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this path runs on one GPU "
              "(each GPU box has one GPU; ensembles shard by member with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import swe_pkg
    pkg = swe_pkg.load()
    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        # a plain copy of known size: each dispatch reads and writes 1 GiB
        a = torch.ones(2**27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate": "copy", "dispatches": 4, "bytes_read_per_dispatch": 2**30,
                          "bytes_written_per_dispatch": 2**30}), flush=True)
        return

    t0 = time.time()
    path = pkg.mesh.case_path(args.nodes, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout=args.layout, simd=4, fp32=args.fp32)
    print(f"bench: input ready N={prob.n_nodes} ({time.time() - t0:.1f}s)", file=sys.stderr,
          flush=True)
    ctx = pkg.runtime.HipContext(prob, members=args.members, exact=args.exact,
                                 member_block=args.member_block, kernel=args.kernel,
                                 prefetch=args.prefetch, order=args.order, tiled=args.tiled,
                                 slot_block=args.slot_block, tile_nodes=args.tile_nodes,
                                 tile_rows=args.tile_rows, fp32=args.fp32)
    for m in range(args.members):
        ctx.set_state(prob.H_init, m)
    stream = torch.cuda.ExternalStream(ctx.stream())

    ctx.rk_steps(args.warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    w0 = time.perf_counter()
    ctx.rk_steps(args.steps)
    ev1.record(stream)
    ev1.synchronize()
    wall = time.perf_counter() - w0
    torch.cuda.synchronize()
    gpu_ms = ev0.elapsed_time(ev1)
    H = ctx.get_state(0)
    finite = bool(np.isfinite(H).all())

    N = prob.padded
    launches = 4 * args.steps
    launch_s = gpu_ms * 1e-3 / launches
    achieved = stage_bytes(N, args.members, args.fp32) / launch_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_n{args.nodes}_b{args.members}.json")
    if os.path.exists(tpath):
        with open(tpath) as fp:
            traffic = json.load(fp).get("hbm_bytes_per_launch")
    node_rhs = float(prob.n_nodes) * 4 * args.steps * args.members
    value = node_rhs / (gpu_ms * 1e-3) / 1e9
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cpu = run_reference_cpu(path, prob.n_nodes, args.cpu_seconds)
    ensemble = None
    if not args.no_ensemble:
        ensemble = ensemble_leg(pkg, torch)
    serving = None
    if args.serve:
        serving = serve_leg(pkg, args.serve_requests, args.serve_rate, args.serve_kernel)
    out = {
        "metric": "eval_RHS GNodes/s + HBM GB/s; ensemble steps/s and p50/p99 latency",
        "value": value,
        "unit": "GNode-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded icosahedral mesh + Gaussian RBF-FD weights, TC5-like state)",
        "config": {"workload": f"RBF-FD SWE RK4, N={prob.n_nodes} nodes, stencil 31, "
                               f"{args.members} member(s), fused stage kernels + HIP graphs",
                   "nodes": prob.n_nodes, "members": args.members, "layout": args.layout,
                   "member_block": args.member_block, "kernel": args.kernel,
                   "prefetch": args.prefetch, "order": args.order, "tiled": args.tiled,
                   "slot_block": args.slot_block,
                   "exact": bool(args.exact), "dt": prob.dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dominant_kernel(args),
                     "bytes_per_launch": stage_bytes(N, args.members, args.fp32),
                     "avg_launch_us": launch_s * 1e6},
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "serving": serving,
        "member_steps_per_s": args.members * args.steps / (gpu_ms * 1e-3),
        "wall_s_timed_region": wall,
        "state_finite": finite,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
