#!/usr/bin/env python3
"""Benchmark of the RBF-FD shallow-water RK4 hot path on one MI355X.

Prints one JSON line.  Headline (the contract's value/roofline/cpu_baseline):
the N=40962 icosahedral mesh, 31-point stencils, FP64, one state, RK4 on the
kernel the library's `auto` rule selects (the resident whole-step kernel at
this size; --kernel names another); one "step" is one RK4 timestep = 4
eval_RHS passes over every node; inputs are generated (seeded) in the
reference's binary format, prepared by the C host (RCM, layout) and resident
in HBM before the timed region.
  value         eval_RHS throughput, GNode-RHS/s (N * 4 * steps / time)
  roofline      the stage kernel against HBM peak: algorithmic bytes per
                launch / average launch duration, timed with HIP events on
                the context's own stream around the timed region
  cpu_baseline  the reference's own main() (oracle/_ref/swe_ref_cfdl.exe,
                built from the reference sources with -O3 -fopenmp
                -DRHS_INNER_SIMD) on the same input, timed by its own
                profiling report on a bounded step sample; cpu_baseline_sfdl
                the same for the sfdl layout
Further legs (each its own key, none of them the headline):
  c3         N=655362, 200 steps, the single-state kernel the autotuner picks
  ensemble   N=163842 x 64 members, 500 steps, the ensemble configuration the
             autotuner picks; member-steps/s, ratio to the single-member path
             on the same mesh, and its HBM and FP64 roofline fractions
  serving    2000-request Poisson mix (N in {2562, 10242, 40962, 163842},
             1-32 members, 50-500 steps), HBM cap 200 GB: p50/p99 latency,
             aggregate member-steps/s, makespan; every result checked finite
             and a sample checked bitwise against solo runs
  fp32       the headline workload in the reference's FP32 model type

Flags: --gpus N (only 1 is supported: the path is built and measured on one GPU),
--steps K, --warmup W, --nodes N, --members B, --exact, --no-cpu,
--no-ensemble, --no-serve, --no-c3, --no-fp32.
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
    ap.add_argument("--kernel", default="auto",
                    choices=["auto", "quad", "lane", "quadw", "pair", "tile", "oct", "resident"])
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
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving leg (Poisson mix of forecast requests)")
    ap.add_argument("--no-c3", action="store_true", help="skip the N=655362 single-state leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 leg")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=9.0,
                    help="Poisson arrival rate, requests/s (9/s: about the rate one MI355X "
                         "completes this mix at)")
    ap.add_argument("--serve-kernel", default="tiles", choices=["tiles", "pairs"],
                    help="batch kernel of the serving leg: member-lane LDS tiles (member "
                         "block 32) or member pairs (member block 64)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    return ap.parse_args()


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() else n


def run_reference_cpu(path: str, n_nodes: int, target_s: float, layout: str = "cfdl"):
    """Times the reference build's own main() (Eval_Rhs seconds/timestep
    from its profiling report, attempts >= 2) on a bounded step sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"swe_ref_{layout}.exe")
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
            "sample": (f"reference main() (-O3 -fopenmp -DRHS_INNER_SIMD, FP64, {layout}, RCM) on the "
                       f"same N={n_nodes} input: {attempts} attempts x {steps} steps, Eval_Rhs "
                       f"s/step averaged over attempts >= 2 = {per_step:.4e} s "
                       f"({wall:.1f} s wall)"),
            "verified": "VERIFICATION" in out}


def dominant_kernel(args):
    """Name of the stage kernel the options select (the one timed)."""
    if args.fp32:
        k = args.kernel if args.kernel != "auto" else \
            ("resident" if args.nodes <= 256 * 256 else
             "lane" if args.exact or args.nodes > 262144 else "pair")
        return {"lane": "k_stage_lane32", "tile": "k_stage_tile32",
                "resident": "k_steps_resident32"}.get(k, "k_stage_split32")
    if args.members == 1:
        if args.kernel == "auto":
            # the library's auto rule (256 nodes per CU on the 256 CUs of MI355X)
            return "k_steps_resident" if args.nodes <= 256 * 256 else "k_stage_pair"
        return {"quadw": "k_stage_quadw", "pair": "k_stage_pair",
                "tile": "k_stage_ptile", "resident": "k_steps_resident"}[args.kernel]
    if args.tiled:
        return "k_stage_tspec"
    return "k_stage_mpair" if args.member_block in (0, 64) and args.members % 64 == 0 \
        else "k_stage_members"


RESIDENT_NOTE = ("k_steps_resident runs every stage of a call in one launch: 'launch' here is "
                 "one RK stage (launch time / 4 steps), 'achieved' the stage kernels' "
                 "algorithmic bytes per stage over that time; the kernel reads the static "
                 "stencil data (1252 of the 1388 B per node) once per launch, not per stage, "
                 "so frac can exceed 1; its HBM traffic is 'traffic' (rocprofv3 counts)")

FP64_PEAK_TFLOPS = 78.6
FLOP_PER_NODE_RHS = 32 * NNBR + 66   # 16 stencil sums of 31 terms + the pointwise RHS


def timed_steps(torch, ctx, k, warm=2):
    """ms per fused RK step over k steps, HIP events on the context's stream."""
    st = torch.cuda.ExternalStream(ctx.stream())
    ctx.rk_steps(warm)
    ctx.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    ctx.rk_steps(k)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / k


def stored_traffic(kernel, nodes, members):
    """HBM bytes per launch of `kernel` on this workload as counted by
    rocprofv3 (FETCH_SIZE, WRITE_SIZE) outside this process, with its source;
    (None, None) when no count was taken for exactly this kernel."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{kernel}_n{nodes}_b{members}.json")
    if not os.path.exists(tpath):
        return None, None
    with open(tpath) as fp:
        rec = json.load(fp)
    if rec.get("kernel") != kernel:
        return None, None
    return (rec.get("hbm_bytes_per_launch"),
            f"stored rocprofv3 --pmc count of {kernel} ({os.path.relpath(tpath, ROOT)}), "
            "not measured in this run")


def leg_kernel(opts):
    """Short name of the stage kernel a tuned option set runs."""
    if opts.get("members", 1) == 1:
        return {"tile": "k_stage_ptile", "resident": "k_steps_resident"}.get(opts["kernel"],
                                                                             "k_stage_pair")
    if opts["tiled"]:
        return "k_stage_tspec"
    return "k_stage_mpair" if opts["member_block"] == 64 else "k_stage_members"


def describe(opts):
    if opts.get("members", 1) == 1:
        if opts["kernel"] == "tile":
            return (f"k_stage_ptile (pair mapping over LDS tiles of {opts['tile_nodes']} nodes, "
                    f"row staging {opts['prefetch']})")
        if opts["kernel"] == "resident":
            return "k_steps_resident (whole steps in one cooperative launch)"
        return "k_stage_pair"
    if opts["tiled"]:
        return (f"k_stage_tspec (member-lane LDS tiles, {opts['tile_nodes'] or 8} nodes x "
                f"{opts['tile_rows'] or 70} rows, loader and compute waves, direct-to-LDS loads)")
    if opts["member_block"] == 64:
        return "k_stage_mpair (member pairs, 16-byte gathers)"
    return f"k_stage_members (member block {opts['member_block']})"


def c3_leg(pkg, torch, steps=200):
    """The largest mesh (N=655362), single state, 200 steps, on the kernel
    the autotuner picks for it (Hilbert-ordered device copy)."""
    import numpy as np
    path = pkg.mesh.case_path(655362, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    with pkg.runtime.HipContext(prob, order="hilbert") as base:
        opts, _ = base.tune(1, steps=3)
        with pkg.runtime.HipContext(prob, share=base, **opts) as ctx:
            ctx.set_state(prob.H_init)
            ms = timed_steps(torch, ctx, steps)
            finite = bool(np.isfinite(ctx.get_state()).all())
    n = prob.n_nodes
    launch_us = ms * 1e3 / 4
    achieved = stage_bytes(prob.padded, 1) / (launch_us * 1e-6) / 1e9
    traffic, traffic_source = stored_traffic(leg_kernel(opts), n, 1)
    return {"workload": f"N={n}, 1 member, {steps} RK4 steps, FP64, fused stages + HIP graphs",
            "kernel": describe(opts), "options": opts, "ms_per_step": ms,
            "gnode_rhs_per_s": n * 4 / (ms * 1e-3) / 1e9,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_launch": stage_bytes(prob.padded, 1),
                         "avg_launch_us": launch_us, "traffic": traffic,
                         "traffic_source": traffic_source},
            "state_finite": finite}


def ensemble_leg(pkg, torch, steps=500):
    """64-member ensemble at N=163842 against the single-state path on the
    same mesh; the ensemble configuration is the autotuner's pick."""
    import numpy as np
    prob = pkg.host.Problem(pkg.mesh.case_path(163842, seed=1), use_rcm=True, layout="cfdl")
    prob.path = pkg.mesh.case_path(163842, seed=1)
    inits = pkg.serve.member_pool(pkg.mesh, prob, 64, seed=4)
    with pkg.runtime.HipContext(prob) as single:
        single.set_state(inits[0])
        ms1 = timed_steps(torch, single, steps)
    with pkg.runtime.HipContext(prob, order="hilbert") as base:
        opts, _ = base.tune(64, steps=3)
        with pkg.runtime.HipContext(prob, share=base, **opts) as ens:
            for m in range(64):
                ens.set_state(inits[m], m)
            ms = timed_steps(torch, ens, steps)
            finite = all(bool(np.isfinite(ens.get_state(m)).all()) for m in range(64))
    n, N = prob.n_nodes, prob.padded
    launch_us = ms * 1e3 / 4
    bytes_launch = stage_bytes(N, 64)
    flops_launch = float(N) * 64 * FLOP_PER_NODE_RHS
    achieved = bytes_launch / (launch_us * 1e-6) / 1e9
    tflops = flops_launch / (launch_us * 1e-6) / 1e12
    traffic, traffic_source = stored_traffic(leg_kernel(opts), n, 64)
    return {"workload": f"N={n}, 64 members (seeded 1e-3 perturbations), {steps} RK4 steps, "
                        "Hilbert-ordered device copy, fused stages + HIP graphs",
            "kernel": describe(opts), "options": opts,
            "member_steps_per_s": 64 / (ms * 1e-3), "ms_per_step": ms,
            "gnode_rhs_per_s": n * 4 * 64 / (ms * 1e-3) / 1e9,
            "single_member_steps_per_s": 1 / (ms1 * 1e-3),
            "ratio_vs_single": (64 / ms) / (1 / ms1),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_launch": bytes_launch, "avg_launch_us": launch_us,
                         "traffic": traffic, "traffic_source": traffic_source,
                         "fp64_tflops": tflops, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                         "fp64_frac": tflops / FP64_PEAK_TFLOPS},
            "state_finite": finite}


def serve_leg(pkg, n_requests, rate, kernel="tiles", seed=2026, sample_every=250):
    """Continuous-batching mix: Poisson arrivals, N in {2562, 10242, 40962,
    163842}, 1-32 members, 50-500 steps, HBM cap 200 GB.  Every returned
    state is checked finite; every sample_every-th request is kept and
    re-run alone afterwards, which must give the same bits."""
    import numpy as np
    probs, pools = [], []
    for n in (2562, 10242, 40962, 163842):
        path = pkg.mesh.case_path(n, seed=1)
        p = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
        p.path = path
        probs.append(p)
        pools.append(pkg.serve.member_pool(pkg.mesh, p, 32, seed=n))
    reqs = pkg.serve.poisson_mix(pools, n_requests, rate, seed)
    lat, que, done, kept = [], [], [0], []

    def on_done(r):
        lat.append(r.latency_s)
        que.append(r.queue_s)
        for x in r.results:
            if not np.isfinite(x).all():
                raise RuntimeError(f"request {r.tag} returned a non-finite state")
        if r.tag % sample_every == 0:
            kept.append((r.mesh, r.nsteps, list(r.inits), list(r.results)))
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
    checked = 0
    ctxs = [pkg.runtime.HipContext(p) for p in probs]
    try:
        for mesh, nsteps, inits, results in kept:
            for H0, got in zip(inits, results):
                ctxs[mesh].set_state(H0)
                ctxs[mesh].rk_steps(nsteps)
                if not np.array_equal(ctxs[mesh].get_state(), got):
                    raise RuntimeError("a served result differs from its solo run")
                checked += 1
    finally:
        for c in ctxs:
            c.close()
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
            "batches": stats["batches"], "chunks": stats["chunks_launched"],
            "evictions": stats["evictions"],
            "results_checked": "every state finite; "
                               f"{checked} members of {len(kept)} sampled requests bitwise "
                               "equal to their solo runs"}


FP32_KERNELS = {"lane": "k_stage_lane32 (one lane per node)",
                "pair": "k_stage_split32 (2 lanes per node split the stencil slots)",
                "quad": "k_stage_split32 (4 lanes per node split the stencil slots)",
                "oct": "k_stage_split32 (8 lanes per node split the stencil slots)",
                "tile": "k_stage_tile32 (one lane per node over LDS tiles of {tn} nodes)",
                "resident": "k_steps_resident32 (whole steps in one cooperative launch, one lane "
                            "per node)"}


def fp32_leg(pkg, torch, nodes, steps, warmup):
    """The headline workload in the reference's FP32 model type (its shipped
    precision, RUN_32BIT)."""
    import numpy as np
    prob = pkg.host.Problem(pkg.mesh.case_path(nodes, seed=1), use_rcm=True, layout="cfdl",
                            fp32=True)
    with pkg.runtime.HipContext(prob, fp32=True, kernel="lane", order="hilbert") as base:
        opts, _ = base.tune(members=1, steps=20)   # lanes per node 1, 2, 4, 8 or LDS tiles
        with pkg.runtime.HipContext(prob, fp32=True, share=base, kernel=opts["kernel"],
                                    tile_nodes=opts["tile_nodes"]) as ctx:
            ctx.set_state(prob.H_init)
            ms = timed_steps(torch, ctx, steps, warm=warmup)
            finite = bool(np.isfinite(ctx.get_state()).all())
    launch_us = ms * 1e3 / 4
    achieved = stage_bytes(prob.padded, 1, fp32=True) / (launch_us * 1e-6) / 1e9
    return {"workload": f"N={prob.n_nodes}, 1 member, {steps} RK4 steps, FP32 "
                        "(the reference's RUN_32BIT type)",
            "kernel": FP32_KERNELS[opts["kernel"]].format(tn=opts["tile_nodes"]), "dtype": "f32",
            "value": prob.n_nodes * 4 / (ms * 1e-3) / 1e9, "unit": "GNode-RHS/s",
            "ms_per_step": ms,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_launch": stage_bytes(prob.padded, 1, fp32=True),
                         "avg_launch_us": launch_us,
                         "note": RESIDENT_NOTE if opts["kernel"] == "resident"
                         else "one launch = one fused RK stage"},
            "state_finite": finite}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this path runs on one GPU "
              "(ensembles and serving would shard by member across GPUs, one process "
              "per GPU, with no exchange)",
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
    # HBM bytes per launch are counted by rocprofv3 (FETCH_SIZE, WRITE_SIZE)
    # outside this process; the stored count is reported with its source and
    # only when it was taken for the kernel this run selects
    traffic, traffic_source = (None, None) if args.fp32 else \
        stored_traffic(dominant_kernel(args), args.nodes, args.members)
    node_rhs = float(prob.n_nodes) * 4 * args.steps * args.members
    value = node_rhs / (gpu_ms * 1e-3) / 1e9
    ctx.close()
    cpu = cpu_sfdl = None
    if not args.no_cpu:
        cpu = run_reference_cpu(path, prob.n_nodes, args.cpu_seconds, "cfdl")
        cpu_sfdl = run_reference_cpu(path, prob.n_nodes, args.cpu_seconds, "sfdl")
        print("bench: CPU baselines done", file=sys.stderr, flush=True)
    fp32 = None
    if not args.no_fp32 and not args.fp32 and args.members == 1:
        fp32 = fp32_leg(pkg, torch, args.nodes, args.steps, args.warmup)
    c3 = None
    if not args.no_c3:
        c3 = c3_leg(pkg, torch)
        print("bench: C3 leg done", file=sys.stderr, flush=True)
    ensemble = None
    if not args.no_ensemble:
        ensemble = ensemble_leg(pkg, torch)
        print("bench: ensemble leg done", file=sys.stderr, flush=True)
    serving = None
    if not args.no_serve:
        serving = serve_leg(pkg, args.serve_requests, args.serve_rate, args.serve_kernel)
    # the best published figure for this workload: the reference's own run log
    # (Eval_Rhs 7.371213e-04 s per step, N = 40962, sfdl layout, OpenACC on a
    # V100, 2 MPI ranks), which times the RHS alone (value also includes the
    # fused RK4 updates)
    vs_baseline, vs_basis = None, None
    if prob.n_nodes == 40962 and args.members == 1 and not args.fp32:
        published = 4 * 40962 / 7.371213e-04 / 1e9
        vs_baseline = value / published
        vs_basis = (f"{published:.3f} GNode-RHS/s: Eval_Rhs of the reference's run log "
                    "run/psg/output/output_sfdl.txt:41 (N=40962, sfdl, OpenACC on a V100, "
                    "2 MPI ranks, precision not recorded); it times the RHS alone")
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
        "vs_baseline": vs_baseline,
        "vs_baseline_basis": vs_basis,
        "dtype": "f32" if args.fp32 else "f64",
        "data": "synthetic (seeded icosahedral mesh + Gaussian RBF-FD weights, TC5-like state)",
        "config": {"workload": f"RBF-FD SWE RK4, N={prob.n_nodes} nodes, stencil 31, "
                               f"{args.members} member(s), {dominant_kernel(args)}"
                               + (" (the 40962-node, 1000-step single-member case timed "
                                  "against the reference CPU path)"
                                  if prob.n_nodes == 40962 and args.members == 1 else ""),
                   "nodes": prob.n_nodes, "members": args.members, "layout": args.layout,
                   "member_block": args.member_block, "kernel": args.kernel,
                   "prefetch": args.prefetch, "order": args.order, "tiled": args.tiled,
                   "slot_block": args.slot_block,
                   "exact": bool(args.exact), "dt": prob.dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": dominant_kernel(args),
                     "bytes_per_launch": stage_bytes(N, args.members, args.fp32),
                     "avg_launch_us": launch_s * 1e6,
                     "note": RESIDENT_NOTE if dominant_kernel(args) == "k_steps_resident"
                     else "one launch = one fused RK stage"},
        "cpu_baseline": cpu,
        "cpu_baseline_sfdl": cpu_sfdl,
        "c3": c3,
        "ensemble": ensemble,
        "serving": serving,
        "fp32": fp32,
        "member_steps_per_s": args.members * args.steps / (gpu_ms * 1e-3),
        "wall_s_timed_region": wall,
        "state_finite": finite,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
