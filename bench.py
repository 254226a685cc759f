"""Benchmark: chain-steps/sec of the multi-chain annealer on one MI355X.

Headline workload (one graph per launch): 2D toroidal grid 100 x 200
(n = 20000, m = 40000, degree 4), iid uniform +/-1 weights, 1024 chains,
linear beta 0.1 -> 3.0, group-count ladder P = 128, 64, ..., 1 with equal
sweeps per phase.

A "step" is one block of `--sweeps-per-step` (default 2000) sweeps of every
chain.  W warmup steps run first (untimed, at P = 128); then exactly K steps
run as ONE persistent kernel launch timed with HIP events on the library's
stream (inputs and chain state already resident in HBM), with the P ladder
spread evenly over the K timed blocks.  The launch uses the threads-per-chain
heuristic (256 at 4 chains per CU), which is the tile the live autotuner
picks for every phase of this workload.

The same line carries the batched and served modes against the plain
one-graph-per-launch mode (`batched_served`): a 256-graph mixed batch in one
hipGraph-captured launch, and a Poisson stream of solve requests through the
continuous-batching server, with latencies observed for both the server and
a plain first-come-first-served loop.

Prints one JSON line; diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# VALU issue peak: one wave64 instruction per CU per clock at the 2.4 GHz peak
# engine clock (64 int32 lanes per CU per clock).
VALU_PEAK_GINSTR_PER_CU = 2.4
COUNTERS = ROOT / "profiles" / "annealSlice_torus100x200_1024chains_counters.json"
LADDER = [128, 64, 32, 16, 8, 4, 2, 1]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_chain(n: int, mean_deg_winner: float, p_per_sweep) -> float:
    """Sum over sweeps of 5n + 6 P_w + 16 * sum(deg(winners)), with every group
    non-empty (P <= n) and sum(deg) = P * deg for a regular graph."""
    p = np.asarray(p_per_sweep, dtype=np.float64)
    return float(np.sum(5.0 * n + 6.0 * p + 16.0 * mean_deg_winner * p))


def host_cpu_share() -> tuple[int, int, str]:
    """(threads to use, CPUs visible to this process, CPU model).  The share is
    the CPUs this process may run on, capped by a cgroup CPU quota and by
    OMP_NUM_THREADS when either is set (a GPU host gives each GPU's jobs a
    fixed share of its cores)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = visible
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            share = min(share, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, share), visible, model


def cpu_baseline(csr, chains: int, sweeps: int, threads: int):
    import rfmf_amd as R
    from oracle import oracle as O
    iters = [sweeps // 8] * 8
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), chains)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in range(1, chains + 1)])
    seeds = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in range(1, chains + 1)]
    _, _, _, _, sec = O.anneal_many(csr.offsets, csr.edge_to, csr.edge_weight, init, seeds, beta.run_end,
                                    beta.run_val, beta.chain_run_off, LADDER, iters, threads=threads)
    return chains * sum(iters) / sec, sec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (blocks of --sweeps-per-step sweeps)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps before the timed launch")
    ap.add_argument("--sweeps-per-step", type=int, default=2000,
                    help="sweeps per chain in one step (a multiple of 8, so every P of the ladder gets "
                         "an equal share of the timed sweeps)")
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--rows", type=int, default=100)
    ap.add_argument("--cols", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-small-config", action="store_true",
                    help="leave the ER n=800 GPU/CPU comparison out of the cpu_baseline leg")
    ap.add_argument("--no-batched-served", action="store_true",
                    help="headline only: skip the batched and served legs")
    # CPU sample: 4 chains per host thread (at least 64) x 64000 sweeps, about
    # 10-20 s on 16 host threads
    ap.add_argument("--cpu-sweeps", type=int, default=64000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (counter calibration) and exit")
    ap.add_argument("--config", default="torus",
                    choices=["torus", "er800", "sparse", "batch", "serve", "gset", "large"],
                    help="torus: the headline plain run (default); the others are the other workloads")
    ap.add_argument("--threads", type=int, default=0, help="threads per chain (0: autotune)")
    ap.add_argument("--qps", type=float, default=120.0, help="serve: Poisson arrival rate")
    ap.add_argument("--requests", type=int, default=96, help="serve: number of requests")
    ap.add_argument("--slice", type=int, default=500, help="serve: sweeps per serving slice")
    ap.add_argument("--no-plain", action="store_true", help="serve: served leg only (no plain FIFO comparison)")
    ap.add_argument("--serve-sweeps", default="10000,100000", metavar="LO,HI",
                    help="serve: sweeps per request, log-uniform in [LO, HI]")
    ap.add_argument("--plain-stride", type=int, default=1, metavar="S",
                    help="batch: run the plain one-graph-per-launch leg on every S-th graph only (its rate "
                         "is per graph and sweep, so a stratified sample measures it at full length)")
    ap.add_argument("--graphs", type=int, default=256, help="batch: number of graphs")
    ap.add_argument("--batch-sweeps", type=int, default=20000, help="batch: sweeps per chain")
    ap.add_argument("--gset-sweeps", type=int, default=100000, help="gset: sweeps per chain")
    ap.add_argument("--instances", default="", help="gset: comma-separated instance names (default: all 12)")
    ap.add_argument("--results-csv", default="", help="gset: write the results table here")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the annealer shards by chains and graphs with no cross-GPU "
              f"exchange, and this library runs on one MI355X only (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    if args.copy_probe:
        import rfmf_amd as R
        with R.Context(0) as ctx:
            ms = ctx.copy_probe(args.copy_probe, iters=5)
        log(f"[bench] copy probe {args.copy_probe} B: {ms:.3f} ms/launch, "
            f"{2 * args.copy_probe / ms / 1e6:.1f} GB/s read+write")
        return 0

    if args.config != "torus":
        return {"er800": bench_er800, "sparse": bench_sparse, "batch": bench_batch,
                "serve": bench_serve, "gset": bench_gset, "large": bench_large}[args.config](args)

    if args.steps < 1 or args.warmup < 0:
        print("--steps must be at least 1 and --warmup at least 0", file=sys.stderr)
        return 2
    if args.sweeps_per_step < 8 or args.sweeps_per_step % 8:
        print("--sweeps-per-step must be a positive multiple of 8", file=sys.stderr)
        return 2

    import rfmf_amd as R
    from rfmf_amd import workloads as W

    sps = args.sweeps_per_step
    K, Wm = args.steps, args.warmup
    timed_sweeps, warm_sweeps = K * sps, Wm * sps
    csr = W.torus(args.rows, args.cols, seed=0)
    n, C = csr.n, args.chains
    B = R.max_weighted_degree(csr)
    log(f"[bench] torus {args.rows}x{args.cols} n={n} m2={csr.m2} B={B} digest={W.csr_digest(csr)} chains={C} "
        f"step={sps} sweeps, warmup={Wm} steps, timed={K} steps ({timed_sweeps} sweeps)")

    phases = list(LADDER)
    iters = [timed_sweeps // 8] * 8
    iters[0] += warm_sweeps
    total = sum(iters)
    beta = R.build_linear_beta_table(0.1, 3.0, total, C)
    init = np.stack([R.generate_random_spins(s, n) for s in range(1, C + 1)])
    seeds = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in range(1, C + 1)]

    with R.Context(0) as ctx:
        info = ctx.device_info()
        log(f"[bench] device {info}")
        gid = ctx.upload_graph(csr)
        job = R.Job(csr, init, seeds, beta, phases, iters)
        t0 = time.perf_counter()
        warm_ms, timed_ms = run_two_slices(ctx, gid, job, warm_sweeps, timed_sweeps)
        wall = time.perf_counter() - t0
        geo = ctx.last_launch()
        log(f"[bench] warmup {warm_ms:.1f} ms, timed {timed_ms:.1f} ms ({timed_ms * 1e3 / timed_sweeps:.2f} us "
            f"per sweep of {C} chains), host wall {wall:.2f} s, {geo}")

    best = job.best_vals
    for c in range(C):
        assert R.cut_value_csr(csr, job.best_spins[c]) == best[c], "incremental best != recomputed cut"
    chain_steps = C * timed_sweeps
    value = chain_steps / (timed_ms / 1e3)
    p_timed = np.repeat(LADDER, timed_sweeps // 8)
    bytes_per_chain = algorithmic_bytes_per_chain(n, 4.0, p_timed)
    algorithmic_gbs = C * bytes_per_chain / (timed_ms / 1e3) / 1e9
    roofline = roofline_of(args, info, value, timed_ms, timed_sweeps, warm_sweeps, C, algorithmic_gbs,
                           bytes_per_chain / timed_sweeps)

    cpu = None
    if not args.no_cpu_baseline:
        threads, visible, model = host_cpu_share()
        chains = max(64, 4 * threads)
        sweeps = args.cpu_sweeps // 8 * 8
        rate, sec = cpu_baseline(csr, chains, sweeps, threads)
        cpu = {"value": rate, "unit": "chain-steps/s", "cores": threads, "kind": "port",
               "cpus_visible": visible, "cpu_model": model,
               "sample": f"oracle (serial C restatement, one chain per thread), torus 100x200, {chains} chains x "
                         f"{sweeps} sweeps, same P ladder and linear beta, {sec:.2f} s wall on {threads} threads "
                         f"(the host's CPU share for this process; {visible} CPUs visible)"}
        log(f"[bench] cpu baseline {rate:.3e} chain-steps/s on {threads} threads of {visible} visible "
            f"({model}, {sec:.2f} s)")
        if not args.no_small_config:
            # The small configuration in full (10^6 sweeps of 64 chains) on
            # both paths, partitions compared per seed.
            cpu["small_config"] = er800_leg(1_000_000)

    modes = None
    if not args.no_batched_served:
        modes = batched_served_legs(args)

    out = {
        "metric": "chain-steps/sec on 1 MI355X at equal best cut; p50/p99 solve latency",
        "value": value,
        "unit": "chain-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": timed_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded torus graph, seeded initial spins)",
        "config": {"workload": "torus100x200_pm1_1024chains_plain", "n": n, "m": csr.m2 // 2, "chains": C,
                   "sweeps_per_step": sps, "timed_sweeps": timed_sweeps, "warmup_sweeps": warm_sweeps,
                   "schedule": "linear beta 0.1->3.0 over warmup + timed sweeps, P ladder 128..1 equal over "
                               "the timed sweeps (warmup at P=128)", "parallelism": "chains",
                   "threads_per_chain": geo["threads_per_chain"], "wg_per_cu": geo["wg_per_cu"],
                   "best_cut_median": float(np.median(best)), "best_cut_max": int(best.max())},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched_served": modes,
    }
    print(json.dumps(out), flush=True)
    return 0


def roofline_of(args, info, value, timed_ms, timed_sweeps, warm_sweeps, chains, algorithmic_gbs,
                algo_bytes_per_chain_step):
    """VALU-issue roofline of the timed launch, from the committed rocprofv3
    counter run of the same command: wave64 VALU instructions per chain-step
    x chain-steps/s against one instruction per CU per clock.  HBM traffic
    (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy kernel) applies to
    a launch of the profiled size."""
    out = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINSTR_PER_CU * info["cu"],
           "unit": "G wave64-instr/s", "frac": None, "traffic": None,
           "algorithmic_gbs": algorithmic_gbs, "algorithmic_bytes_per_chain_step": algo_bytes_per_chain_step,
           "note": "VALU-issue bound: the chain state is LDS-resident, so HBM carries only the state in and "
                   "out, CSR rows and beta runs (`traffic`); algorithmic_gbs counts the reference layout's "
                   "bytes (5n + 6P + 16 sum deg per chain-step), which here are LDS traffic"}
    if not COUNTERS.exists() or (args.rows, args.cols) != (100, 200):
        return out
    t = json.loads(COUNTERS.read_text())
    ipc = t["valu_instr_per_chain_step"]
    got = ipc * value / 1e9
    out.update({"achieved": got, "frac": got / out["peak"], "instr_per_chain_step": ipc,
                "profile": t.get("source"),
                "profiled_shape": {k: t.get(k) for k in ("chains", "timed_sweeps", "warmup_sweeps")}})
    if (t.get("chains"), t.get("timed_sweeps"), t.get("warmup_sweeps")) == (chains, timed_sweeps, warm_sweeps):
        traffic = t["bytes_per_launch"]
        gbs = traffic / (timed_ms / 1e3) / 1e9
        out.update({"traffic": traffic, "traffic_per_sweep": traffic / timed_sweeps, "traffic_gbs": gbs,
                    "traffic_hbm_frac": gbs / HBM_PEAK_GBS})
    return out


def batched_served_legs(args) -> dict:
    """Batched and served modes against the plain one-graph-per-launch mode."""
    import rfmf_amd as R
    out = {}
    b = run_batch(R, args.graphs, args.batch_sweeps, args.threads)
    out["batched"] = b
    s = run_serve(R, args.requests, args.qps, args.slice)
    out["served"] = s
    out["target"] = ">= 5x the plain one-graph-per-launch chain-steps/s at equal best-cut medians"
    out["batched_vs_plain"] = b["speedup"]
    out["served_vs_plain"] = s["speedup"]
    return out


def _job(R, csr, seeds, iters, beta_range=(0.1, 3.0)):
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    beta = R.build_linear_beta_table(beta_range[0], beta_range[1], sum(iters), len(seeds))
    return R.Job(csr, init, solver, beta, LADDER, iters)


def _emit(obj):
    print(json.dumps(obj), flush=True)


def er800_leg(sweeps: int, threads_per_chain: int = 0, check: bool = True) -> dict:
    """The small configuration: ER n=800 p=0.06 unit weights, 64 chains,
    `sweeps` sweeps (linear beta 0.1->3.0, P ladder 128..1) in one plain GPU
    launch with autotuned geometry; with `check`, the CPU path (the oracle,
    one chain per thread on the host's CPU share) runs every chain too, timed,
    and the two are compared bit for bit per seed."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    csr = W.erdos_renyi(800, 0.06, seed=0)
    C = 64
    iters = [sweeps // 8] * 8
    job = _job(R, csr, range(1, C + 1), iters)
    log(f"[er800] n={csr.n} m={csr.m2 // 2} B={R.max_weighted_degree(csr)} chains={C} sweeps={sum(iters)}")
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ms = ctx.batch_run([(gid, job)], autotune=threads_per_chain == 0, threads=threads_per_chain)
        geo = ctx.last_launch()
    gpu_rate = C * sum(iters) / (ms / 1e3)
    out = {"config": "er800_p0.06_64chains", "sweeps": sum(iters), "gpu_value": gpu_rate, "gpu_kernel_ms": ms,
           "geometry": geo, "best_cut_median": float(np.median(job.best_vals)),
           "best_cut_max": int(job.best_vals.max())}
    if not check:
        log(f"[er800] gpu {ms:.1f} ms ({gpu_rate:.3e} chain-steps/s, {geo}); oracle check skipped")
        return out
    from oracle import oracle as O
    threads, visible, model = host_cpu_share()
    end, best, cv, bv, sec = O.anneal_many(csr.offsets, csr.edge_to, csr.edge_weight, job.init, job.seeds,
                                           job.beta.run_end, job.beta.run_val, job.beta.chain_run_off,
                                           LADDER, iters, threads=threads)
    same = (np.array_equal(end, job.end_spins) and np.array_equal(best, job.best_spins) and
            np.array_equal(cv, job.final_vals) and np.array_equal(bv, job.best_vals))
    recomputed = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(C))
    log(f"[er800] gpu {ms:.1f} ms ({gpu_rate:.3e} chain-steps/s, {geo}); cpu {sec:.1f} s on {threads} threads; "
        f"bit-exact per seed: {same}")
    out.update({"cpu_value": C * sum(iters) / sec, "cpu_threads": threads, "cpu_visible": visible,
                "cpu_model": model, "cpu_seconds": sec, "gpu_vs_cpu": gpu_rate / (C * sum(iters) / sec),
                "bit_exact_per_seed": bool(same), "best_cut_recomputed": bool(recomputed)})
    return out


def bench_er800(args) -> int:
    """The small configuration on its own line (1M sweeps unless --steps is
    given): GPU run and the CPU oracle on every chain, compared bit for bit
    per seed."""
    sweeps = args.steps * args.sweeps_per_step if args.steps != 20 else 1_000_000
    r = er800_leg(sweeps, args.threads, check=not args.no_cpu_baseline)
    _emit({"metric": "chain-steps/sec", **r})
    return 0 if r.get("bit_exact_per_seed", True) and r.get("best_cut_recomputed", True) else 1


def bench_sparse(args) -> int:
    """Sparse random graphs, 4096 chains each in one persistent launch:
    n=10000 m=10000 unit weights, and n=7000 m=17150 +/-1.  Flips per sweep
    come from the quality-gated tuner (256 pilot chains per candidate ladder
    over the full budget; the fastest ladder whose median best cut is within
    0.1% of the reference's 128..1 ladder), threads per chain from the
    geometry autotuner; the 128..1 ladder runs beside it for comparison."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    sweeps = args.steps * args.sweeps_per_step if args.steps != 20 else 100_000
    C = 4096
    tops = [8, 16, 32, 64, 128, 256, 512]
    rc = 0
    for name, csr in (("sparse_n10000_m10000_unit", W.sparse_random(10000, 10000, False, seed=0)),
                      ("sparse_n7000_m17150_pm1", W.sparse_random(7000, 17150, True, seed=0))):
        with R.Context(0) as ctx:
            gid = ctx.upload_graph(csr)
            t0 = time.perf_counter()
            tune = ctx.tune_flips(gid, sweeps, tops, pilot_chains=256, seed_base=1_000_000, reference_top_p=128)
            tune_s = time.perf_counter() - t0
            out = {}
            for label, top in (("tuned", tune["chosen_top_p"]), ("ladder128", 128)):
                phases, iters = R.flips_ladder(top, sweeps)
                job = R.Job(csr, np.stack([R.generate_random_spins(s, csr.n) for s in range(1, C + 1)]),
                            [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in range(1, C + 1)],
                            R.build_linear_beta_table(0.1, 3.0, sweeps, C), phases, iters)
                ms = ctx.batch_run([(gid, job)], autotune=args.threads == 0, threads=args.threads)
                geo = ctx.last_launch()
                ok = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(0, C, 97))
                rc |= 0 if ok else 1
                out[label] = {"top_p": top, "value": C * sweeps / (ms / 1e3), "kernel_ms": ms, "geometry": geo,
                              "best_cut_median": float(np.median(job.best_vals)),
                              "best_cut_max": int(job.best_vals.max()), "cuts_recomputed": bool(ok)}
        log(f"[sparse] {name}: n={csr.n} tuner {tune_s:.1f} s -> top P {tune['chosen_top_p']} "
            f"(pilot medians {tune['median_best']}); tuned {out['tuned']['value']:.3e} chain-steps/s median "
            f"{out['tuned']['best_cut_median']}, ladder128 {out['ladder128']['value']:.3e} median "
            f"{out['ladder128']['best_cut_median']}")
        _emit({"metric": "chain-steps/sec", "config": name, "n": csr.n, "m": csr.m2 // 2, "chains": C,
               "sweeps": sweeps, "value": out["tuned"]["value"], "flips_tuner": tune, "tuner_seconds": tune_s,
               "runs": out})
    return rc


def run_batch(R, graphs_count: int, sweeps: int, threads: int = 0, plain_stride: int = 1) -> dict:
    """Batched mixed instances: `graphs_count` independent graphs (n
    log-uniform in [800, 20000], family uniform over ER mean degree 6 / torus
    +/-1 / random degree 2-6 +/-1), 64 chains each, in one hipGraph-captured
    persistent launch, against the plain mode: one launch per graph, graphs
    one after another.  Every chain must be identical in both modes.  With
    plain_stride S > 1 the plain leg runs every S-th graph only, and its rate
    is that sample's work over its time."""
    from rfmf_amd import workloads as W
    iters = [sweeps // 8] * 8
    t0 = time.perf_counter()
    graphs = W.batch_instances(graphs_count, seed=0)
    log(f"[batch] generated {len(graphs)} graphs in {time.perf_counter() - t0:.1f} s; "
        f"sum n = {sum(g.n for _, g in graphs)}")
    jobs_b = [_job(R, g, range(64 * i + 1, 64 * i + 65), iters) for i, (_, g) in enumerate(graphs)]
    sample = list(range(0, len(graphs), max(1, plain_stride)))
    jobs_p = {i: _job(R, graphs[i][1], range(64 * i + 1, 64 * i + 65), iters) for i in sample}
    work = 64 * len(graphs) * sum(iters)
    work_p = 64 * len(sample) * sum(iters)
    with R.Context(0) as ctx:
        gids = [ctx.upload_graph(g) for _, g in graphs]
        ms_b = ctx.batch_run(list(zip(gids, jobs_b)), slice_sweeps=sum(iters), use_hip_graph=True,
                             autotune=threads == 0, threads=threads)
        geo = ctx.last_launch()
        log(f"[batch] batched {ms_b:.1f} ms; plain leg on {len(sample)} of {len(graphs)} graphs")
        ms_p = 0.0
        for k, i in enumerate(sample):
            ms_p += ctx.batch_run([(gids[i], jobs_p[i])])
            if (k + 1) % 16 == 0:
                log(f"[batch] plain {k + 1}/{len(sample)} graphs, {ms_p / 1e3:.1f} s of kernels")
    same = all(np.array_equal(jobs_b[i].end_spins, j.end_spins) and
               np.array_equal(jobs_b[i].best_spins, j.best_spins) and
               np.array_equal(jobs_b[i].best_vals, j.best_vals) for i, j in jobs_p.items())
    rb, rp = work / (ms_b / 1e3), work_p / (ms_p / 1e3)
    med_b = float(np.median(np.concatenate([j.best_vals for j in jobs_b])))
    med_p = float(np.median(np.concatenate([j.best_vals for j in jobs_p.values()])))
    med_bs = float(np.median(np.concatenate([jobs_b[i].best_vals for i in sample])))
    log(f"[batch] batched {ms_b:.1f} ms {rb:.3e} chain-steps/s ({geo}); plain {ms_p:.1f} ms {rp:.3e}; "
        f"x{rb / rp:.2f}; identical: {same}")
    out = {"config": f"batch_{len(graphs)}graphs_64chains", "sweeps": sum(iters), "chains": 64 * len(graphs),
           "batched_value": rb, "plain_value": rp, "speedup": rb / rp, "batched_kernel_ms": ms_b,
           "plain_kernel_ms": ms_p, "geometry": geo, "identical_results": bool(same),
           "best_cut_median_batched": med_b, "best_cut_median_plain": med_p}
    if len(sample) < len(graphs):
        out.update({"plain_graphs": len(sample), "best_cut_median_batched_same_graphs": med_bs})
    return out


def bench_batch(args) -> int:
    import rfmf_amd as R
    sweeps = args.steps * args.sweeps_per_step if args.steps != 20 else args.batch_sweeps
    r = run_batch(R, args.graphs, sweeps, args.threads, args.plain_stride)
    _emit({"metric": "chain-steps/sec", **r})
    return 0 if r["identical_results"] else 1


def run_serve(R, nreq: int, qps: float, slice_sweeps: int, lo: int = 10_000, hi: int = 100_000,
              plain_leg: bool = True) -> dict:
    """Serving: the request stream into the continuous-batching server (HBM
    footprint admission), against plain per-request launches served one at a
    time in arrival order (upload, one-graph launch, download per request).
    Both modes are driven by the same real arrival clock and their latencies
    observed (completion minus arrival).  Every request's result must be
    identical in both modes."""
    from rfmf_amd import workloads as W
    stream, arrivals = W.request_stream(nreq, qps, lo, hi)
    reqs = [(g, c, s) for _, g, c, s in stream]
    work = float(sum(c * s for _, c, s in reqs))

    def mk(i, g, chains, sweeps):
        return _job(R, g, range(1000 * i + 1, 1000 * i + 1 + chains), [sweeps // 8] * 8)

    served = [mk(i, *r) for i, r in enumerate(reqs)]
    plain = [mk(i, *r) for i, r in enumerate(reqs)]
    log(f"[serve] {nreq} requests, {qps} qps, sweeps log-uniform [{lo}, {hi}], total {work:.3e} chain-steps")
    with R.Context(0) as ctx:
        # plain: one request at a time, first come first served
        t0 = time.perf_counter()
        finish = []
        for i, j in enumerate(plain if plain_leg else []):
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            gid = ctx.upload_graph(j.csr)
            ctx.batch_run([(gid, j)])
            ctx.release_graph(gid)
            finish.append(time.perf_counter() - t0)
            if (i + 1) % 4 == 0:
                log(f"[serve] plain {i + 1}/{nreq} requests done at {finish[-1]:.1f} s")
        if plain_leg:
            lat_p = np.array(finish) - arrivals
            span_p = finish[-1] - arrivals[0]
            log(f"[serve] plain FIFO done in {finish[-1]:.1f} s")
        else:
            lat_p, span_p = np.full(nreq, np.nan), float("nan")
        with R.Server(ctx, hbm_budget_bytes=64 << 30, slice_sweeps=slice_sweeps) as srv:
            t0 = time.perf_counter()
            ids = []
            for i, j in enumerate(served):
                delay = arrivals[i] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(j))
            for k, rid in enumerate(ids):
                srv.wait(rid, 1200.0)
                if (k + 1) % 4 == 0:
                    log(f"[serve] served {k + 1}/{nreq} requests done at {time.perf_counter() - t0:.1f} s")
            polls = [srv.poll(rid) for rid in ids]
            st = srv.stats()
    lat_s = np.array([td - ts for _, ts, _, td in polls])
    span_s = max(td for *_, td in polls) - min(ts for _, ts, _, _ in polls)
    same = plain_leg and all(np.array_equal(a.end_spins, b.end_spins) and
                             np.array_equal(a.best_spins, b.best_spins) and
                             np.array_equal(a.best_vals, b.best_vals) for a, b in zip(served, plain))
    med_s = float(np.median(np.concatenate([j.best_vals for j in served])))
    med_p = float(np.median(np.concatenate([j.best_vals for j in plain])))
    out = {"config": f"serve_{nreq}req_{qps:g}qps_sweeps{lo}-{hi}", "served_value": work / span_s,
           "plain_value": work / span_p, "speedup": span_p / span_s,
           "served_p50_s": float(np.percentile(lat_s, 50)), "served_p99_s": float(np.percentile(lat_s, 99)),
           "plain_p50_s": float(np.percentile(lat_p, 50)), "plain_p99_s": float(np.percentile(lat_p, 99)),
           "slices": st["slices"], "busy_kernel_ms": st["busy_kernel_ms"], "served_span_s": span_s,
           "plain_span_s": span_p, "identical_results": bool(same),
           "best_cut_median_served": med_s, "best_cut_median_plain": med_p}
    log(f"[serve] {out}")
    return out


def bench_serve(args) -> int:
    import rfmf_amd as R
    lo, hi = (int(x) for x in args.serve_sweeps.split(","))
    r = run_serve(R, args.requests, args.qps, args.slice, lo, hi, plain_leg=not args.no_plain)
    _emit({"metric": "chain-steps/sec and solve latency", **r})
    return 0 if r["identical_results"] or args.no_plain else 1


def bench_gset(args) -> int:
    """The reference's own benchmark table on MI355X: for each of the 12 Gset
    instances, the 80 chains of the reference's V100-SXM2 run (the same raw
    seeds, t0 ladder 35..400 and P ladder 128..1 with equal sweeps per phase,
    tests/golden/reference_runs.json from results/V100-SXM2-10M/results.csv),
    one launch per instance with the per-phase autotuned geometry, at
    `--gset-sweeps` sweeps per chain (the reference ran 8e7).  Writes a
    results.csv-style table (`--results-csv`) and one JSON line per
    instance with chain-sweeps/s beside the reference's."""
    import rfmf_amd as R
    ref = json.loads((ROOT / "tests" / "golden" / "reference_runs.json").read_text())["V100-SXM2-10M"]
    sweeps = args.gset_sweeps // 8 * 8
    iters = [sweeps // 8] * 8
    names = [s for s in args.instances.split(",") if s] or sorted(ref, key=lambda k: int(k[1:]))
    rows = []
    rc = 0
    for inst in names:
        r = ref[inst]
        csr, _ = R.read_instance_from_file(str(ROOT / "tests" / "golden" / "gset" / inst))
        seeds, t0s = r["seeds"], r["t0s"]
        job = R.Job(csr, np.stack([R.generate_random_spins(s, csr.n) for s in seeds]),
                    [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds],
                    R.build_beta_run_tables(t0s, sum(iters)), LADDER, iters)
        t_wall = time.perf_counter()
        with R.Context(0) as ctx:
            gid = ctx.upload_graph(csr)
            ms = ctx.batch_run([(gid, job)], autotune=args.threads == 0, threads=args.threads)
            geo = ctx.last_launch()
        wall = time.perf_counter() - t_wall
        ok = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(len(seeds)))
        rc |= 0 if ok else 1
        chains = len(seeds)
        rate = chains * sum(iters) / (ms / 1e3)
        ref_rate = chains * sum(r["iters"]) / r["kernel_s"]
        best = int(job.best_vals.max())
        row = {"instance": inst, "n": csr.n, "m": csr.m2 // 2, "chains": chains, "best_cut": best,
               "hits_at_best": int((job.best_vals == best).sum()), "best_chain": int(np.argmax(job.best_vals)),
               "best_t0": float(t0s[int(np.argmax(job.best_vals))]), "kernel_s": ms / 1e3, "wall_s": wall,
               "sweeps": sum(iters), "chain_sweeps_per_s": rate, "us_per_sweep": ms * 1e3 / sum(iters),
               "ref_sweeps": sum(r["iters"]), "ref_kernel_s": r["kernel_s"], "ref_chain_sweeps_per_s": ref_rate,
               "ratio_vs_reference": rate / ref_rate, "ref_best_cut": r["best_cut"], "cuts_recomputed": bool(ok),
               "threads_last_phase": geo["threads_per_chain"]}
        rows.append(row)
        log(f"[gset] {inst}: n={csr.n} {ms:.1f} ms {rate:.3e} chain-sweeps/s {row['us_per_sweep']:.2f} us/sweep "
            f"(V100-SXM2 {ref_rate:.2e}: {rate / ref_rate:.2f}x); best {best} after {sum(iters)} sweeps "
            f"(reference {r['best_cut']} after {sum(r['iters'])})")
        _emit({"metric": "chain-sweeps/sec", "config": f"gset_{inst}_80chains", **row})
    if args.results_csv:
        cols = list(rows[0].keys())
        with open(args.results_csv, "w") as f:
            f.write(",".join(cols) + "\n")
            for row in rows:
                f.write(",".join(f"{row[c]:.6g}" if isinstance(row[c], float) else str(row[c]) for c in cols) + "\n")
        log(f"[gset] wrote {args.results_csv}")
    return rc


def bench_large(args) -> int:
    """A graph whose chain state does not fit LDS: torus 1000x1000 +/-1
    (n = 10^6, 1 MB of 8-bit words per chain), 256 chains x 800 sweeps,
    ladder 128..1, linear beta: the chains work on their state in place in
    HBM.  Beside it, the headline torus 100x200 (1024 chains x 8000 sweeps)
    with every chain forced to HBM-resident state, against its LDS-resident
    rate.  Geometry from the built-in heuristic (runs too short for the
    autotuner's trial slices)."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    rc = 0
    for name, csr, C, sweeps, kw in (
            ("torus1000x1000_256chains", W.torus(1000, 1000, seed=1), 256, 800, {}),
            ("torus100x200_1024chains_lds", W.torus(100, 200, seed=0), 1024, 8000, {}),
            ("torus100x200_1024chains_hbm", W.torus(100, 200, seed=0), 1024, 8000, {"state_in_hbm": True})):
        iters = [sweeps // 8] * 8
        job = _job(R, csr, range(1, C + 1), iters)
        with R.Context(0) as ctx:
            gid = ctx.upload_graph(csr)
            ms = ctx.batch_run([(gid, job)], threads=args.threads, **kw)
            geo = ctx.last_launch()
        ok = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(0, C, 64))
        rc |= 0 if ok else 1
        rate = C * sum(iters) / (ms / 1e3)
        log(f"[large] {name}: n={csr.n} {ms:.1f} ms {rate:.3e} chain-steps/s {rate * csr.n:.3e} vertex-steps/s "
            f"({geo})")
        _emit({"metric": "chain-steps/sec", "config": name, "n": csr.n, "chains": C, "sweeps": sum(iters),
               "value": rate, "vertex_steps_per_s": rate * csr.n, "geometry": geo, "cuts_recomputed": bool(ok)})
    return rc


def run_two_slices(ctx, gid, job, warm: int, timed: int):
    """Admits the job, runs `warm` sweeps in one launch and `timed` sweeps in
    the next (each bracketed by HIP events on the library's stream); returns
    (warm ms, timed ms).  With no warmup the first window is empty."""
    ms = ctx.batch_run_windows([(gid, job)], [warm, warm + timed])
    return ms[0], ms[1]


if __name__ == "__main__":
    sys.exit(main())
