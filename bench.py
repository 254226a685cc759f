"""Benchmark: chain-steps/sec of the multi-chain annealer on one MI355X.

Workload (one graph per launch, the headline configuration): 2D toroidal
grid 100 x 200 (n = 20000, m = 40000, degree 4), iid uniform +/-1 weights,
1024 chains, linear beta 0.1 -> 3.0, group-count ladder P = 128, 64, ..., 1
with equal sweeps per phase.  A "step" is one sweep of every chain.

W warmup sweeps run first (untimed, at P = 128); then exactly K sweeps per
chain run as ONE persistent kernel launch timed with HIP events on the
library's stream (inputs and chain state already resident in HBM), with the
P ladder spread evenly over the K timed sweeps.  The launch uses the
threads-per-chain heuristic (256 at 4 chains per CU), which is the tile the
live autotuner picks for every phase of this workload; autotuning the window
instead measures the same kernel plus its trial launches.

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
    ap.add_argument("--steps", type=int, default=40000, help="timed sweeps per chain")
    ap.add_argument("--warmup", type=int, default=2000, help="untimed sweeps per chain before the timed launch")
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--rows", type=int, default=100)
    ap.add_argument("--cols", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    # CPU sample: 64 chains x 64000 sweeps = 4.1e6 chain-steps, about 10 s on 16 host threads
    ap.add_argument("--cpu-chains", type=int, default=64)
    ap.add_argument("--cpu-sweeps", type=int, default=64000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (counter calibration) and exit")
    ap.add_argument("--config", default="torus",
                    choices=["torus", "er800", "sparse", "batch", "serve", "gset", "large"],
                    help="torus: the headline plain run (default); the others are the other workloads")
    ap.add_argument("--threads", type=int, default=0, help="threads per chain (0: autotune)")
    ap.add_argument("--qps", type=float, default=20.0, help="serve: Poisson arrival rate")
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--slice", type=int, default=500, help="serve: sweeps per serving slice")
    ap.add_argument("--graphs", type=int, default=256, help="batch: number of graphs")
    args = ap.parse_args()

    if args.config != "torus" and not args.copy_probe:
        return {"er800": bench_er800, "sparse": bench_sparse, "batch": bench_batch,
                "serve": bench_serve, "gset": bench_gset, "large": bench_large}[args.config](args)

    if args.copy_probe:
        import rfmf_amd as R
        with R.Context(0) as ctx:
            ms = ctx.copy_probe(args.copy_probe, iters=5)
        log(f"[bench] copy probe {args.copy_probe} B: {ms:.3f} ms/launch, "
            f"{2 * args.copy_probe / ms / 1e6:.1f} GB/s read+write")
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the annealer shards by chains and graphs with no cross-GPU "
              f"exchange, and this build runs on one MI355X only (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    import rfmf_amd as R
    from rfmf_amd import workloads as W

    K = (args.steps // 8) * 8
    Wm = args.warmup
    csr = W.torus(args.rows, args.cols, seed=0)
    n, C = csr.n, args.chains
    B = R.max_weighted_degree(csr)
    log(f"[bench] torus {args.rows}x{args.cols} n={n} m2={csr.m2} B={B} digest={W.csr_digest(csr)} chains={C} "
        f"warmup={Wm} timed={K}")

    phases = list(LADDER)
    iters = [K // 8] * 8
    iters[0] += Wm
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
        ms_total, slices = run_two_slices(ctx, gid, job, Wm, K)
        wall = time.perf_counter() - t0
        warm_ms, timed_ms = slices
        log(f"[bench] warmup {warm_ms:.1f} ms, timed {timed_ms:.1f} ms, host wall {wall:.2f} s")

    best = job.best_vals
    for c in range(min(C, 8)):
        assert R.cut_value_csr(csr, job.best_spins[c]) == best[c], "incremental best != recomputed cut"
    chain_steps = C * K
    value = chain_steps / (timed_ms / 1e3)
    p_timed = np.repeat(LADDER, K // 8)
    bytes_per_chain = algorithmic_bytes_per_chain(n, 4.0, p_timed)
    achieved = C * bytes_per_chain / (timed_ms / 1e3) / 1e9
    # HBM bytes and VALU instructions come from the committed rocprofv3
    # counter runs of this same workload (FETCH_SIZE x2.0 + WRITE_SIZE x1.0,
    # calibrated on a plain copy kernel); traffic only applies to a launch of
    # the profiled size.
    traffic, valu = None, None
    if COUNTERS.exists():
        t = json.loads(COUNTERS.read_text())
        if t.get("steps") == K and t.get("chains") == C and (args.rows, args.cols) == (100, 200):
            traffic = t.get("bytes_per_launch")
        if (args.rows, args.cols) == (100, 200):
            peak = VALU_PEAK_GINSTR_PER_CU * info["cu"]
            got = t["valu_instr_per_chain_step"] * value / 1e9
            valu = {"bound": "valu", "achieved": got, "peak": peak, "unit": "G wave64-instr/s", "frac": got / peak,
                    "instr_per_chain_step": t["valu_instr_per_chain_step"]}

    cpu = None
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(16, cores))
        rate, sec = cpu_baseline(csr, args.cpu_chains, args.cpu_sweeps, threads)
        cpu = {"value": rate, "unit": "chain-steps/s", "cores": threads, "kind": "port",
               "sample": f"oracle (serial C restatement), torus 100x200, {args.cpu_chains} chains x "
                         f"{args.cpu_sweeps} sweeps, same P ladder and linear beta, {sec:.2f} s wall"}
        log(f"[bench] cpu baseline {rate:.3e} chain-steps/s on {threads} threads ({sec:.2f} s)")

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
                   "schedule": "linear beta 0.1->3.0, P ladder 128..1 equal sweeps", "parallelism": "chains",
                   "best_cut_median": float(np.median(best)), "best_cut_max": int(best.max())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "note": "achieved = algorithmic bytes (5n + 6P + 16 sum deg per chain-step); the chain "
                             "state is LDS-resident, so real HBM traffic is `traffic` and the binding "
                             "limit is VALU issue (`valu`)",
                     "valu": valu},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def _job(R, csr, seeds, iters, beta_range=(0.1, 3.0)):
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    beta = R.build_linear_beta_table(beta_range[0], beta_range[1], sum(iters), len(seeds))
    return R.Job(csr, init, solver, beta, LADDER, iters)


def _emit(obj):
    print(json.dumps(obj), flush=True)


def bench_er800(args) -> int:
    """ER n=800 p=0.06 unit weights, 64 chains, 1M sweeps (linear beta 0.1->3.0,
    P ladder 128..1): GPU run and the CPU oracle on every chain, compared bit
    for bit per seed."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    from oracle import oracle as O
    sweeps = args.steps if args.steps != 40000 else 1_000_000
    csr = W.erdos_renyi(800, 0.06, seed=0)
    C = 64
    iters = [sweeps // 8] * 8
    job = _job(R, csr, range(1, C + 1), iters)
    log(f"[er800] n={csr.n} m={csr.m2 // 2} B={R.max_weighted_degree(csr)} chains={C} sweeps={sum(iters)}")
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ms = ctx.batch_run([(gid, job)], autotune=args.threads == 0, threads=args.threads)
        geo = ctx.last_launch()
    gpu_rate = C * sum(iters) / (ms / 1e3)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, cores))
    end, best, cv, bv, sec = O.anneal_many(csr.offsets, csr.edge_to, csr.edge_weight, job.init, job.seeds,
                                           job.beta.run_end, job.beta.run_val, job.beta.chain_run_off,
                                           LADDER, iters, threads=threads)
    same = (np.array_equal(end, job.end_spins) and np.array_equal(best, job.best_spins) and
            np.array_equal(cv, job.final_vals) and np.array_equal(bv, job.best_vals))
    recomputed = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(C))
    log(f"[er800] gpu {ms:.1f} ms ({gpu_rate:.3e} chain-steps/s, {geo}); cpu {sec:.1f} s on {threads} threads; "
        f"bit-exact per seed: {same}")
    _emit({"metric": "chain-steps/sec", "config": "er800_p0.06_64chains", "sweeps": sum(iters),
           "gpu_value": gpu_rate, "gpu_kernel_ms": ms, "geometry": geo,
           "cpu_value": C * sum(iters) / sec, "cpu_threads": threads, "cpu_seconds": sec,
           "bit_exact_per_seed": bool(same), "best_cut_recomputed": bool(recomputed),
           "best_cut_median": float(np.median(job.best_vals)), "best_cut_max": int(job.best_vals.max())})
    return 0 if same and recomputed else 1


def bench_sparse(args) -> int:
    """Sparse random graphs, 4096 chains each in one persistent launch:
    n=10000 m=10000 unit weights, and n=7000 m=17150 +/-1.  Flips per sweep
    come from the quality-gated tuner (256 pilot chains per candidate ladder
    over the full budget; the fastest ladder whose median best cut is within
    0.1% of the reference's 128..1 ladder), threads per chain from the
    geometry autotuner; the 128..1 ladder runs beside it for comparison."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    sweeps = args.steps if args.steps != 40000 else 100_000
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


def bench_batch(args) -> int:
    """Batched mixed instances: `graphs` independent graphs (n log-uniform in
    [800, 20000], family uniform over ER mean degree 6 / torus +/-1 / random
    degree 2-6 +/-1), 64 chains each, in one hipGraph-captured persistent
    launch, against the plain mode: one launch per graph, graphs one after
    another.  Every chain must be identical in both modes."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    sweeps = args.steps if args.steps != 40000 else 20_000
    iters = [sweeps // 8] * 8
    t0 = time.perf_counter()
    graphs = W.batch_instances(args.graphs, seed=0)
    log(f"[batch] generated {len(graphs)} graphs in {time.perf_counter() - t0:.1f} s; "
        f"sum n = {sum(g.n for _, g in graphs)}")
    jobs_b = [_job(R, g, range(64 * i + 1, 64 * i + 65), iters) for i, (_, g) in enumerate(graphs)]
    jobs_p = [_job(R, g, range(64 * i + 1, 64 * i + 65), iters) for i, (_, g) in enumerate(graphs)]
    work = 64 * len(graphs) * sum(iters)
    with R.Context(0) as ctx:
        gids = [ctx.upload_graph(g) for _, g in graphs]
        ms_b = ctx.batch_run(list(zip(gids, jobs_b)), slice_sweeps=sum(iters), use_hip_graph=True,
                             autotune=args.threads == 0, threads=args.threads)
        geo = ctx.last_launch()
        ms_p = 0.0
        for gid, j in zip(gids, jobs_p):
            ms_p += ctx.batch_run([(gid, j)])
    same = all(np.array_equal(a.end_spins, b.end_spins) and np.array_equal(a.best_spins, b.best_spins) and
               np.array_equal(a.best_vals, b.best_vals) for a, b in zip(jobs_b, jobs_p))
    rb, rp = work / (ms_b / 1e3), work / (ms_p / 1e3)
    med_b = float(np.median(np.concatenate([j.best_vals for j in jobs_b])))
    med_p = float(np.median(np.concatenate([j.best_vals for j in jobs_p])))
    log(f"[batch] batched {ms_b:.1f} ms {rb:.3e} chain-steps/s ({geo}); plain {ms_p:.1f} ms {rp:.3e}; "
        f"x{rb / rp:.2f}; identical: {same}")
    _emit({"metric": "chain-steps/sec", "config": f"batch_{len(graphs)}graphs_64chains", "sweeps": sum(iters),
           "batched_value": rb, "plain_value": rp, "speedup": rb / rp, "batched_kernel_ms": ms_b,
           "plain_kernel_ms": ms_p, "geometry": geo, "identical_results": bool(same),
           "best_cut_median_batched": med_b, "best_cut_median_plain": med_p})
    return 0 if same else 1


def bench_serve(args) -> int:
    """Serving: a Poisson stream of solve requests at `qps`, each a seeded
    graph from the batched-mix generator (n 800-20000) with 32-256 chains and
    a log-uniform sweep count, admitted by HBM footprint into a continuously
    batched server, against plain per-request launches served one at a time
    in arrival order.  Reports p50/p99 latency and aggregate chain-steps/s;
    every request's result must be identical in both modes."""
    import rfmf_amd as R
    from rfmf_amd import workloads as W
    rng = np.random.default_rng(2026)
    nreq = args.requests
    graphs = W.batch_instances(nreq, seed=7)
    lo, hi = 10_000, 100_000
    reqs = []
    for i, (_, g) in enumerate(graphs):
        chains = int(rng.integers(32, 257))
        sweeps = int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) // 8 * 8
        reqs.append((g, chains, sweeps))
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, nreq))
    work = float(sum(c * s for _, c, s in reqs))

    def mk(i, g, chains, sweeps):
        return _job(R, g, range(1000 * i + 1, 1000 * i + 1 + chains), [sweeps // 8] * 8)

    served = [mk(i, *r) for i, r in enumerate(reqs)]
    plain = [mk(i, *r) for i, r in enumerate(reqs)]
    log(f"[serve] {nreq} requests, {args.qps} qps, sweeps log-uniform [{lo}, {hi}], total {work:.3e} chain-steps")
    with R.Context(0) as ctx:
        # plain: each request alone, wall time of its whole call (upload, run, download)
        svc = []
        for j in plain:
            t = time.perf_counter()
            gid = ctx.upload_graph(j.csr)
            ctx.batch_run([(gid, j)])
            ctx.release_graph(gid)
            svc.append(time.perf_counter() - t)
        with R.Server(ctx, hbm_budget_bytes=64 << 30, slice_sweeps=args.slice) as srv:
            t0 = time.perf_counter()
            ids = []
            for i, j in enumerate(served):
                delay = arrivals[i] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(j))
            for rid in ids:
                srv.wait(rid, 1200.0)
            polls = [srv.poll(rid) for rid in ids]
            st = srv.stats()
    lat_s = np.array([td - ts for _, ts, _, td in polls])
    span_s = max(td for *_, td in polls) - min(ts for _, ts, _, _ in polls)
    finish, start_free = [], 0.0
    for a, s in zip(arrivals, svc):
        start = max(a, start_free)
        start_free = start + s
        finish.append(start_free)
    lat_p = np.array(finish) - arrivals
    span_p = finish[-1] - arrivals[0]
    same = all(np.array_equal(a.end_spins, b.end_spins) and np.array_equal(a.best_spins, b.best_spins)
               for a, b in zip(served, plain))
    out = {"metric": "chain-steps/sec and solve latency", "config": f"serve_{nreq}req_{args.qps}qps",
           "served_value": work / span_s, "plain_value": work / span_p, "speedup": span_p / span_s,
           "served_p50_s": float(np.percentile(lat_s, 50)), "served_p99_s": float(np.percentile(lat_s, 99)),
           "plain_p50_s": float(np.percentile(lat_p, 50)), "plain_p99_s": float(np.percentile(lat_p, 99)),
           "slices": st["slices"], "busy_kernel_ms": st["busy_kernel_ms"], "identical_results": bool(same)}
    log(f"[serve] {out}")
    _emit(out)
    return 0 if same else 1


# Reference runs on Tesla V100-SXM2, 80 chains, 8e7 sweeps per chain:
# chain-sweeps/s from results/V100-SXM2-10M/results.csv:3,4,10,13.
V100_SXM2 = {"G61": 7.40e6, "G62": 8.09e6, "G70": 7.31e6, "G81": 5.32e6}


def bench_gset(args) -> int:
    """The reference's own benchmark shape on the Gset instances held here:
    80 chains (t0 in {35, 50, 70, 100, 140, 200, 280, 400} x 10 replicas),
    ladder 128..1 with equal sweeps per phase, log schedule, one launch per
    instance with the per-phase autotuned geometry; chain-sweeps/s next to
    the reference's published V100-SXM2 figure."""
    import rfmf_amd as R
    sweeps = args.steps if args.steps != 40000 else 100_000
    iters = [sweeps // 8] * 8
    t0s = [float(t) for t in (35, 50, 70, 100, 140, 200, 280, 400) for _ in range(10)]
    rc = 0
    for inst in ("G61", "G62", "G70", "G81"):
        csr, _ = R.read_instance_from_file(str(ROOT / "tests" / "golden" / "gset" / inst))
        seeds = list(range(1, 81))
        job = R.Job(csr, np.stack([R.generate_random_spins(s, csr.n) for s in seeds]),
                    [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds],
                    R.build_beta_run_tables(t0s, sum(iters)), LADDER, iters)
        with R.Context(0) as ctx:
            gid = ctx.upload_graph(csr)
            ms = ctx.batch_run([(gid, job)], autotune=args.threads == 0, threads=args.threads)
            geo = ctx.last_launch()
        ok = all(R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c] for c in range(80))
        rc |= 0 if ok else 1
        rate = 80 * sum(iters) / (ms / 1e3)
        log(f"[gset] {inst}: n={csr.n} {ms:.1f} ms {rate:.3e} chain-sweeps/s {ms * 1e3 / sum(iters):.2f} us/sweep "
            f"(V100-SXM2 {V100_SXM2[inst]:.2e}: {rate / V100_SXM2[inst]:.2f}x) last phase {geo['threads_per_chain']} "
            f"threads")
        _emit({"metric": "chain-sweeps/sec", "config": f"gset_{inst}_80chains", "n": csr.n, "sweeps": sum(iters),
               "value": rate, "us_per_sweep": ms * 1e3 / sum(iters), "reference_v100_sxm2": V100_SXM2[inst],
               "ratio_vs_reference": rate / V100_SXM2[inst], "best_cut_max": int(job.best_vals.max()),
               "cuts_recomputed": bool(ok)})
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
    the next; returns (total ms, (warm ms, timed ms))."""
    ms = ctx.batch_run_windows([(gid, job)], [warm, warm + timed])
    return ms[0] + ms[1], (ms[0], ms[1])


if __name__ == "__main__":
    sys.exit(main())
