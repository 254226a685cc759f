"""Batched multi-graph launches, hipGraph replay, slicing, chain isolation,
the step-wise run API and the serving loop: every chain's result must be
byte-identical to its one-shot run (and to the oracle)."""
import time

import numpy as np
import pytest

import rfmf_amd as R
from rfmf_amd import workloads as W
from conftest import run_oracle_chains

pytestmark = pytest.mark.gpu

LADDER = [128, 64, 32, 16, 8, 4, 2, 1]


def make_job(csr, seeds, phases, iters, linear=True, t0=None):
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    if linear:
        beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), len(seeds))
    else:
        beta = R.build_beta_run_tables(t0, sum(iters))
    return R.Job(csr, init, solver, beta, phases, iters)


def one_shot(job):
    return R.anneal_gpu_tables(job.csr, R.max_weighted_degree(job.csr), job.init, job.beta, job.phases,
                               job.iters, job.seeds)


def same(job, res, label):
    assert np.array_equal(job.end_spins, res.end_spins), f"{label}: final spins differ"
    assert np.array_equal(job.best_spins, res.best_spins), f"{label}: best spins differ"
    assert np.array_equal(job.best_vals, res.kernel_best_vals), f"{label}: best cuts differ"
    assert np.array_equal(job.final_vals, res.kernel_final_vals), f"{label}: final cuts differ"


@pytest.fixture(scope="module")
def graphs(gset):
    return {
        "G11": gset("G11"),
        "ER800": W.erdos_renyi(800, 0.06, seed=3),      # 16-bit words
        "G70": gset("G70"),
        "deg": W.degree_random(1500, 2, 6, seed=4),
        "torus": W.torus(30, 60, seed=5),
    }


def _mixed_jobs(graphs):
    return [
        ("G11", make_job(graphs["G11"], [1, 2, 3], [8, 4, 2, 1], [60] * 4)),
        ("ER800", make_job(graphs["ER800"], [4, 5], LADDER, [25] * 8)),
        ("G70", make_job(graphs["G70"], [6], [8, 2, 1], [50, 50, 50], linear=False, t0=[105.0])),
        ("deg", make_job(graphs["deg"], [7, 8, 9, 10], [3, 64, 1], [40, 40, 40])),
        ("torus", make_job(graphs["torus"], [11, 12], LADDER, [30] * 8)),
        # beta_max * B + |gumbel| max >= 2^31 - 1: int64 scores, its own launch group
        ("G11", make_job(graphs["G11"], [15, 16], [8, 4, 2, 1], [60] * 4, linear=False, t0=[0.005, 0.01])),
    ]


def test_batch_mixed_graphs_equal_one_shot(graphs):
    print("[batch] 5 graphs, mixed word widths and schedules, one batch", flush=True)
    jobs = _mixed_jobs(graphs)
    with R.Context(0) as ctx:
        ids = {k: ctx.upload_graph(g) for k, g in graphs.items()}
        ctx.batch_run([(ids[k], j) for k, j in jobs])
    for k, j in jobs:
        same(j, one_shot(j), f"batch {k}")
    k, j = jobs[3]
    end, best, cv, bv = run_oracle_chains(j.csr, j.init, j.seeds, j.beta, j.phases, j.iters)
    assert np.array_equal(j.end_spins, end) and np.array_equal(j.best_spins, best)


@pytest.mark.parametrize("slice_sweeps,use_graph", [(1, False), (37, False), (50, True), (999, True),
                                                     (5000, True)])
def test_slicing_and_graph_replay_invariance(graphs, slice_sweeps, use_graph):
    # 5000-sweep slices are long enough for size classes: the captured graph
    # then forks the classes onto the concurrent launch lanes.
    print(f"[batch] slice={slice_sweeps} hipGraph={use_graph}", flush=True)
    jobs = _mixed_jobs(graphs)
    with R.Context(0) as ctx:
        ids = {k: ctx.upload_graph(g) for k, g in graphs.items()}
        ctx.batch_run([(ids[k], j) for k, j in jobs], slice_sweeps=slice_sweeps, use_hip_graph=use_graph)
    for k, j in jobs:
        same(j, one_shot(j), f"slice {slice_sweeps} {k}")


@pytest.mark.parametrize("threads,autotune", [(128, False), (512, False), (1024, False), (0, True)])
def test_threads_per_chain_and_autotune_invariance(graphs, threads, autotune):
    print(f"[batch] threads={threads} autotune={autotune}", flush=True)
    jobs = _mixed_jobs(graphs)
    with R.Context(0) as ctx:
        ids = {k: ctx.upload_graph(g) for k, g in graphs.items()}
        ctx.batch_run([(ids[k], j) for k, j in jobs], slice_sweeps=45, threads=threads, autotune=autotune)
        geo = ctx.last_launch()
    if threads:
        assert geo["threads_per_chain"] == threads
    else:
        assert len(geo["trial_ms_per_sweep"]) >= 2
    for k, j in jobs:
        same(j, one_shot(j), f"threads {threads} {k}")


@pytest.mark.parametrize("score_table,mode", [(-1, "int32-mul24"), (0, "mul24-table"), (1, "mul24-table")])
def test_score_table_choice_invariance(graphs, score_table, mode):
    # 8-bit words with 24-bit scores run either with the per-sweep biased LUT
    # or with the static LDS score table; the choice never changes a result.
    print(f"[batch] score_table={score_table}", flush=True)
    jobs = _mixed_jobs(graphs)
    torus = make_job(graphs["torus"], [21, 22, 23], LADDER, [20] * 8)
    with R.Context(0) as ctx:
        ids = {k: ctx.upload_graph(g) for k, g in graphs.items()}
        ctx.batch_run([(ids[k], j) for k, j in jobs], slice_sweeps=45, score_table=score_table)
        ctx.batch_run([(ids["torus"], torus)], score_table=score_table)
        assert ctx.last_launch()["score_mode"] == mode
    for k, j in jobs:
        same(j, one_shot(j), f"score_table {score_table} {k}")
    end, best, cv, bv = run_oracle_chains(torus.csr, torus.init, torus.seeds, torus.beta, torus.phases,
                                          torus.iters)
    assert np.array_equal(torus.end_spins, end) and np.array_equal(torus.best_spins, best)
    assert np.array_equal(torus.best_vals, bv)


@pytest.mark.parametrize("slice_sweeps,use_graph", [(0, False), (37, True)])
def test_state_in_hbm_invariance(graphs, slice_sweeps, use_graph):
    # Every chain working on its state in place in HBM (the path of graphs too
    # large for LDS) gives the same chains as the LDS-resident launches.
    print(f"[batch] state_in_hbm slice={slice_sweeps} hipGraph={use_graph}", flush=True)
    jobs = _mixed_jobs(graphs)
    with R.Context(0) as ctx:
        ids = {k: ctx.upload_graph(g) for k, g in graphs.items()}
        ctx.batch_run([(ids[k], j) for k, j in jobs], slice_sweeps=slice_sweeps, use_hip_graph=use_graph,
                      state_in_hbm=True)
        assert ctx.last_launch()["state_in_hbm"]
    for k, j in jobs:
        same(j, one_shot(j), f"state_in_hbm {k}")


def test_chain_isolation_beyond_resident_capacity(gset):
    # 3000 chains of an 800-vertex graph exceed the resident workgroups of
    # 256 CUs, so the persistent workgroups take several chains each.
    csr = gset("G11")
    seeds = list(range(501, 3501))
    print(f"[isolation] {len(seeds)} chains of G11", flush=True)
    job = make_job(csr, seeds, [8, 1], [60, 40])
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)])
    for c in (0, 1, 1500, 2999):
        alone = make_job(csr, [seeds[c]], [8, 1], [60, 40])
        r = one_shot(alone)
        assert np.array_equal(job.end_spins[c], r.end_spins[0]), f"chain {c} differs from its C=1 run"
        assert np.array_equal(job.best_spins[c], r.best_spins[0])
        assert job.best_vals[c] == r.kernel_best_vals[0]


def test_run_api_prefix_matches_oracle(graphs):
    csr = graphs["torus"]
    phases, iters = LADDER, [40] * 8
    job = make_job(csr, [21, 22, 23], phases, iters)
    print("[run] advance 100 + collect, then to completion", flush=True)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        with R.Run(ctx, [(gid, job)]) as run:
            _, fin = run.advance(100)
            assert not fin
            run.collect()
            mid_end, mid_best, mid_bv = job.end_spins.copy(), job.best_spins.copy(), job.best_vals.copy()
            done = False
            while not done:
                _, done = run.advance(70, use_hip_graph=True)
            run.collect()
    # prefix: the first 100 sweeps are phases 128, 64 (40 each) and 20 of P=32
    end, best, cv, bv = run_oracle_chains(csr, job.init, job.seeds, job.beta, [128, 64, 32], [40, 40, 20])
    assert np.array_equal(mid_end, end) and np.array_equal(mid_best, best) and np.array_equal(mid_bv, bv)
    end, best, cv, bv = run_oracle_chains(csr, job.init, job.seeds, job.beta, phases, iters)
    assert np.array_equal(job.end_spins, end) and np.array_equal(job.best_spins, best)
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv)


def test_server_results_equal_one_shot(graphs):
    reqs = [
        ("G11", make_job(graphs["G11"], [31, 32], [8, 4, 2, 1], [80] * 4)),
        ("torus", make_job(graphs["torus"], [33], LADDER, [50] * 8)),
        ("ER800", make_job(graphs["ER800"], [34, 35, 36], LADDER, [20] * 8)),
        ("deg", make_job(graphs["deg"], [37], [3, 64, 1], [100, 100, 100])),
        ("G70", make_job(graphs["G70"], [38, 39], [8, 2, 1], [60, 60, 60], linear=False, t0=[90.0, 100.0])),
        ("G11b", make_job(graphs["G11"], [40], [16, 1], [300, 300])),
        # the int32 gate fails (t0 = 0.005, B = 4): int64 scores
        ("G11w", make_job(graphs["G11"], [42, 43], [8, 4, 2, 1], [80] * 4, linear=False, t0=[0.005, 0.01])),
    ]
    fps = [R.footprint_bytes(j.csr.n, j.csr.m2, j.chains, j.beta.run_end.size, len(j.phases)) for _, j in reqs]
    budget = sum(sorted(fps)[-2:])   # at most about two requests resident: queueing is exercised
    print(f"[serve] {len(reqs)} requests, budget {budget} B", flush=True)
    with R.Context(0) as ctx:
        with R.Server(ctx, budget, slice_sweeps=64) as srv:
            ids = []
            for k, j in reqs:
                ids.append(srv.submit(j))
                time.sleep(0.002)
            big = make_job(graphs["torus"], list(range(100, 400)), LADDER, [10] * 8)
            with pytest.raises(R.RfmfError) as ei:
                srv.submit(big)
            assert ei.value.status == 10
            # P above the 4096-group limit: refused at submit, before any
            # device work.
            too_many_groups = make_job(graphs["G11"], [41], [5000, 1], [10, 10])
            with pytest.raises(R.RfmfError) as ei:
                srv.submit(too_many_groups)
            assert ei.value.status == 2
            for rid in ids:
                srv.wait(rid, 120.0)
            lat = [srv.poll(rid) for rid in ids]
            st = srv.stats()
    assert all(d for d, *_ in lat)
    assert st["completed"] == len(reqs) and st["rejected"] == 2
    for (k, j) in reqs:
        same(j, one_shot(j), f"served {k}")


def test_flips_tuner_gate_and_pilots_are_ordinary_runs(gset):
    csr = gset("G11")
    sweeps, tops, pilots = 1600, [4, 16, 64], 24
    print(f"[tune-flips] G11 sweeps={sweeps} candidates={tops} pilots={pilots}", flush=True)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        for ref in (None, 16):
            rep = ctx.tune_flips(gid, sweeps, tops, pilot_chains=pilots, seed_base=100, quality_tol=0.001,
                                 reference_top_p=ref)
            # the chosen candidate is within tolerance of the floor median
            # (the reference ladder's, else the best) and the fastest such one
            med, rate = rep["median_best"], rep["chain_steps_per_s"]
            floor = med[ref] if ref else max(med.values())
            ok = [t for t in tops if med[t] >= floor - 0.001 * abs(floor)]
            assert rep["chosen_top_p"] in ok
            assert rate[rep["chosen_top_p"]] == max(rate[t] for t in ok)
        # each pilot is an ordinary run: the same seeds and ladder through the
        # batch path give the same median
        for t in tops:
            phases, iters = R.flips_ladder(t, sweeps)
            seeds = range(101, 101 + pilots)
            job = make_job(csr, seeds, phases, iters)
            ctx.batch_run([(gid, job)])
            assert float(np.median(job.best_vals)) == med[t]


def test_windows_with_per_phase_autotune_equal_one_shot(graphs):
    csr = graphs["torus"]
    job = make_job(csr, [51, 52, 53, 54], LADDER, [40] * 8)
    print("[windows] torus, bounds 50 / 170, per-phase autotune", flush=True)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ms = ctx.batch_run_windows([(gid, job)], [50, 170], autotune=True)
        geo = ctx.last_launch()
    assert len(ms) == 2 and all(m > 0 for m in ms)
    # the report describes the last launch made, not the empty final call
    assert geo["threads_per_chain"] in (128, 256, 512, 1024) and geo["wg_per_cu"] >= 1, geo
    same(job, one_shot(job), "windows")


def test_score_table_lean_and_rich_builds():
    """The score table's 4-wave tile has a lean build (6 workgroups per CU)
    for launches with more chains than the register-rich build (4 per CU)
    keeps resident, and the rich build otherwise; with few chains per CU the
    rich build also takes the graph's rows into LDS.  A chain's result is the
    same in every form, and the oracle's."""
    csr = W.torus(40, 50, seed=3)   # n = 2000, 8-bit words, +/-1 weights
    phases, iters = [128, 8, 1], [40, 40, 40]
    with R.Context(0) as ctx:
        cu = ctx.device_info()["cu"]
        big = make_job(csr, list(range(1, 6 * cu + 1)), phases, iters)
        small = make_job(csr, list(range(1, 49)), phases, iters)
        rows = make_job(csr, list(range(1, 49)), phases, iters)
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, big)], threads=256, score_table=1)
        geo_big = ctx.last_launch()
        ctx.batch_run([(gid, small)], threads=256, score_table=1, rows_in_lds=-1)
        geo_small = ctx.last_launch()
        ctx.batch_run([(gid, rows)], threads=256, score_table=1)
        geo_rows = ctx.last_launch()
    assert geo_big["score_mode"] == geo_small["score_mode"] == geo_rows["score_mode"] == "mul24-table"
    assert geo_big["wg_per_cu"] == 6 and geo_small["wg_per_cu"] == 4, (geo_big, geo_small)
    assert not geo_big["rows_in_lds"] and not geo_small["rows_in_lds"] and geo_rows["rows_in_lds"], geo_rows
    for other in (small, rows):
        for a, b in ((big.end_spins, other.end_spins), (big.best_spins, other.best_spins),
                     (big.best_vals, other.best_vals), (big.final_vals, other.final_vals)):
            assert np.array_equal(a[:48], b), "lean, rich and rows-in-LDS builds differ"
    init = small.init[:8]
    solver = small.seeds[:8]
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), 8)
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(small.end_spins[:8], end) and np.array_equal(small.best_spins[:8], best)
    assert np.array_equal(small.final_vals[:8], cv) and np.array_equal(small.best_vals[:8], bv)


@pytest.mark.parametrize("slice_sweeps", [0, 1, 3, 37])
def test_deferred_best_state_across_slices(slice_sweeps):
    """The register-rich build without rows in LDS packs an improving sweep's
    state only when the next sweep fails to improve, from the words with that
    sweep's winners set back (at the end of a slice too).  A descent at high
    beta (P = 8, then 1, most sweeps improving) run whole and in slices of 1,
    3 and 37 sweeps gives the oracle's best and final states."""
    csr = W.torus(40, 50, seed=3)   # n = 2000, 8-bit words, +/-1 weights
    phases, iters = [8, 1], [150, 250]
    seeds = list(range(21, 29))
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    beta = R.build_linear_beta_table(1.5, 3.0, sum(iters), len(seeds))
    job = R.Job(csr, init, solver, beta, phases, iters)
    print(f"[batch] deferred best state, slice={slice_sweeps}", flush=True)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], slice_sweeps=slice_sweeps, threads=256, score_table=1, rows_in_lds=-1)
        geo = ctx.last_launch()
    assert geo["wg_per_cu"] == 4 and not geo["rows_in_lds"], geo   # the register-rich build
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end) and np.array_equal(job.best_spins, best)
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv)


def test_server_admission_passes_a_blocked_request_at_most_eight_times(graphs):
    # A long request holds most of the budget; a large request behind it
    # cannot fit until it retires; ten small requests arrive after the large
    # one and fit beside the long one.  The first eight pass the large
    # request, the last two wait behind it; every result equals its one-shot
    # run.
    long_req = make_job(graphs["torus"], list(range(601, 801)), LADDER, [12500] * 8)
    large = make_job(graphs["torus"], list(range(801, 1101)), LADDER, [10] * 8)
    small = [make_job(graphs["G11"], [1200 + i], [8, 1], [25, 25]) for i in range(10)]
    fp = lambda j: R.footprint_bytes(j.csr.n, j.csr.m2, j.chains, j.beta.run_end.size, len(j.phases))
    budget = fp(long_req) + fp(large) - 1
    assert fp(long_req) + 10 * fp(small[0]) <= budget and fp(large) <= budget
    print(f"[serve-bypass] budget {budget} B: long {fp(long_req)} B, large {fp(large)} B, small {fp(small[0])} B",
          flush=True)
    with R.Context(0) as ctx:
        with R.Server(ctx, budget, slice_sweeps=500) as srv:
            ids = [srv.submit(long_req), srv.submit(large)] + [srv.submit(j) for j in small]
            for rid in ids:
                srv.wait(rid, 120.0)
            t = [srv.poll(rid) for rid in ids]
    t_long_done, t_large_admit = t[0][3], t[1][2]
    admits = [x[2] for x in t[2:]]
    print(f"[serve-bypass] long done {t_long_done:.4f}, large admitted {t_large_admit:.4f}, small admitted "
          f"{[round(a, 4) for a in admits]}", flush=True)
    assert t_large_admit >= t_long_done
    assert all(a < t_large_admit for a in admits[:8])
    assert all(a >= t_large_admit for a in admits[8:])
    for j in [long_req, large] + small:
        same(j, one_shot(j), "bypass")
