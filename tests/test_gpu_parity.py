"""HIP path vs the CPU oracle, bit-exact, through the C ABI.

Each case runs C chains on the GPU and the same chains in the oracle (the
sequential group-order restatement of the reference kernel) and requires
identical final spins, best spins, incremental final and best cuts, and
incremental cuts equal to the host recompute of the returned spins."""
import numpy as np
import pytest

import rfmf_amd as R
from rfmf_amd import workloads as W
from conftest import run_oracle_chains

pytestmark = pytest.mark.gpu


def _case(name, csr, seeds, beta, phases, iters, B=None):
    print(f"[gpu-parity] {name}: n={csr.n} m2={csr.m2} chains={len(seeds)} phases={list(phases)} "
          f"iters={list(iters)}", flush=True)
    n = csr.n
    init = np.stack([R.generate_random_spins(s, n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    B = R.max_weighted_degree(csr) if B is None else B
    res = R.anneal_gpu_tables(csr, B, init, beta, phases, iters, solver)
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    for c in range(len(seeds)):
        assert np.array_equal(res.end_spins[c], end[c]), f"{name} chain {c}: final spins differ"
        assert np.array_equal(res.best_spins[c], best[c]), f"{name} chain {c}: best spins differ"
        assert res.kernel_final_vals[c] == cv[c], f"{name} chain {c}: final cut differs"
        assert res.kernel_best_vals[c] == bv[c], f"{name} chain {c}: best cut differs"
        assert R.cut_value_csr(csr, res.end_spins[c]) == res.kernel_final_vals[c]
        assert R.cut_value_csr(csr, res.best_spins[c]) == res.kernel_best_vals[c]
    return res


def _log_tables(t0s, iters):
    return R.build_beta_run_tables(t0s, int(sum(iters)))


def test_g11_four_chains_log_schedule(gset):
    csr = gset("G11")
    phases, iters = [8, 4, 2, 1], [150] * 4
    _case("G11", csr, [42, 7, 42, 9], _log_tables([105, 105, 90, 120], iters), phases, iters)


def test_g1_dense_u16_words(gset):
    csr = gset("G1")
    phases, iters = [64, 32, 16, 8, 4, 2, 1], [120] * 7
    _case("G1", csr, [42, 43, 42], _log_tables([105, 105, 90], iters), phases, iters)


def test_g62_full_p_ladder(gset):
    csr = gset("G62")
    phases, iters = [128, 64, 32, 16, 8, 4, 2, 1], [100] * 8
    _case("G62", csr, [4247410429, 1771247888], _log_tables([35.0, 50.0], iters), phases, iters)


def test_g70_sparse(gset):
    csr = gset("G70")
    phases, iters = [8, 2, 1], [150] * 3
    _case("G70", csr, [13, 14], _log_tables([105, 105], iters), phases, iters)


def test_g61_pm1(gset):
    csr = gset("G61")
    phases, iters = [64, 16], [150, 150]
    _case("G61", csr, [9, 10], _log_tables([105, 95], iters), phases, iters)


def test_er800_linear_schedule():
    csr = W.erdos_renyi(800, 0.06, seed=1)
    phases, iters = [128, 64, 32, 16, 8, 4, 2, 1], [250] * 8
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), 8)
    _case("ER800-linear", csr, list(range(1, 9)), beta, phases, iters)


@pytest.mark.parametrize("P", [3, 5, 12, 100, 1000, 1024, 2048])
def test_general_group_counts(P):
    # odd n (tail block), weights up to +/-5, P not a power of two / above
    # the lane-reduce limit / above n: the LDS-atomic argmax path.
    rng = np.random.default_rng(P)
    n, m = 613, 2400
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    w = rng.integers(1, 6, m) * np.where(rng.random(m) < 0.5, -1, 1)
    csr = R.graph_from_edges(u[keep], v[keep], w[keep])
    phases, iters = [P, 1], [90, 30]
    _case(f"P={P}", csr, [5, 6, 7], _log_tables([40.0, 60.0, 80.0], iters), phases, iters)


@pytest.mark.parametrize("threads", [128, 512])
def test_u16_words_lane_summed_winner_update(threads):
    # weighted degree about 10^4: 16-bit words, where 3B exceeds the field's
    # range, so FLIP sums each winner's couplings over its lanes and adds them
    # in one update instead of one atomic per adjacent winner
    rng = np.random.default_rng(100)
    n, m = 509, 1800
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    w = rng.integers(1, 1200, m) * np.where(rng.random(m) < 0.5, -1, 1)
    csr = R.graph_from_edges(u[keep], v[keep], w[keep])
    B = R.max_weighted_degree(csr)
    assert 16383 // 3 < B <= 16383, B
    phases, iters = [64, 8, 1], [60, 60, 60]
    seeds = [21, 22, 23]
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    beta = _log_tables([3000.0, 5000.0, 8000.0], iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        job = R.Job(csr, init, solver, beta, phases, iters)
        ctx.batch_run([(gid, job)], threads=threads)
        geo = ctx.last_launch()
    print(f"[gpu-parity] u16 lane-summed: B={B} threads={threads} {geo}", flush=True)
    assert geo["word_bits"] == 16 and geo["threads_per_chain"] == threads, geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end) and np.array_equal(job.best_spins, best)
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv)


def test_wide_words_u32():
    # a weighted degree above 16383 forces 32-bit state words
    rng = np.random.default_rng(11)
    n = 300
    u = np.arange(n - 1)
    v = u + 1
    w = rng.integers(-3000, 3001, n - 1)
    w[w == 0] = 1
    u = np.concatenate([u, np.zeros(10, np.int64)])
    v = np.concatenate([v, np.arange(2, 12)])
    w = np.concatenate([w, np.full(10, 2000)])
    csr = R.graph_from_edges(u, v, w)
    assert R.max_weighted_degree(csr) > 16383
    phases, iters = [16, 4, 1], [60, 60, 60]
    beta = R.build_linear_beta_table(0.0, 0.02, sum(iters), 2)
    _case("u32-words", csr, [3, 4], beta, phases, iters)


def test_wide_ladder_register_argmax(gset):
    # P = 512 and 256 keep the register argmax (lanes congruent mod 64 / 32
    # share a group), next to the 128..1 steps already covered
    csr = gset("G62")
    phases, iters = [512, 256, 128, 1], [60, 60, 60, 60]
    _case("G62-P512", csr, [3, 4], _log_tables([40.0, 60.0], iters), phases, iters)


def test_p_zero_phase_skipped(gset):
    csr = gset("G11")
    phases, iters = [8, 0, 4], [100, 50, 100]
    _case("P0-phase", csr, [1, 2], _log_tables([100.0, 100.0], iters), phases, iters)


# ---------------------------------------------------------- int64 scores
# Schedules the int32 gate refuses (beta_max * B + |gumbel| max >= 2^31 - 1)
# take exact int64 scores; the reference's own isolation test runs one at
# t0 = 0.01 (tests/test_chain_isolation.sh:131).

def test_int64_scores_small_t0(gset):
    csr = gset("G11")
    phases, iters = [8, 4, 2, 1], [100] * 4
    assert not R.anneal_score_safe_int32(sum(iters), 0.005, R.max_weighted_degree(csr))
    _case("int64-G11", csr, [1, 2, 3], _log_tables([0.005, 0.01, 0.02], iters), phases, iters)


def test_int64_scores_u16_words(gset):
    csr = gset("G1")
    phases, iters = [64, 16, 4, 1], [60] * 4
    assert not R.anneal_score_safe_int32(sum(iters), 0.02, R.max_weighted_degree(csr))
    _case("int64-G1", csr, [5, 6], _log_tables([0.02, 0.05], iters), phases, iters)


@pytest.mark.parametrize("P", [3, 100, 1000])
def test_int64_scores_general_group_counts(P):
    # the two-pass argmax over LDS slots (max score, then the smallest id)
    rng = np.random.default_rng(100 + P)
    n, m = 613, 2400
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    w = rng.integers(1, 6, m) * np.where(rng.random(m) < 0.5, -1, 1)
    csr = R.graph_from_edges(u[keep], v[keep], w[keep])
    phases, iters = [P, 1], [60, 30]
    assert not R.anneal_score_safe_int32(sum(iters), 0.02, R.max_weighted_degree(csr))
    _case(f"int64-P={P}", csr, [5, 6, 7], _log_tables([0.02, 0.03, 0.05], iters), phases, iters)


def test_int64_scores_u32_words():
    rng = np.random.default_rng(12)
    n = 300
    u = np.arange(n - 1)
    v = u + 1
    w = rng.integers(-3000, 3001, n - 1)
    w[w == 0] = 1
    u = np.concatenate([u, np.zeros(10, np.int64)])
    v = np.concatenate([v, np.arange(2, 12)])
    w = np.concatenate([w, np.full(10, 2000)])
    csr = R.graph_from_edges(u, v, w)
    phases, iters = [16, 4, 1], [60, 60, 60]
    beta = R.build_linear_beta_table(0.0, 1.0, sum(iters), 2)   # beta up to 2^20: 2^20 * B > 2^31
    assert int(beta.run_val.max()) * R.max_weighted_degree(csr) > 2**31
    _case("int64-u32-words", csr, [3, 4], beta, phases, iters)


def test_int64_scores_equal_int32_scores_when_forced():
    # B passed to the drop-in call decides the gate (as annealGpu's
    # maxWeightedDeg does): a B of 2^29 forces int64 scores on a schedule the
    # int32 path handles; both paths must give identical chains.
    csr = W.erdos_renyi(800, 0.06, seed=2)
    phases, iters = [128, 32, 8, 2, 1], [80] * 5
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), 4)
    narrow = _case("int32", csr, [1, 2, 3, 4], beta, phases, iters)
    wide = _case("forced-int64", csr, [1, 2, 3, 4], beta, phases, iters, B=2**29)
    assert np.array_equal(narrow.end_spins, wide.end_spins)
    assert np.array_equal(narrow.best_spins, wide.best_spins)


# ---------------------------------------------------------- state beyond LDS
# A chain whose state words do not fit one workgroup's LDS keeps them in HBM
# and updates them in place (the reference keeps every chain's spins and
# fields in global memory, src/anneal_kernels.cu:221-261); the results must
# not change.  n = 180000 at 8-bit words, 90000 at 16-bit, 45000 at 32-bit.

def _cycle_with_chords(n, wmax, seed):
    rng = np.random.default_rng(seed)
    u = np.arange(n)
    v = (u + 1) % n
    cu = rng.integers(0, n, n // 4)
    cv = (cu + rng.integers(2, n - 1, cu.size)) % n
    u, v = np.concatenate([u, cu]), np.concatenate([v, cv])
    w = rng.integers(1, wmax + 1, u.size) * np.where(rng.random(u.size) < 0.5, -1, 1)
    return R.graph_from_edges(u, v, w)


def _hbm_case(name, csr, seeds, beta, phases, iters, bits):
    print(f"[gpu-parity] {name}: n={csr.n} m2={csr.m2} chains={len(seeds)} phases={list(phases)} "
          f"iters={list(iters)}", flush=True)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)])
        geo = ctx.last_launch()
    assert geo["state_in_hbm"] and geo["word_bits"] == bits, geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end), f"{name}: final spins differ"
    assert np.array_equal(job.best_spins, best), f"{name}: best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), f"{name}: cuts differ"
    for c in range(len(seeds)):
        assert R.cut_value_csr(csr, job.best_spins[c]) == job.best_vals[c]


def test_state_beyond_lds_u8_words():
    csr = W.torus(400, 450, seed=9)
    phases, iters = [128, 8, 1], [30, 30, 20]
    _hbm_case("torus-180k", csr, [1, 2], R.build_linear_beta_table(0.5, 3.0, sum(iters), 2), phases, iters, 8)


def test_state_beyond_lds_u16_words():
    csr = _cycle_with_chords(90000, 40, seed=13)
    assert 63 < R.max_weighted_degree(csr) <= 16383
    phases, iters = [64, 3, 1], [30, 30, 20]
    _hbm_case("cycle-90k-u16", csr, [3, 4], R.build_linear_beta_table(0.0, 0.2, sum(iters), 2), phases, iters,
              16)


def test_state_beyond_lds_u32_words():
    csr = _cycle_with_chords(45000, 9000, seed=14)
    assert R.max_weighted_degree(csr) > 16383
    phases, iters = [32, 1], [40, 40]
    _hbm_case("cycle-45k-u32", csr, [5, 6], R.build_linear_beta_table(0.0, 0.001, sum(iters), 2), phases, iters,
              32)


# ---------------------------------------------------------- 64-bit words
# A max weighted degree of 2^30 or more needs 64-bit state words (the
# reference's 64-bit S, main.cu:51); they score in int64.

def test_words_u64_in_lds():
    csr = _cycle_with_chords(500, 1 << 29, seed=15)
    assert R.max_weighted_degree(csr) >= 2**30
    phases, iters = [16, 4, 1], [60, 60, 60]
    beta = R.build_linear_beta_table(0.0, 2e-6, sum(iters), 3)
    _case("u64-words", csr, [7, 8, 9], beta, phases, iters)


def test_words_u64_in_hbm():
    csr = _cycle_with_chords(30000, 1 << 29, seed=16)
    assert R.max_weighted_degree(csr) >= 2**30
    phases, iters = [64, 1], [30, 30]
    _hbm_case("cycle-30k-u64", csr, [10, 11], R.build_linear_beta_table(0.0, 2e-6, sum(iters), 2), phases, iters,
              64)


def _weighted_torus(rows, cols, wmax, seed):
    rng = np.random.default_rng(seed)
    idx = np.arange(rows * cols).reshape(rows, cols)
    u = np.concatenate([idx.ravel(), idx.ravel()])
    v = np.concatenate([np.roll(idx, -1, axis=1).ravel(), np.roll(idx, -1, axis=0).ravel()])
    w = rng.integers(1, wmax + 1, u.size) * np.where(rng.random(u.size) < 0.5, -1, 1)
    return R.graph_from_edges(u, v, w)


@pytest.mark.parametrize("threads,score_table,hbm", [(128, 1, False), (256, 1, False), (256, -1, False),
                                                     (128, 0, True)])
def test_packed_lane_keys_switch_off_mid_run(threads, score_table, hbm):
    """A lane's running argmax packs (score, block) into one int32 key while
    |beta| * B + max|Gumbel| < 2^(31 - key bits), the key bits numbering the
    lane's blocks.  With B near 60 and 20000 vertices the linear schedule
    crosses that bound mid-run, so every chain runs sweeps of both forms."""
    csr = _weighted_torus(100, 200, 15, seed=21)
    B = R.max_weighted_degree(csr)
    assert 40 < B <= 63
    lane_blocks = -(-((csr.n + 7) // 8) // threads)
    key_bits = (lane_blocks - 1).bit_length()
    bound = 2 ** (31 - key_bits)
    lo, hi = round(0.1 * 2**20), round(3.0 * 2**20)
    assert lo * B + 3617613 < bound <= hi * B + 3617613
    seeds = [1, 2, 3]
    phases, iters = [128, 16, 1], [60, 60, 60]
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), len(seeds))
    print(f"[gpu-parity] packed keys: threads={threads} score_table={score_table} hbm={hbm} B={B}", flush=True)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], threads=threads, score_table=score_table, state_in_hbm=hbm)
        geo = ctx.last_launch()
    assert geo["threads_per_chain"] == threads and geo["word_bits"] == 8 and geo["state_in_hbm"] == hbm, geo
    assert geo["score_mode"] == ("mul24-table" if score_table == 1 else "int32-mul24"), geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end), "final spins differ"
    assert np.array_equal(job.best_spins, best), "best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), "cuts differ"


@pytest.mark.parametrize("threads,score_table", [(128, 1), (256, -1)])
def test_packed_lane_keys_ragged_tail(threads, score_table):
    """Packed per-lane keys on a chain whose last block is partial (n mod 8 =
    7) and whose full blocks do not divide evenly over the lanes."""
    csr = _cycle_with_chords(9999, 10, seed=23)
    assert csr.n % 8 == 7 and R.max_weighted_degree(csr) <= 63
    assert (csr.n // 8) // threads >= 3 and (csr.n // 8) % threads != 0
    seeds = [4, 5]
    phases, iters = [64, 8, 1], [50, 50, 50]
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), len(seeds))
    print(f"[gpu-parity] packed keys ragged: threads={threads} score_table={score_table}", flush=True)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], threads=threads, score_table=score_table)
        geo = ctx.last_launch()
    assert geo["threads_per_chain"] == threads and geo["word_bits"] == 8, geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end), "final spins differ"
    assert np.array_equal(job.best_spins, best), "best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), "cuts differ"


def test_packed_lane_keys_u32_words():
    """Packed per-lane keys over full 32-bit multiplies (32-bit words, int32
    scores without the 24-bit multiplier): B = 43893 puts the key bound at
    beta = 0.0028, so a linear schedule to 0.006 runs both forms."""
    csr = _cycle_with_chords(10003, 9000, seed=31)
    B = R.max_weighted_degree(csr)
    assert B > 16383
    seeds = [6, 7]
    phases, iters = [64, 4, 1], [40, 40, 40]
    beta = R.build_linear_beta_table(0.0, 0.006, sum(iters), len(seeds))
    print(f"[gpu-parity] packed keys u32: n={csr.n} B={B}", flush=True)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], threads=128)
        geo = ctx.last_launch()
    assert geo["threads_per_chain"] == 128 and geo["word_bits"] == 32 and geo["score_mode"] == "int32", geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end), "final spins differ"
    assert np.array_equal(job.best_spins, best), "best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), "cuts differ"


def test_packed_lane_keys_u16_words():
    """Packed per-lane keys over 16-bit words (24-bit multiplies, LUT
    scores): B = 176 puts the key bound at beta = 0.71, inside the linear
    schedule 0.1 -> 2.0, and n mod 8 = 5 leaves a partial last block."""
    csr = _cycle_with_chords(12005, 40, seed=37)
    B = R.max_weighted_degree(csr)
    assert 63 < B <= 16383
    seeds = [8, 9]
    phases, iters = [128, 16, 1], [50, 50, 50]
    beta = R.build_linear_beta_table(0.1, 2.0, sum(iters), len(seeds))
    print(f"[gpu-parity] packed keys u16: n={csr.n} B={B}", flush=True)
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], threads=128)
        geo = ctx.last_launch()
    assert geo["threads_per_chain"] == 128 and geo["word_bits"] == 16 and geo["score_mode"] == "int32-mul24", geo
    end, best, cv, bv = run_oracle_chains(csr, init, solver, beta, phases, iters)
    assert np.array_equal(job.end_spins, end), "final spins differ"
    assert np.array_equal(job.best_spins, best), "best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), "cuts differ"


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9, 17, 63])
def test_tiny_graphs(n):
    """Fewer vertices than one block, one lane or one wave: a partial first
    block, groups beyond n (P = 64 and 16 leave most groups empty), and the
    ladder down to a single group."""
    rng = np.random.default_rng(100 + n)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.size) < 0.6
    keep[0] = True
    u, v = iu[keep], ju[keep]
    w = rng.integers(1, 4, u.size) * np.where(rng.random(u.size) < 0.5, -1, 1)
    csr = R.graph_from_edges(u, v, w)
    phases, iters = [64, 16, 4, 1], [25, 25, 25, 25]
    beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), 3)
    _case(f"tiny n={csr.n}", csr, [11, 12, 13], beta, phases, iters)


# Graphs whose weights are all +1/-1 (n <= 32768) carry compact 2-byte rows;
# launches with LDS to spare (about one chain per CU) copy them and the
# offsets into LDS and walk winner rows there.  Forced on and off, every tile
# (the 2- and 4-wave tiles in their register-rich builds, which carry the
# path, 8 and 16 waves in every build), against the oracle.
_ROWS_CASES = {
    # name: (graph, phases, iters, t0s, expected score mode)
    "G62-table": ("G62", [128, 64, 8, 1], [50] * 4, [35.0, 50.0], "mul24-table"),
    "G1-u16": ("G1", [64, 16, 4, 1], [50] * 4, [105.0, 90.0], "int32-mul24"),
    "G11-int64": ("G11", [8, 4, 2, 1], [50] * 4, [0.005, 0.01], "int64"),
}


@pytest.fixture(scope="module")
def rows_oracle(gset):
    cache = {}

    def get(name):
        if name not in cache:
            g, phases, iters, t0s, _ = _ROWS_CASES[name]
            csr = gset(g)
            seeds = [3, 4]
            init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
            solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
            beta = _log_tables(t0s, iters)
            cache[name] = (csr, init, solver, beta, run_oracle_chains(csr, init, solver, beta, phases, iters))
        return cache[name]
    return get


@pytest.mark.parametrize("name", sorted(_ROWS_CASES))
@pytest.mark.parametrize("threads", [128, 256, 512, 1024])
@pytest.mark.parametrize("rows", [-1, 1])
def test_rows_in_lds_against_oracle(rows_oracle, name, threads, rows):
    _, phases, iters, _, mode = _ROWS_CASES[name]
    csr, init, solver, beta, (end, best, cv, bv) = rows_oracle(name)
    capable = threads >= 512 or mode != "int64"   # int64 scores have no register-rich build
    print(f"[gpu-parity] rows in LDS: {name} threads={threads} rows_in_lds={rows}", flush=True)
    job = R.Job(csr, init, solver, beta, phases, iters)
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        ctx.batch_run([(gid, job)], threads=threads, rows_in_lds=rows)
        geo = ctx.last_launch()
    assert geo["threads_per_chain"] == threads and geo["score_mode"] == mode, geo
    assert geo["rows_in_lds"] == (rows == 1 and capable), geo
    assert np.array_equal(job.end_spins, end), "final spins differ"
    assert np.array_equal(job.best_spins, best), "best spins differ"
    assert np.array_equal(job.final_vals, cv) and np.array_equal(job.best_vals, bv), "cuts differ"
