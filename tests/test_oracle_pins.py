"""Pins the CPU oracle (oracle/rfmf_oracle.c) before it is trusted as the
parity reference: published known-answer vectors, closed forms, brute force on
tiny inputs, vectors written by the reference's own host code
(tests/golden/ref_vectors.json) and the reference's committed benchmark
partitions (tests/golden/committed_chains.json)."""
import hashlib
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN, GSET
from oracle import oracle as O

REF = json.loads((GOLDEN / "ref_vectors.json").read_text())
COMMITTED = json.loads((GOLDEN / "committed_chains.json").read_text())["chains"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- Philox / Gumbel

def test_philox_random123_known_answers():
    # Published Philox4x32-10 known-answer vectors (Random123 kat_vectors).
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_matches_reference_code():
    for v in REF["philox"]:
        assert O.philox(v["ctr"], v["key"]) == v["out"]


def test_gumbel_lut_closed_form():
    lut = O.gumbel_lut()
    closed = [int(round(-math.log(-math.log((2 * i + 1) / 32.0)) * 2**20)) for i in range(16)]
    assert lut == closed
    assert lut[15] + 127 == 3617613
    assert all(a < b for a, b in zip(lut, lut[1:]))


def test_gumbel_from_bits_all_2048():
    lut = O.gumbel_lut()
    for r in range(2048):
        g = O.gumbel_from_bits(r)
        assert g == lut[r & 15] + (r >> 4)
        assert g == REF["gumbel_bits"][r]


def test_vertex_gumbels_match_reference_code():
    for rec in REF["gumbels"]:
        got = [O.vertex_gumbel(rec["seed"], rec["step"], v) for v in range(64)]
        assert got == rec["gumbels"], (rec["seed"], rec["step"])


def test_vertex_gumbel_field_layout():
    # lane j of block b takes 16-bit field j of the 128-bit output, low 11 bits
    seed, step = 0x1234_5678_9ABC_DEF0, 77
    key = [seed & 0xFFFFFFFF, seed >> 32]
    lut = O.gumbel_lut()
    for b in (0, 1, 5, 2**32 + 3):
        out = O.philox([step & 0xFFFFFFFF, step >> 32, b & 0xFFFFFFFF, b >> 32], key)
        for j in range(8):
            word = out[j >> 1]
            r = ((word >> 16) if j & 1 else word) & 0x7FF
            assert O.vertex_gumbel(seed, step, 8 * b + j) == lut[r & 15] + (r >> 4)


def test_packed_key_order_bruteforce():
    rng = np.random.default_rng(1)
    for rec in REF["pack"]:
        assert O.pack_score_id(rec["score"], rec["id"]) == rec["key"]
    for _ in range(300):
        k = int(rng.integers(2, 7))
        scores = rng.integers(-3, 4, k).tolist()
        ids = rng.choice(50, k, replace=False).tolist()
        win_key = max(range(k), key=lambda i: O.pack_score_id(scores[i], ids[i]))
        win_lex = min(range(k), key=lambda i: (-scores[i], ids[i]))
        assert win_key == win_lex
    assert O.pack_score_id(-(2**31), 2**31 - 1) > 0   # any real candidate beats the empty key 0


# ---------------------------------------------------------------- initial spins

def test_splitmix64_published_first_output():
    assert O.splitmix64_first(0) == 0xE220A8397B1DCDAF


def test_initial_spins_match_reference_code():
    for rec in REF["spins"]:
        s = O.initial_spins(rec["seed"], rec["n"])
        assert "".join("1" if x > 0 else "0" for x in s) == rec["bits"]


def test_initial_spins_consume_reservoir_lsb_first():
    # 64 spins per xoshiro256++ output, least significant bit first.
    s = O.initial_spins(0, 128)
    state = [0]
    st = np.zeros(4, np.uint64)
    sm = 0
    words = []
    import ctypes
    for i in range(4):
        sm = (sm + 0x9E3779B97F4A7C15) % 2**64
        z = sm
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % 2**64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % 2**64
        words.append(z ^ (z >> 31))
    s0, s1, s2, s3 = words
    outs = []
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) % 2**64
    for _ in range(2):
        outs.append((rotl((s0 + s3) % 2**64, 23) + s0) % 2**64)
        t = (s1 << 17) % 2**64
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t; s3 = rotl(s3, 45)
    expect = [1 if (outs[i // 64] >> (i % 64)) & 1 else -1 for i in range(128)]
    assert s.tolist() == expect
    del state, st, ctypes


# ---------------------------------------------------------------- beta schedule

def test_beta_fixed_closed_form_and_reference():
    for rec in REF["beta_fixed"]:
        assert O.beta_fixed(rec["step"], rec["t0"]) == rec["beta"]
        assert rec["beta"] == int(math.floor(math.log(1 + rec["step"]) / rec["t0"] * 2**20 + 0.5))


def test_beta_run_tables_match_reference_code():
    for rec in REF["beta_tables"]:
        ends, vals, seen = [], [], []
        for t0 in rec["t0s"]:
            if t0 in seen:
                continue
            seen.append(t0)
            e, v = O.beta_runs_log(t0, rec["total"], cap=max(rec["n_runs"] + 2, 16))
            ends.append(e)
            vals.append(v)
        e = np.concatenate(ends)
        v = np.concatenate(vals)
        assert e.size == rec["n_runs"] and len(seen) == rec["n_tables"]
        assert sha(e.astype(np.int64)) == rec["run_end_sha"]
        assert sha(v.astype(np.int32)) == rec["run_val_sha"]


def test_beta_runs_decode_to_per_step_values():
    for t0, total in [(35.0, 5000), (0.7, 3000), (400.0, 20000)]:
        e, v = O.beta_runs_log(t0, total)
        per_step = np.repeat(v, np.diff(np.concatenate([[0], e])))
        assert per_step.size == total
        assert all(per_step[s - 1] == O.beta_fixed(s, t0) for s in range(1, total + 1, 7))


def test_linear_schedule_closed_form():
    b0, b1, T = 104858, 3145728, 1000
    e, v = O.beta_runs_linear(b0, b1, T)
    per_step = np.repeat(v, np.diff(np.concatenate([[0], e])))
    assert per_step[0] == b0 and per_step[-1] == b1
    for s in range(1, T + 1):
        assert per_step[s - 1] == b0 + ((b1 - b0) * (s - 1)) // (T - 1)


# ---------------------------------------------------------------- gate

def test_gate_matches_reference_code():
    for rec in REF["gate"]:
        assert O.score_safe_int32(rec["total"], rec["t0"], rec["B"]) == bool(rec["safe"])


def test_gate_boundary_closed_form():
    for T, t0 in [(10**6, 1.0), (80_000_000, 35.0), (1000, 0.05)]:
        bmf = int(math.floor(math.log(1 + T) / t0 * 2**20 + 0.5))
        bstar = (2**31 - 1 - 3617613 - 1) // bmf   # largest B with bmf*B + 3617613 < 2^31 - 1
        assert O.score_safe_int32(T, t0, bstar)
        assert not O.score_safe_int32(T, t0, bstar + 1)


# ---------------------------------------------------------------- graphs and cuts

@pytest.mark.parametrize("inst", sorted((p.name for p in GSET.iterdir()), key=lambda s: int(s[1:])))
def test_csr_matches_reference_code(inst):
    off, to, w, hn, hm = O.read_gset(GSET / inst)
    rec = next(r for r in REF["csr"] if r["instance"] == inst)
    assert off.size - 1 == rec["n"] and to.size == rec["m2"]
    assert sha(off) == rec["offsets_sha"] and sha(to) == rec["to_sha"] and sha(w) == rec["w_sha"]
    assert O.max_weighted_degree(off, to, w) == rec["B"]


def test_dense_ids_are_endpoint_ranks():
    # Dense id = rank of the original id among the vertices an edge touches
    # (the reference's eval_cut.py rank map).
    lines = [l for l in (GSET / "G70").read_text().splitlines() if l.strip() and not l.lstrip().startswith("#")]
    edges = [tuple(int(x) for x in l.split()) for l in lines[1:]]
    ids = sorted({u for u, _, _ in edges} | {v for _, v, _ in edges})
    rank = {o: i for i, o in enumerate(ids)}
    off, to, w, _, _ = O.read_gset(GSET / "G70")
    assert off.size - 1 == len(ids)
    for u, v, ww in edges[:500]:
        ru, rv = rank[u], rank[v]
        row = to[off[ru]:off[ru + 1]]
        assert rv in row and w[off[ru] + int(np.where(row == rv)[0][0])] == ww
        assert np.all(np.diff(row) > 0)


def _bits(hexstr, n):
    return np.unpackbits(np.frombuffer(bytes.fromhex(hexstr), np.uint8))[:n]


def test_committed_partitions_rescore_to_committed_values():
    csrs = {}
    for c in COMMITTED:
        if c["instance"] not in csrs:
            csrs[c["instance"]] = O.read_gset(GSET / c["instance"])[:3]
        off, to, w = csrs[c["instance"]]
        for key, val in (("best_partition", "best_cut"), ("final_partition", "final_cut")):
            spin = np.where(_bits(c[key], c["n"]) > 0, 1, -1).astype(np.int8)
            assert O.cut_value(off, to, w, spin) == c[val], (c["instance"], c["chain"], key)
        assert c["kernel_best"] == c["best_cut"] and c["kernel_final"] == c["final_cut"]
    for rec in REF["committed_cuts"]:
        c = next(x for x in COMMITTED if x["instance"] == rec["instance"] and x["run"] == rec["run"]
                 and x["chain"] == rec["chain"])
        assert rec["cut"] == c["best_cut" if rec["which"] == "best_partition" else "final_cut"]


def test_committed_digest_fixture_agrees_with_partitions_and_results_tables():
    # The digest fixture (every chain of both 8e7-sweep runs) against the
    # full-partition fixture and against the runs' results.csv rows: same
    # seeds and t0s in order, the table's best cut and its hit count.
    dig = json.loads((GOLDEN / "committed_digests.json").read_text())
    runs = json.loads((GOLDEN / "reference_runs.json").read_text())
    fields = dig["fields"]
    for run, insts in dig["runs"].items():
        assert len(insts) == 12, run
        for inst, rec in insts.items():
            chains = [dict(zip(fields, c)) for c in rec["chains"]]
            row = runs[run][inst]
            assert [c["seed"] for c in chains] == row["seeds"] and [c["t0"] for c in chains] == row["t0s"]
            assert rec["phases"] == row["phases"] and rec["iters"] == row["iters"]
            best = max(c["best_cut"] for c in chains)
            assert best == row["best_cut"], (run, inst)
            assert sum(c["best_cut"] == best for c in chains) == row["hits_at_best"], (run, inst)
            assert all(c["kernel_best"] == c["best_cut"] and c["kernel_final"] == c["final_cut"] for c in chains)
    for c in COMMITTED:
        d = dict(zip(fields, dig["runs"][c["run"]][c["instance"]]["chains"][c["chain"]]))
        assert d["seed"] == c["seed"] and d["best_cut"] == c["best_cut"]
        for key in ("best_partition", "final_partition"):
            packed = np.packbits(_bits(c[key], c["n"])).tobytes()
            assert hashlib.sha256(packed).hexdigest()[:16] == d[key + "_sha256_16"], (c["instance"], c["chain"])


def test_cut_bruteforce_edge_list():
    lines = [l for l in (GSET / "G11").read_text().splitlines() if l.strip()]
    edges = [tuple(int(x) for x in l.split()) for l in lines[1:]]
    ids = sorted({u for u, _, _ in edges} | {v for _, v, _ in edges})
    rank = {o: i for i, o in enumerate(ids)}
    off, to, w, _, _ = O.read_gset(GSET / "G11")
    rng = np.random.default_rng(3)
    for _ in range(5):
        spin = rng.choice([-1, 1], off.size - 1).astype(np.int8)
        brute = sum(ww for u, v, ww in edges if spin[rank[u]] != spin[rank[v]])
        assert O.cut_value(off, to, w, spin) == brute


# ---------------------------------------------------------------- the annealing engine

def _tiny_graph(rng, n, p, wmax):
    u, v = np.triu_indices(n, 1)
    keep = rng.random(u.size) < p
    u, v = u[keep], v[keep]
    w = rng.integers(1, wmax + 1, u.size) * rng.choice([-1, 1], u.size)
    if u.size == 0:
        u, v, w = np.array([0]), np.array([1]), np.array([1])
    return O.build_csr(u, v, w)[:3]


def _corr_formulation(off, to, w, init, seed, beta_per_step, phases, iters):
    """The reference kernel's formulation (src/anneal_kernels.cu:292-484),
    restated in Python for tiny graphs: every winner is picked from the
    pre-sweep state, the cut advances by sum(mDelta + mSpin*corr) with corr over
    EARLIER adjacent winners, then all flips and field updates are applied."""
    n = off.size - 1
    spin = init.astype(np.int64).copy()
    S = np.array([sum(int(w[i]) * spin[to[i]] for i in range(off[v], off[v + 1])) for v in range(n)])
    cur = sum(int(w[i]) for v in range(n) for i in range(off[v], off[v + 1]) if to[i] > v and spin[v] != spin[to[i]])
    best, best_spin, step = cur, spin.copy(), 0
    for P, it in zip(phases, iters):
        if P == 0:
            continue
        for _ in range(it):
            beta = beta_per_step[step]
            step += 1
            cand = {}
            for v in range(n):
                sc = O.vertex_gumbel(seed, step, v) + beta * int(spin[v] * S[v])
                g = v % P
                if g not in cand or sc > cand[g][0] or (sc == cand[g][0] and v < cand[g][1]):
                    cand[g] = (sc, v)
            picks = {g: cand[g][1] for g in cand}
            winner_group = {b: g for g, b in picks.items()}
            m_spin = {g: int(spin[b]) for g, b in picks.items()}
            total = 0
            for g, b in picks.items():
                corr = 0
                for i in range(off[b], off[b + 1]):
                    ga = winner_group.get(int(to[i]), -1)
                    if 0 <= ga < g:
                        corr += 2 * int(w[i]) * (-m_spin[ga])
                total += m_spin[g] * int(S[b]) + m_spin[g] * corr
            cur += total
            for g, b in picks.items():
                spin[b] = -m_spin[g]
                for i in range(off[b], off[b + 1]):
                    S[to[i]] += 2 * int(w[i]) * (-m_spin[g])
            if cur > best:
                best, best_spin = cur, spin.copy()
    return spin.astype(np.int8), best_spin.astype(np.int8), cur, best


@pytest.mark.parametrize("case", range(12))
def test_sequential_engine_equals_corr_formulation(case):
    rng = np.random.default_rng(100 + case)
    n = int(rng.integers(6, 25))
    off, to, w = _tiny_graph(rng, n, float(rng.uniform(0.15, 0.6)), int(rng.integers(1, 4)))
    n = off.size - 1
    phases = [int(x) for x in rng.choice([1, 2, 3, 5, 8, 40], 3)]
    iters = [int(x) for x in rng.integers(3, 12, 3)]
    total = sum(iters)
    t0 = float(rng.uniform(0.02, 3.0)) if case % 3 else 1e9   # t0 = 1e9: beta 0, many score ties
    e, v = O.beta_runs_log(t0, total, cap=total + 2)
    per_step = np.repeat(v, np.diff(np.concatenate([[0], e])))
    init = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int8)
    seed = int(rng.integers(0, 2**63))
    ref = _corr_formulation(off, to, w, init, seed, per_step, phases, iters)
    end, best, cur, bestv, steps = O.anneal_chain(off, to, w, init, seed, e, v, phases, iters)
    assert steps == total
    assert np.array_equal(end, ref[0]) and np.array_equal(best, ref[1])
    assert cur == ref[2] and bestv == ref[3]
    assert O.cut_value(off, to, w, end) == cur and O.cut_value(off, to, w, best) == bestv


def test_incremental_cut_equals_recompute_on_gset():
    off, to, w, _, _ = O.read_gset(GSET / "G11")
    n = off.size - 1
    e, v = O.beta_runs_log(60.0, 2000)
    for seed in (1, 2):
        init = O.initial_spins(seed, n)
        end, best, cur, bestv, _ = O.anneal_chain(off, to, w, init, seed + 17, e, v, [16, 4, 1], [700, 700, 600])
        assert O.cut_value(off, to, w, end) == cur
        assert O.cut_value(off, to, w, best) == bestv
        assert bestv >= O.cut_value(off, to, w, init)


@pytest.mark.slow
def test_cpu_replay_of_a_committed_chain():
    """80M sweeps of G62 chain 0 on one core (~45 min): the oracle alone
    reproduces the committed partitions."""
    c = next(x for x in COMMITTED if x["instance"] == "G62" and x["chain"] == 0)
    off, to, w, _, _ = O.read_gset(GSET / "G62")
    e, v = O.beta_runs_log(c["t0"], sum(c["iters"]))
    init = O.initial_spins(c["seed"], off.size - 1)
    end, best, cur, bestv, _ = O.anneal_chain(off, to, w, init, (c["seed"] + 0xC2B2AE3D27D4EB4F) % 2**64, e, v,
                                             c["phases"], c["iters"])
    assert bestv == c["best_cut"] and cur == c["final_cut"]
    assert np.array_equal((best > 0).astype(np.uint8), _bits(c["best_partition"], c["n"]))
    assert np.array_equal((end > 0).astype(np.uint8), _bits(c["final_partition"], c["n"]))
