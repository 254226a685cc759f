"""Host side of the product library on the CPU: the C ABI's exports, the load
path (Gset reader, dedup, degree-zero removal, dense CSR), beta tables, the
int32 gate, initial spins and cut recompute against the oracle and the
reference's own host code; error codes of the drop-in entry point; the CLI's
argument handling; synthetic workload digests."""
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import rfmf_amd as R
from rfmf_amd import _capi, workloads as W
from conftest import GOLDEN, GSET, ROOT
from oracle import oracle as O

REF = json.loads((GOLDEN / "ref_vectors.json").read_text())


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "rfmf.h").read_text()
    names = set(re.findall(r"\b(rfmf_[a-z0-9_]+)\s*\(", header))
    lib = C.CDLL(str(_capi.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names <= set(_capi.SIGNATURES), sorted(names - set(_capi.SIGNATURES))
    assert R.lib().rfmf_version().decode().startswith("rfmf")


@pytest.mark.parametrize("inst", sorted((p.name for p in GSET.iterdir()), key=lambda s: int(s[1:])))
def test_csr_product_equals_oracle(inst):
    csr, st = R.read_instance_from_file(str(GSET / inst))
    off, to, w, hn, hm = O.read_gset(GSET / inst)
    assert np.array_equal(csr.offsets, off) and np.array_equal(csr.edge_to, to)
    assert np.array_equal(csr.edge_weight, w)
    assert st.num_vertices == hn and st.num_edges == hm and st.scale == 1.0
    assert R.max_weighted_degree(csr) == st.max_abs_weighted_degree == O.max_weighted_degree(off, to, w)
    assert np.all(np.diff(csr.original_ids) > 0)


def test_graph_from_edges_first_weight_wins_and_dense_ranks():
    u = [10, 30, 10, 30, 7]
    v = [30, 10, 20, 20, 10]
    w = [5, -9, 2, 3, 1]          # (30, 10) repeats (10, 30): first weight 5 wins
    csr = R.graph_from_edges(u, v, w)
    assert csr.original_ids.tolist() == [7, 10, 20, 30]
    assert csr.offsets.tolist() == [0, 1, 4, 6, 8]
    assert csr.edge_to.tolist() == [1, 0, 2, 3, 1, 3, 1, 2]
    assert csr.edge_weight.tolist() == [1, 1, 2, 5, 2, 3, 5, 3]
    off, to, ww, ids = O.build_csr(u, v, w)
    assert np.array_equal(off, csr.offsets) and np.array_equal(to, csr.edge_to)
    assert np.array_equal(ww, csr.edge_weight) and ids.tolist() == [7, 10, 20, 30]


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return p


READER_CASES = {
    "crlf_comments": "# header comment\r\n\r\n4 4\r\n1 2 1\r\n  # inner\r\n2 3 -1\r\n3 4 1\r\n4 1 2\r\n",
    "float_weights": "3 3\n1 2 0.5\n2 3 0.25\n3 1 -1.75\n",
    "zero_dropped": "4 4\n1 2 0\n2 3 1\n3 4 0.0\n1 4 3\n",
    "duplicates": "3 4\n1 2 4\n2 1 7\n2 3 1\n3 2 9\n",
    "sparse_ids": "100 3\n5 90 1\n90 17 1\n17 5 -1\n",
    "tiny_floats": "2 1\n1 2 0.000123\n",
}


@pytest.mark.parametrize("case", sorted(READER_CASES))
def test_reader_matches_reference_loader(tmp_path, case):
    p = _write(tmp_path, case, READER_CASES[case])
    csr, st = R.read_instance_from_file(str(p))
    if not O.ref_driver_available():
        pytest.skip("reference host build not present")
    lines = O.ref_driver("csr", p).splitlines()
    n, m2, B, scale, zero, maxabs = lines[0].split()
    assert csr.n == int(n) and csr.m2 == int(m2)
    assert st.scale == float(scale) and st.zero_dropped == int(zero)
    assert st.max_abs_weighted_degree == int(maxabs) == R.max_weighted_degree(csr) == int(B)
    assert csr.offsets.tolist() == [int(x) for x in lines[1].split()]
    assert csr.edge_to.tolist() == [int(x) for x in lines[2].split()] if lines[2] else csr.m2 == 0
    assert csr.edge_weight.tolist() == [int(x) for x in lines[3].split()] if lines[3] else csr.m2 == 0


def test_reader_scale_values(tmp_path):
    # maxB = 2.0 (vertex 1: 0.25 + 1.75), maxW = 1.75, maxDeg = 2: the largest
    # 10^k with 2*10^k + 1 <= INT32_MAX/4 is 10^8.
    _, st = R.read_instance_from_file(str(_write(tmp_path, "f", READER_CASES["float_weights"])))
    assert st.float_weights and st.scale == 1e8
    csr, st = R.read_instance_from_file(str(_write(tmp_path, "z", READER_CASES["zero_dropped"])))
    assert st.zero_dropped == 2 and csr.n == 4 and csr.m2 == 4


def test_reader_errors(tmp_path):
    with pytest.raises(R.RfmfError) as e:
        R.read_instance_from_file(str(_write(tmp_path, "bad", "3 2\n1 2 1\n")))
    assert e.value.status == 8
    with pytest.raises(R.RfmfError):
        R.read_instance_from_file(str(tmp_path / "missing"))


def test_spins_product_equals_oracle_and_reference():
    for rec in REF["spins"]:
        s = R.generate_random_spins(rec["seed"], rec["n"])
        assert "".join("1" if x > 0 else "0" for x in s) == rec["bits"]
        assert np.array_equal(s, O.initial_spins(rec["seed"], rec["n"]))


@pytest.mark.parametrize("t0s,total", [([35.0, 50.0, 35.0], 200_000), ([0.5], 30_000), ([400.0, 1.0], 123_457),
                                       ([2.0], 0), ([7.0], 1)])
def test_log_tables_product_equals_oracle_bruteforce(t0s, total):
    t = R.build_beta_run_tables(t0s, total)
    for c, t0 in enumerate(t0s):
        e, v = O.beta_runs_log(t0, total, cap=total + 2)
        te, tv = t.table(c)
        assert np.array_equal(te, e) and np.array_equal(tv, v)


def test_log_tables_equal_reference_digests():
    import hashlib
    for rec in REF["beta_tables"]:
        t = R.build_beta_run_tables(rec["t0s"], rec["total"])
        assert t.run_end.size == rec["n_runs"] and t.n_tables == rec["n_tables"]
        assert t.chain_run_off.tolist() == rec["chain_run_off"]
        assert hashlib.sha256(t.run_end.tobytes()).hexdigest() == rec["run_end_sha"]
        assert hashlib.sha256(t.run_val.tobytes()).hexdigest() == rec["run_val_sha"]


def test_linear_table_product_equals_oracle():
    for b0, b1, T in [(0.1, 3.0, 10_000), (0.1, 3.0, 2_000_000), (1.0, 1.0, 50), (2.0, 0.5, 777)]:
        t = R.build_linear_beta_table(b0, b1, T, 3)
        e, v = O.beta_runs_linear(int(round(b0 * 2**20)), int(round(b1 * 2**20)), T)
        assert np.array_equal(t.run_end, e) and np.array_equal(t.run_val, v)


def test_gate_and_beta_fixed_product_equals_oracle():
    rng = np.random.default_rng(5)
    for rec in REF["gate"]:
        assert R.anneal_score_safe_int32(rec["total"], rec["t0"], rec["B"]) == bool(rec["safe"])
    for _ in range(500):
        T = int(10 ** rng.uniform(0, 12))
        t0 = float(10 ** rng.uniform(-3, 3))
        B = int(10 ** rng.uniform(0, 9))
        assert R.anneal_score_safe_int32(T, t0, B) == O.score_safe_int32(T, t0, B)
        s = int(rng.integers(1, 2**40))
        assert R.beta_fixed_for_step(s, t0) == O.beta_fixed(s, t0)


def test_cut_product_equals_oracle():
    csr, _ = R.read_instance_from_file(str(GSET / "G61"))
    rng = np.random.default_rng(9)
    for _ in range(5):
        s = rng.choice([-1, 1], csr.n).astype(np.int8)
        assert R.cut_value_csr(csr, s) == O.cut_value(csr.offsets, csr.edge_to, csr.edge_weight, s)


def test_total_steps_cap():
    assert R.anneal_total_steps([1, 2, 3]) == 6
    with pytest.raises(R.RfmfError) as e:
        R.anneal_total_steps([10**12, 1])
    assert e.value.status == 7


# ---------------------------------------------------------------- drop-in error behaviour (no device needed)

def _g11():
    return R.read_instance_from_file(str(GSET / "G11"))[0]


def test_anneal_rejects_bad_inputs_before_device_work():
    csr = _g11()
    init = np.ones((2, csr.n), np.int8)
    seeds = [1, 2]
    with pytest.raises(R.RfmfError) as e:
        R.anneal_gpu(csr, 4, init, [1.0, 0.0], [8], [10], seeds)
    assert e.value.status == 1                         # t0 <= 0
    with pytest.raises(R.RfmfError) as e:
        R.anneal_gpu(csr, 4, init, [1.0, 2.0], [8, 1], [10**12, 1], seeds)
    assert e.value.status == 7                         # total-sweep cap
    # A schedule the int32 gate refuses is accepted (int64 scores): on a host
    # without a GPU the only error left is the missing device.
    assert not R.anneal_score_safe_int32(100, 0.001, 4)
    with pytest.raises(R.RfmfError) as e:
        R.anneal_gpu(csr, 4, init, [0.001, 2.0], [8], [100], seeds)
    assert e.value.status == 4


def test_noalloc_entry_point_refuses_calls_beyond_its_array_before_device_work():
    # 4096 chains of the 100 x 200 torus need 82 MB of initial spins alone,
    # more than the allocation-free entry point's fixed 64 MiB array: refused
    # with RFMF_ERR_UNSUPPORTED before any device call (this runs without a GPU).
    csr = W.torus(100, 200, seed=0)
    C_ = 4096
    init = np.ones((C_, csr.n), np.int8)
    beta = R.build_linear_beta_table(0.1, 3.0, 8, C_)
    with pytest.raises(R.RfmfError) as ei:
        R.anneal_gpu_tables_noalloc(csr, R.max_weighted_degree(csr), init, beta, [8, 1], [4, 4],
                                    list(range(1, C_ + 1)))
    assert ei.value.status == 2 and "scratch" in str(ei.value)
    # invalid inputs are refused by the same checks as the allocating path
    with pytest.raises(R.RfmfError) as ei:
        R.anneal_gpu_tables_noalloc(csr, 4, init[:2] * 0, R.build_linear_beta_table(0.1, 3.0, 8, 2), [8, 1],
                                    [4, 4], [1, 2])
    assert ei.value.status == 1 and "spins" in str(ei.value)


def test_anneal_passthrough_without_phases():
    csr = _g11()
    init = np.stack([R.generate_random_spins(s, csr.n) for s in (3, 4)])
    r = R.anneal_gpu(csr, 4, init, [1.0, 2.0], [], [], [5, 6])
    assert np.array_equal(r.end_spins, init) and np.array_equal(r.best_spins, init)
    assert r.kernel_ms == 0.0
    assert [int(x) for x in r.kernel_best_vals] == [R.cut_value_csr(csr, init[c]) for c in range(2)]


def test_anneal_passthrough_on_an_empty_graph(tmp_path):
    """A graph left with no vertices (every edge weight zero, so every vertex
    has degree zero and is removed) anneals to a pass-through, as
    src/anneal_kernels.cu:856-867 returns it: no device work, zero time, the
    empty partitions and cuts of 0."""
    p = tmp_path / "zero_weights"
    p.write_text("3 2\n1 2 0\n2 3 0\n")
    csr, st = R.read_instance_from_file(str(p))
    assert (st.num_vertices, st.zero_dropped, csr.n, csr.m2) == (3, 2, 0, 0)
    init = np.zeros((2, 0), dtype=np.int8)
    r = R.anneal_gpu(csr, R.max_weighted_degree(csr), init, [1.0, 2.0], [8, 1], [4, 4], [5, 6])
    assert r.end_spins.shape == (2, 0) and r.best_spins.shape == (2, 0)
    assert r.kernel_ms == 0.0
    assert [int(x) for x in r.kernel_best_vals] == [0, 0] and [int(x) for x in r.kernel_final_vals] == [0, 0]


def test_context_without_device_reports_no_device():
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(R.RfmfError) as e:
        R.Context(0)
    assert e.value.status == 4


# ---------------------------------------------------------------- CLI

CLI = ROOT / "rfmf_amd" / "bin" / "rfmf_main"


@pytest.mark.parametrize("args,needle", [
    ([], "Usage"),
    (["x", "--seed", "1"], "one (seed, t0) pair"),
    (["x", "--seeds", "1,2", "--t0s", "1", "--phases", "1", "--iters", "1"], "--seeds has 2 entries but --t0s has 1"),
    (["x", "--seeds", "1", "--t0s", "0", "--phases", "1", "--iters", "1"], "must all be positive"),
    (["x", "--seeds", "1", "--t0s", "1", "--phases", "1,2", "--iters", "1"], "--iters has 1 entries but --phases has 2"),
    (["x", "--bogus", "1"], "unrecognized option"),
    (["x", "--seeds", "1"], "--t0s missing (required)"),
])
def test_cli_argument_errors(args, needle):
    r = subprocess.run([str(CLI), *args], capture_output=True, text=True)
    assert r.returncode == 1 and needle in r.stderr


# ---------------------------------------------------------------- workloads

def test_workload_digests_are_stable():
    assert W.csr_digest(W.torus(100, 200, seed=0)) == "52f133f8b6e30961"
    er = W.erdos_renyi(800, 0.06, seed=0)
    assert er.n == 800 and abs(er.m2 / 2 - 0.06 * 800 * 799 / 2) < 600
    sp = W.sparse_random(10000, 10000, False, seed=0)
    assert sp.m2 == 20000 and 8000 < sp.n < 9300
    pm = W.sparse_random(7000, 17150, True, seed=0)
    assert pm.m2 == 34300 and set(np.unique(pm.edge_weight)) == {-1, 1}
    dg = W.degree_random(2000, 2, 6, seed=0)
    assert dg.degrees().max() <= 6


def test_footprint_grows_with_chains():
    a = R.footprint_bytes(20000, 80000, 64, 1000, 8)
    b = R.footprint_bytes(20000, 80000, 128, 1000, 8)
    assert b > a > 20000 * 64


def test_flips_ladder_halves_to_one_and_splits_sweeps_evenly():
    assert R.flips_ladder(128, 1000) == ([128, 64, 32, 16, 8, 4, 2, 1], [125] * 8)
    assert R.flips_ladder(1, 7) == ([1], [7])
    ph, it = R.flips_ladder(16, 1003)
    assert ph == [16, 8, 4, 2, 1] and it == [201, 201, 201, 200, 200] and sum(it) == 1003
    for bad, status in ((3, 1), (0, 1), (8192, 2)):
        with pytest.raises(R.RfmfError) as e:
            R.flips_ladder(bad, 10)
        assert e.value.status == status


# C struct name -> ctypes mirror used by the Python package.
_MIRRORS = {"rfmf_csr": "Csr", "rfmf_read_stats": "ReadStats", "rfmf_beta_tables": "BetaTables",
            "rfmf_launch_cfg": "LaunchCfg", "rfmf_job": "Job", "rfmf_tuning": "Tuning",
            "rfmf_flips_tune": "FlipsTune", "rfmf_flips_report": "FlipsReport",
            "rfmf_server_stats": "ServerStats"}


def test_header_is_c_and_struct_layouts_match_ctypes(tmp_path):
    # The boundary is a plain C header: it compiles as C11 on its own, and
    # every struct's size and field offsets equal the ctypes mirror's.
    import shutil
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "rfmf.h"', 'int main(void) {']
    for cname, pyname in _MIRRORS.items():
        mirror = getattr(_capi, pyname)
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in mirror._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run([cc, "-std=c11", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, what, val = line.split()
        got[(cname, what)] = int(val)
    for cname, pyname in _MIRRORS.items():
        mirror = getattr(_capi, pyname)
        assert got[(cname, "size")] == C.sizeof(mirror), cname
        for field, _ in mirror._fields_:
            assert got[(cname, field)] == getattr(mirror, field).offset, f"{cname}.{field}"


def test_group_map_reciprocal_needs_one_correction():
    """The kernel's x mod P for P not a power of two: q = mulhi(x, m) with
    m = floor((2^32 - 1) / P) is floor(x / P) or one less for every x < 2^31,
    so r = x - q * P needs at most one subtraction of P (anneal_slice.hpp,
    modBy).  Checked for every P up to 4096 on random and edge values of x."""
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.integers(0, 2**31, 20000, dtype=np.int64),
                         np.array([0, 1, 2**31 - 1, 2**31 - 2, 2**24, 2**16 - 1], dtype=np.int64)])
    xs = xs.astype(object)
    for p in range(3, 4097):
        if p & (p - 1) == 0:
            continue
        m = (2**32 - 1) // p
        for x in xs[:: max(1, p // 64)] if p > 64 else xs:
            q = (x * m) >> 32
            r = x - q * p
            assert 0 <= r < 2 * p, (x, p)
            assert (r - p if r >= p else r) == x % p, (x, p)


def test_packed_lane_key_order_and_decode():
    """The scoring loop's per-lane key (anneal_slice.hpp, scorePackedLanes):
    with the table variant's score t = beta * (delta + 64) + gumbel, the key
    ((t << k) + (K - 1 - i) - ((64 * beta) << k)) mod 2^32, read as int32,
    equals s * 2^k + (K - 1 - i) for the true score s whenever
    |beta| * B + max|gumbel| < 2^(31 - k); its maximum over a lane's blocks
    is the largest score, ties to the earliest block; and decoding recovers
    s and i."""
    rng = np.random.default_rng(11)
    gmin, gmax = -1303301, 3617613
    for k in range(0, 8):
        K = 1 << k
        for _ in range(300):
            B = int(rng.integers(1, 64))
            lim = (2 ** (31 - k) - gmax - 1) // B
            beta = int(rng.integers(0, min(lim, 2**23 - 1) + 1))
            items = []
            for i in range(K):
                delta = int(rng.integers(-B, B + 1))
                g = int(rng.integers(gmin, gmax + 1))
                s = beta * delta + g
                assert abs(s) < 2 ** (31 - k)
                t = beta * (delta + 64) + g
                key = (((t << k) + (K - 1 - i) - ((64 * beta) << k)) % 2**32)
                key = key - 2**32 if key >= 2**31 else key
                assert key == s * 2**k + (K - 1 - i)
                assert -(2**31) < key < 2**31
                assert key >> k == s and K - 1 - (key & (K - 1)) == i
                items.append((s, i, key))
            best = max(items, key=lambda e: e[2])
            want = max(items, key=lambda e: (e[0], -e[1]))
            assert best[:2] == want[:2]


def test_lane_keys_fold_as_64_bit_keys():
    """The cross-lane fold (anneal_slice.hpp, the packed branch of select and
    RESOLVE): each lane slot's int32 key s * 2^k + (K - 1 - i) becomes the u64
    key ((key ^ 2^31) << 32) | ~(8 tid + j).  Over the slots of every lane of a
    group, the unsigned maximum is the vertex of largest score with the
    smallest id, where vertex = 8 (tid + i * threads) + j, and RESOLVE's
    decode ~lo + ((K - 1 - (hi & (2^k - 1))) << (3 + log2 threads)) returns
    that vertex.  Scores are drawn from a narrow range so that ties across
    lanes, slots and blocks are common."""
    rng = np.random.default_rng(12)
    for lg_nt in (7, 8, 9, 10):
        nt = 1 << lg_nt
        for K in (1, 2, 3, 5, 10):
            k = 0 if K <= 1 else (K - 1).bit_length()
            for _ in range(40):
                cands = []
                for _c in range(64):
                    tid = int(rng.integers(0, nt))
                    j = int(rng.integers(0, 8))
                    i = int(rng.integers(0, K))
                    s = int(rng.integers(-3, 4)) * 1000 + int(rng.integers(-2, 3))
                    key32 = s * 2**k + (K - 1 - i)
                    hi = (key32 % 2**32) ^ 0x80000000
                    lo = ~(8 * tid + j) & 0xFFFFFFFF
                    vertex = 8 * (tid + i * nt) + j
                    cands.append((s, vertex, (hi << 32) | lo))
                key = max(c[2] for c in cands)
                s_want, v_want = max((c[0], -c[1]) for c in cands)
                hi, lo = key >> 32, key & 0xFFFFFFFF
                mask = (1 << k) - 1
                v = (~lo & 0xFFFFFFFF) + ((K - 1 - (hi & mask)) << (3 + lg_nt))
                assert v == -v_want, (nt, K, key)
                assert hi != 0   # no real key is the empty slot's 0


def test_spin_byte_gathers_the_eight_spin_bits():
    """The best-spin snapshot's gather (anneal_slice.hpp, Block8::spinByte):
    for 8-bit words s = ((d0 >> 7) & 0x01010101) | ((d1 >> 3) & 0x10101010)
    and s | s >> 7 | s >> 14 | s >> 21, for 16-bit words s = OR over dwords q
    of (dq >> (15 - 2q)) & (0x10001 << 2q) and s | s >> 15: the low byte is
    the spin (top) bit of word j at bit j, for random words and the all-set
    and all-clear cases."""
    rng = np.random.default_rng(5)
    d = rng.integers(0, 2**32, (20000, 4), dtype=np.uint64)
    d = np.concatenate([d, np.zeros((1, 4), np.uint64), np.full((1, 4), 2**32 - 1, np.uint64)])
    s8 = ((d[:, 0] >> 7) & 0x01010101) | ((d[:, 1] >> 3) & 0x10101010)
    got8 = (s8 | (s8 >> 7) | (s8 >> 14) | (s8 >> 21)) & 0xFF
    s16 = np.zeros(len(d), np.uint64)
    for q in range(4):
        s16 |= (d[:, q] >> np.uint64(15 - 2 * q)) & np.uint64(0x00010001 << (2 * q))
    got16 = (s16 | (s16 >> 15)) & 0xFF
    want8 = np.zeros(len(d), np.uint64)
    want16 = np.zeros(len(d), np.uint64)
    for j in range(8):
        want8 |= ((d[:, j // 4] >> np.uint64(8 * (j % 4) + 7)) & 1) << np.uint64(j)
        want16 |= ((d[:, j // 2] >> np.uint64(16 * (j % 2) + 15)) & 1) << np.uint64(j)
    assert np.array_equal(got8, want8)
    assert np.array_equal(got16, want16)
    assert got8[-1] == 0xFF and got8[-2] == 0 and got16[-1] == 0xFF and got16[-2] == 0
