"""Seeded synthetic MAX-CUT instances for the benchmark configurations.

Every generator is a pure function of its arguments (numpy PCG64 streams) and
returns an edge list that goes through the library's load path
(graph_from_edges: first weight wins, degree-zero removal, dense ids, sorted
rows).  The instance families:

  * erdos_renyi(n, p)          G(n, p), unit weights (Gset G1-G5 family)
  * torus(rows, cols)          2D toroidal grid, iid uniform +/-1 weights
                               (Gset G62/G81 family)
  * sparse_random(n, m, pm1)   m distinct uniform edges, unit or +/-1 weights
                               (Gset G70 / G61 family)
  * degree_random(n, 2, 6)     each vertex draws a target degree in [2, 6];
                               stubs are paired at random, self loops and
                               repeats discarded; +/-1 weights
"""
from __future__ import annotations

import math

import numpy as np

from . import AnnealCsr, graph_from_edges

MASTER_SEED = 0x5EED_2026


def _rng(*key: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([MASTER_SEED, *key])))


def _finish(u, v, w) -> AnnealCsr:
    return graph_from_edges(np.asarray(u, np.int64), np.asarray(v, np.int64), np.asarray(w, np.int64))


def erdos_renyi(n: int, p: float, seed: int = 0) -> AnnealCsr:
    r = _rng(1, n, int(p * 1e9), seed)
    iu, ju = np.triu_indices(n, k=1)
    keep = r.random(iu.size) < p
    u, v = iu[keep], ju[keep]
    return _finish(u, v, np.ones(u.size, np.int64))


def torus(rows: int, cols: int, seed: int = 0) -> AnnealCsr:
    r = _rng(2, rows, cols, seed)
    idx = np.arange(rows * cols).reshape(rows, cols)
    right = np.roll(idx, -1, axis=1)
    down = np.roll(idx, -1, axis=0)
    u = np.concatenate([idx.ravel(), idx.ravel()])
    v = np.concatenate([right.ravel(), down.ravel()])
    w = np.where(r.random(u.size) < 0.5, -1, 1)
    return _finish(u, v, w)


def sparse_random(n: int, m: int, pm1: bool, seed: int = 0) -> AnnealCsr:
    r = _rng(3, n, m, int(pm1), seed)
    seen: set[tuple[int, int]] = set()
    us, vs = [], []
    while len(us) < m:
        k = 2 * (m - len(us)) + 16
        a = r.integers(0, n, k)
        b = r.integers(0, n, k)
        for x, y in zip(a.tolist(), b.tolist()):
            if x == y:
                continue
            key = (x, y) if x < y else (y, x)
            if key in seen:
                continue
            seen.add(key)
            us.append(key[0])
            vs.append(key[1])
            if len(us) == m:
                break
    w = np.where(r.random(m) < 0.5, -1, 1) if pm1 else np.ones(m, np.int64)
    return _finish(us, vs, w)


def degree_random(n: int, dmin: int = 2, dmax: int = 6, seed: int = 0) -> AnnealCsr:
    r = _rng(4, n, dmin, dmax, seed)
    deg = r.integers(dmin, dmax + 1, n)
    stubs = np.repeat(np.arange(n), deg)
    if stubs.size % 2:
        stubs = stubs[:-1]
    r.shuffle(stubs)
    a, b = stubs[0::2], stubs[1::2]
    keep = a != b
    a, b = a[keep], b[keep]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo.astype(np.int64) * n + hi
    _, first = np.unique(key, return_index=True)
    first.sort()
    u, v = lo[first], hi[first]
    w = np.where(r.random(u.size) < 0.5, -1, 1)
    return _finish(u, v, w)


def batch_instances(count: int = 256, seed: int = 0) -> list[tuple[str, AnnealCsr]]:
    """The mixed multi-graph batch: n log-uniform in [800, 20000], family
    uniform over {ER with mean degree 6, torus +/-1 (factored close to 2:1),
    random degree 2-6 +/-1}."""
    r = _rng(5, count, seed)
    out = []
    for i in range(count):
        n = int(math.floor(math.exp(r.uniform(math.log(800), math.log(20000)))))
        fam = int(r.integers(0, 3))
        if fam == 0:
            out.append(("er", erdos_renyi_sparse(n, 6.0, seed=i)))
        elif fam == 1:
            rows = max(2, int(round(math.sqrt(n / 2))))
            cols = max(2, n // rows)
            out.append(("torus", torus(rows, cols, seed=i)))
        else:
            out.append(("deg2-6", degree_random(n, 2, 6, seed=i)))
    return out


def erdos_renyi_sparse(n: int, mean_degree: float, seed: int = 0) -> AnnealCsr:
    """G(n, p) with p = mean_degree / (n - 1), sampled as a binomial edge
    count of distinct uniform pairs (exactly the G(n, p) law), unit weights."""
    r = _rng(6, n, int(mean_degree * 1000), seed)
    pairs = n * (n - 1) // 2
    m = int(r.binomial(pairs, mean_degree / (n - 1)))
    seen: set[tuple[int, int]] = set()
    us, vs = [], []
    while len(us) < m:
        k = 2 * (m - len(us)) + 16
        a = r.integers(0, n, k)
        b = r.integers(0, n, k)
        for x, y in zip(a.tolist(), b.tolist()):
            if x == y:
                continue
            key = (x, y) if x < y else (y, x)
            if key in seen:
                continue
            seen.add(key)
            us.append(key[0])
            vs.append(key[1])
            if len(us) == m:
                break
    return _finish(us, vs, np.ones(len(us), np.int64))


def csr_digest(csr: AnnealCsr) -> str:
    import hashlib
    h = hashlib.sha256()
    for a in (csr.offsets, csr.edge_to, csr.edge_weight):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def request_stream(count: int, qps: float, sweeps_lo: int, sweeps_hi: int, seed: int = 2026,
                   graph_seed: int = 7, chains_lo: int = 32, chains_hi: int = 256):
    """A serving stream: `count` solve requests with Poisson arrivals at `qps`
    (exponential gaps), each a graph of the mixed-batch generator (n 800 to
    20000) with a uniform chain count in [chains_lo, chains_hi] and a
    log-uniform sweep count in [sweeps_lo, sweeps_hi] rounded down to a
    multiple of 8 (equal sweeps for each P of the ladder 128..1).  Returns
    ([(family, csr, chains, sweeps)], arrival seconds from the first)."""
    rng = np.random.default_rng(seed)
    graphs = batch_instances(count, seed=graph_seed)
    reqs = []
    for fam, g in graphs:
        chains = int(rng.integers(chains_lo, chains_hi + 1))
        sweeps = max(8, int(np.exp(rng.uniform(np.log(sweeps_lo), np.log(sweeps_hi)))) // 8 * 8)
        reqs.append((fam, g, chains, sweeps))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, count))
    return reqs, arrivals - arrivals[0]
