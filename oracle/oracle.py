"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the parity reference or the baseline, never as
the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_DRIVER = HERE / "_ref" / "ref_driver"

_lib = None


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE), "oracle"], check=True, capture_output=True)
        l = C.CDLL(str(LIB_PATH))
        u32, i32, i64, u64 = C.c_uint32, C.c_int32, C.c_int64, C.c_uint64
        P = C.POINTER
        l.or_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        l.or_gumbel_from_bits.argtypes = [u32]
        l.or_gumbel_from_bits.restype = i32
        l.or_gumbel_lut.argtypes = [C.c_int]
        l.or_gumbel_lut.restype = i32
        l.or_pack_score_id.argtypes = [i32, i32]
        l.or_pack_score_id.restype = u64
        l.or_vertex_gumbel.argtypes = [u64, u64, u64]
        l.or_vertex_gumbel.restype = i32
        l.or_splitmix64_next.argtypes = [P(u64)]
        l.or_splitmix64_next.restype = u64
        l.or_initial_spins.argtypes = [u64, i64, P(C.c_int8)]
        l.or_beta_fixed.argtypes = [u64, C.c_double]
        l.or_beta_fixed.restype = i32
        l.or_beta_runs_log.argtypes = [C.c_double, i64, P(i64), P(i32), i64]
        l.or_beta_runs_log.restype = i64
        l.or_beta_linear.argtypes = [u64, i64, i32, i32]
        l.or_beta_linear.restype = i32
        l.or_beta_runs_linear.argtypes = [i32, i32, i64, P(i64), P(i32), i64]
        l.or_beta_runs_linear.restype = i64
        l.or_score_safe_int32.argtypes = [i64, C.c_double, i64]
        l.or_score_safe_int32.restype = C.c_int
        l.or_max_weighted_degree.argtypes = [P(u32), P(i32), P(i32), i64]
        l.or_max_weighted_degree.restype = i64
        l.or_cut_value.argtypes = [P(u32), P(i32), P(i32), i64, P(C.c_int8)]
        l.or_cut_value.restype = i64
        l.or_build_csr.argtypes = [P(i64), P(i64), P(i64), i64, P(P(u32)), P(P(i32)), P(P(i32)), P(P(i64))]
        l.or_build_csr.restype = i64
        l.or_read_gset.argtypes = [C.c_char_p, P(P(u32)), P(P(i32)), P(P(i32)), P(i64), P(i64)]
        l.or_read_gset.restype = i64
        l.or_free.argtypes = [C.c_void_p]
        l.or_anneal_chain.argtypes = [P(u32), P(i32), P(i32), i64, P(C.c_int8), u64, P(i64), P(i32), i64,
                                      P(u64), P(u64), i64, P(C.c_int8), P(C.c_int8), C.c_void_p]
        l.or_anneal_chain.restype = C.c_int
        l.or_anneal_many.argtypes = [P(u32), P(i32), P(i32), i64, i64, P(C.c_int8), P(u64), P(i64), P(i32),
                                     P(i64), i64, P(u64), P(u64), i64, C.c_int, P(C.c_int8), P(C.c_int8),
                                     P(i64), P(i64)]
        l.or_anneal_many.restype = C.c_double
        _lib = l
    return _lib


class ChainResult(C.Structure):
    _fields_ = [("cur", C.c_int64), ("best", C.c_int64), ("steps", C.c_uint64)]


def philox(ctr, key) -> list[int]:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return [int(x) for x in o]


def gumbel_lut() -> list[int]:
    return [lib().or_gumbel_lut(i) for i in range(16)]


def gumbel_from_bits(r: int) -> int:
    return lib().or_gumbel_from_bits(r)


def vertex_gumbel(solver_seed: int, step: int, v: int) -> int:
    return lib().or_vertex_gumbel(solver_seed, step, v)


def pack_score_id(score: int, vid: int) -> int:
    return lib().or_pack_score_id(score, vid)


def splitmix64_first(seed: int) -> int:
    s = C.c_uint64(seed)
    return lib().or_splitmix64_next(C.byref(s))


def initial_spins(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int8)
    lib().or_initial_spins(seed, n, _p(out, C.c_int8))
    return out


def beta_fixed(step: int, t0: float) -> int:
    return lib().or_beta_fixed(step, t0)


def beta_runs_log(t0: float, total: int, cap: int | None = None):
    cap = cap or max(16, int(np.log1p(max(total, 1)) / t0 * (1 << 20)) + 16)
    cap = min(cap, max(total, 1) + 1)
    e = np.empty(cap, np.int64)
    v = np.empty(cap, np.int32)
    n = lib().or_beta_runs_log(t0, total, _p(e, C.c_int64), _p(v, C.c_int32), cap)
    if n < 0:
        raise RuntimeError("beta run cap exceeded")
    return e[:n].copy(), v[:n].copy()


def beta_linear(step: int, total: int, b0: int, b1: int) -> int:
    return lib().or_beta_linear(step, total, b0, b1)


def beta_runs_linear(b0: int, b1: int, total: int):
    cap = min(abs(b1 - b0) + 2, max(total, 1) + 1)
    e = np.empty(cap, np.int64)
    v = np.empty(cap, np.int32)
    n = lib().or_beta_runs_linear(b0, b1, total, _p(e, C.c_int64), _p(v, C.c_int32), cap)
    if n < 0:
        raise RuntimeError("beta run cap exceeded")
    return e[:n].copy(), v[:n].copy()


def score_safe_int32(total: int, t0: float, B: int) -> bool:
    return bool(lib().or_score_safe_int32(total, t0, B))


def max_weighted_degree(off, to, w) -> int:
    return lib().or_max_weighted_degree(_p(off, C.c_uint32), _p(to, C.c_int32), _p(w, C.c_int32), off.size - 1)


def cut_value(off, to, w, spin) -> int:
    s = np.ascontiguousarray(spin, dtype=np.int8)
    return lib().or_cut_value(_p(off, C.c_uint32), _p(to, C.c_int32), _p(w, C.c_int32), off.size - 1,
                              _p(s, C.c_int8))


def _take_csr(n, po, pt, pw, pids=None):
    m2 = po[n] if n > 0 else 0
    off = np.ctypeslib.as_array(po, shape=(n + 1,)).copy()
    to = np.ctypeslib.as_array(pt, shape=(max(m2, 1),))[:m2].copy()
    w = np.ctypeslib.as_array(pw, shape=(max(m2, 1),))[:m2].copy()
    ids = np.ctypeslib.as_array(pids, shape=(max(n, 1),))[:n].copy() if pids else None
    for p in (po, pt, pw) + ((pids,) if pids else ()):
        lib().or_free(C.cast(p, C.c_void_p))
    return off, to, w, ids


def build_csr(u, v, w):
    u = np.ascontiguousarray(u, np.int64)
    v = np.ascontiguousarray(v, np.int64)
    w = np.ascontiguousarray(w, np.int64)
    po, pt, pw = C.POINTER(C.c_uint32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    pids = C.POINTER(C.c_int64)()
    n = lib().or_build_csr(_p(u, C.c_int64), _p(v, C.c_int64), _p(w, C.c_int64), u.size, C.byref(po),
                           C.byref(pt), C.byref(pw), C.byref(pids))
    if n < 0:
        raise RuntimeError("or_build_csr failed")
    return _take_csr(n, po, pt, pw, pids)


def read_gset(path):
    po, pt, pw = C.POINTER(C.c_uint32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    hn, hm = C.c_int64(), C.c_int64()
    n = lib().or_read_gset(str(path).encode(), C.byref(po), C.byref(pt), C.byref(pw), C.byref(hn), C.byref(hm))
    if n < 0:
        raise RuntimeError(f"or_read_gset failed on {path}")
    off, to, w, _ = _take_csr(n, po, pt, pw)
    return off, to, w, hn.value, hm.value


def anneal_chain(off, to, w, init_spin, solver_seed, run_end, run_val, phases, iters):
    n = off.size - 1
    init = np.ascontiguousarray(init_spin, np.int8)
    re_ = np.ascontiguousarray(run_end, np.int64)
    rv = np.ascontiguousarray(run_val, np.int32)
    ph = np.ascontiguousarray(phases, np.uint64)
    it = np.ascontiguousarray(iters, np.uint64)
    end = np.empty(n, np.int8)
    best = np.empty(n, np.int8)
    res = ChainResult()
    rc = lib().or_anneal_chain(_p(off, C.c_uint32), _p(to, C.c_int32), _p(w, C.c_int32), n, _p(init, C.c_int8),
                               solver_seed, _p(re_, C.c_int64), _p(rv, C.c_int32), re_.size, _p(ph, C.c_uint64),
                               _p(it, C.c_uint64), ph.size, _p(end, C.c_int8), _p(best, C.c_int8), C.byref(res))
    if rc != 0:
        raise RuntimeError("or_anneal_chain failed")
    return end, best, res.cur, res.best, res.steps


def anneal_many(off, to, w, init_spins, solver_seeds, run_end, run_val, chain_run_off, phases, iters,
                threads: int = 0):
    """Returns (end, best, cur_vals, best_vals, solver_seconds)."""
    n = off.size - 1
    init = np.ascontiguousarray(init_spins, np.int8).reshape(-1, n)
    Cn = init.shape[0]
    seeds = np.ascontiguousarray(np.asarray(solver_seeds, dtype=np.uint64))
    re_ = np.ascontiguousarray(run_end, np.int64)
    rv = np.ascontiguousarray(run_val, np.int32)
    ro = np.ascontiguousarray(chain_run_off, np.int64)
    ph = np.ascontiguousarray(phases, np.uint64)
    it = np.ascontiguousarray(iters, np.uint64)
    end = np.empty((Cn, n), np.int8)
    best = np.empty((Cn, n), np.int8)
    cv = np.empty(Cn, np.int64)
    bv = np.empty(Cn, np.int64)
    sec = lib().or_anneal_many(_p(off, C.c_uint32), _p(to, C.c_int32), _p(w, C.c_int32), n, Cn,
                               _p(init, C.c_int8), _p(seeds, C.c_uint64), _p(re_, C.c_int64), _p(rv, C.c_int32),
                               _p(ro, C.c_int64), re_.size, _p(ph, C.c_uint64), _p(it, C.c_uint64), ph.size,
                               int(threads), _p(end, C.c_int8), _p(best, C.c_int8), _p(cv, C.c_int64),
                               _p(bv, C.c_int64))
    if sec < 0:
        raise RuntimeError("or_anneal_many failed")
    return end, best, cv, bv, sec


def ref_driver_available() -> bool:
    return REF_DRIVER.exists()


def ref_driver(*args: str) -> str:
    r = subprocess.run([str(REF_DRIVER), *map(str, args)], capture_output=True, text=True, check=True)
    return r.stdout
