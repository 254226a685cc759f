"""rfmf_amd -- MI355X-native multi-chain rejection-free multiple-flip Gumbel
annealer for MAX-CUT.

The Python layer mirrors the reference's host interface (graph loading, CSR,
cut recompute, beta schedules, the int32 gate, initial spins and
gpu3::annealGpu) over the C ABI of librfmf.so; every computation runs in that
library (the annealing itself in HIP kernels on the GPU).  There is no Python
or CPU fallback: a missing library raises on first use.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import RfmfError, check, lib

K_SOLVER_SEED_OFFSET = 0xC2B2AE3D27D4EB4F
Q20 = 1 << 20

__all__ = [
    "AnnealCsr", "ReadStats", "BetaRunTables", "LaunchConfig", "AnnealResult", "Context", "Job", "Run",
    "Server", "RfmfError", "read_instance_from_file", "graph_from_edges", "cut_value_csr",
    "max_weighted_degree", "anneal_score_safe_int32", "beta_fixed_for_step", "anneal_total_steps",
    "generate_random_spins", "build_beta_run_tables", "build_linear_beta_table", "anneal_gpu",
    "anneal_gpu_tables", "footprint_bytes", "flips_ladder", "K_SOLVER_SEED_OFFSET", "Q20",
]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class ReadStats:
    num_vertices: int
    num_edges: int
    zero_dropped: int
    float_weights: bool
    scale: float
    max_abs_weighted_degree: int


class AnnealCsr:
    """Dense CSR snapshot (rows sorted by neighbour id, edges stored twice)."""

    def __init__(self, offsets, edge_to, edge_weight, original_ids=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        self.edge_to = np.ascontiguousarray(edge_to, dtype=np.int32)
        self.edge_weight = np.ascontiguousarray(edge_weight, dtype=np.int32)
        self.n = int(self.offsets.size - 1)
        self.m2 = int(self.edge_to.size)
        self.original_ids = (np.arange(self.n, dtype=np.int64) if original_ids is None
                             else np.ascontiguousarray(original_ids, dtype=np.int64))
        self._view = _capi.Csr(_ptr(self.offsets, C.c_uint32), _ptr(self.edge_to, C.c_int32),
                               _ptr(self.edge_weight, C.c_int32), self.n, self.m2)

    def numVertices(self) -> int:
        return self.n

    def numEdges(self) -> int:
        return self.m2 // 2

    def view(self) -> _capi.Csr:
        return self._view

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))

    @classmethod
    def _from_handle(cls, handle) -> "AnnealCsr":
        v = _capi.Csr()
        check(lib().rfmf_graph_view(handle, C.byref(v)))
        n, m2 = v.n, v.m2
        off = np.ctypeslib.as_array(v.offsets, shape=(n + 1,)).copy()
        to = np.ctypeslib.as_array(v.edge_to, shape=(m2,)).copy() if m2 else np.zeros(0, np.int32)
        w = np.ctypeslib.as_array(v.edge_weight, shape=(m2,)).copy() if m2 else np.zeros(0, np.int32)
        ids_p = lib().rfmf_graph_original_ids(handle)
        ids = np.ctypeslib.as_array(ids_p, shape=(n,)).copy() if n else np.zeros(0, np.int64)
        return cls(off, to, w, ids)


def read_instance_from_file(path: str, forced_scale: float = 0.0) -> tuple[AnnealCsr, ReadStats]:
    """read_instance_from_file + removeDegreeZeroVertices + AnnealCsr."""
    h = C.c_void_p()
    st = _capi.ReadStats()
    check(lib().rfmf_graph_read_gset(str(path).encode(), float(forced_scale), C.byref(h), C.byref(st)))
    try:
        csr = AnnealCsr._from_handle(h)
    finally:
        lib().rfmf_graph_free(h)
    return csr, ReadStats(st.num_vertices, st.num_edges, st.zero_dropped, bool(st.float_weights), st.scale,
                          st.max_abs_weighted_degree)


def graph_from_edges(u, v, w) -> AnnealCsr:
    """Graph::addEdge in order (first weight wins) + degree-zero removal + AnnealCsr."""
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.int64)
    h = C.c_void_p()
    check(lib().rfmf_graph_from_edges(_ptr(u, C.c_int64), _ptr(v, C.c_int64), _ptr(w, C.c_int64), u.size,
                                      C.byref(h)))
    try:
        return AnnealCsr._from_handle(h)
    finally:
        lib().rfmf_graph_free(h)


def cut_value_csr(csr: AnnealCsr, spin) -> int:
    s = np.ascontiguousarray(spin, dtype=np.int8)
    if s.size != csr.n:
        raise ValueError("spin length != n")
    return int(lib().rfmf_cut_value(C.byref(csr.view()), _ptr(s, C.c_int8)))


def max_weighted_degree(csr: AnnealCsr) -> int:
    return int(lib().rfmf_max_weighted_degree(C.byref(csr.view())))


def anneal_score_safe_int32(total_steps: int, t0: float, delta_bound: int) -> bool:
    return bool(lib().rfmf_score_safe_int32(int(total_steps), float(t0), int(delta_bound)))


def beta_fixed_for_step(step: int, t0: float) -> int:
    return int(lib().rfmf_beta_fixed_for_step(int(step), float(t0)))


def anneal_total_steps(iters) -> int:
    it = np.ascontiguousarray(iters, dtype=np.uint64)
    out = C.c_int64()
    check(lib().rfmf_total_steps(_ptr(it, C.c_uint64), it.size, C.byref(out)))
    return out.value


def generate_random_spins(raw_seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int8)
    check(lib().rfmf_initial_spins(int(raw_seed) & (2**64 - 1), int(n), _ptr(out, C.c_int8)))
    return out


@dataclass
class BetaRunTables:
    run_end: np.ndarray      # int64, 0-based exclusive end step per run
    run_val: np.ndarray      # int32, Q20 beta per run
    chain_run_off: np.ndarray  # int64, chain -> first run of its table
    n_tables: int = 1
    _struct: _capi.BetaTables = field(default=None, repr=False)

    def struct(self) -> _capi.BetaTables:
        self.run_end = np.ascontiguousarray(self.run_end, dtype=np.int64)
        self.run_val = np.ascontiguousarray(self.run_val, dtype=np.int32)
        self.chain_run_off = np.ascontiguousarray(self.chain_run_off, dtype=np.int64)
        self._struct = _capi.BetaTables(_ptr(self.run_end, C.c_int64), _ptr(self.run_val, C.c_int32),
                                        _ptr(self.chain_run_off, C.c_int64), self.run_end.size)
        return self._struct

    def table(self, chain: int) -> tuple[np.ndarray, np.ndarray]:
        """(run_end, run_val) of one chain's table (tables are concatenated)."""
        off = int(self.chain_run_off[chain])
        later = self.chain_run_off[self.chain_run_off > off]
        end = int(later.min()) if later.size else int(self.run_end.size)
        return self.run_end[off:end], self.run_val[off:end]

    def beta_at(self, chain: int, step1: int) -> int:
        """Q20 beta of 1-based sweep `step1` for `chain`."""
        ends, vals = self.table(chain)
        r = min(int(np.searchsorted(ends, step1 - 1, side="right")), ends.size - 1)
        return int(vals[r])


def build_beta_run_tables(t0s, total_steps: int) -> BetaRunTables:
    t = np.ascontiguousarray(t0s, dtype=np.float64)
    nr = C.c_int64()
    nt = C.c_int64()
    off = np.empty(t.size, dtype=np.int64)
    check(lib().rfmf_beta_tables_log(_ptr(t, C.c_double), t.size, int(total_steps), None, None, 0,
                                     C.byref(nr), _ptr(off, C.c_int64), C.byref(nt)))
    ends = np.empty(nr.value, dtype=np.int64)
    vals = np.empty(nr.value, dtype=np.int32)
    check(lib().rfmf_beta_tables_log(_ptr(t, C.c_double), t.size, int(total_steps), _ptr(ends, C.c_int64),
                                     _ptr(vals, C.c_int32), nr.value, C.byref(nr), _ptr(off, C.c_int64),
                                     C.byref(nt)))
    return BetaRunTables(ends, vals, off, nt.value)


def build_linear_beta_table(beta0: float, beta1: float, total_steps: int, chains: int) -> BetaRunTables:
    """Linear beta0 -> beta1 (in units of 1, converted to Q20 by rounding)."""
    b0 = int(round(beta0 * Q20))
    b1 = int(round(beta1 * Q20))
    nr = C.c_int64()
    check(lib().rfmf_beta_tables_linear(b0, b1, int(total_steps), None, None, 0, C.byref(nr)))
    ends = np.empty(nr.value, dtype=np.int64)
    vals = np.empty(nr.value, dtype=np.int32)
    check(lib().rfmf_beta_tables_linear(b0, b1, int(total_steps), _ptr(ends, C.c_int64), _ptr(vals, C.c_int32),
                                        nr.value, C.byref(nr)))
    return BetaRunTables(ends, vals, np.zeros(chains, dtype=np.int64), 1)


@dataclass
class LaunchConfig:
    corr: int = 0       # 0 walk, 1 warpwalk (output-invariant)
    device: int = 0
    s_width: int = 32   # 32 / 64 (output-invariant)

    def struct(self) -> _capi.LaunchCfg:
        return _capi.LaunchCfg(self.corr, self.device, self.s_width, 0)


@dataclass
class AnnealResult:
    end_spins: np.ndarray     # C x n int8
    best_spins: np.ndarray    # C x n int8
    kernel_ms: float
    kernel_best_vals: np.ndarray
    kernel_final_vals: np.ndarray


def _prep_common(csr, initial_spins, solver_seeds, phases, iters):
    init = np.ascontiguousarray(initial_spins, dtype=np.int8).reshape(-1, csr.n) if csr.n else \
        np.zeros((len(solver_seeds), 0), np.int8)
    seeds = np.ascontiguousarray(np.asarray(solver_seeds, dtype=np.uint64))
    ph = np.ascontiguousarray(phases, dtype=np.uint64)
    it = np.ascontiguousarray(iters, dtype=np.uint64)
    C_ = seeds.size
    outs = (np.empty((C_, csr.n), np.int8), np.empty((C_, csr.n), np.int8), np.empty(C_, np.int64),
            np.empty(C_, np.int64))
    return init, seeds, ph, it, C_, outs


def anneal_gpu(csr: AnnealCsr, max_weighted_deg: int, initial_spins, t0s, phases, iters, solver_seeds,
               cfg: LaunchConfig | None = None) -> AnnealResult:
    """gpu3::annealGpu: C chains, one (t0, solver seed) each, shared phases/iters."""
    init, seeds, ph, it, C_, (end, best, bv, fv) = _prep_common(csr, initial_spins, solver_seeds, phases, iters)
    t0 = np.ascontiguousarray(t0s, dtype=np.float64)
    if t0.size != C_ or init.shape[0] != C_:
        raise ValueError("t0s, solver_seeds and initial_spins must have one entry per chain")
    ms = C.c_float()
    cs = (cfg or LaunchConfig()).struct()
    check(lib().rfmf_anneal(C.byref(csr.view()), int(max_weighted_deg), _ptr(init, C.c_int8), C_,
                            _ptr(t0, C.c_double), _ptr(ph, C.c_uint64), _ptr(it, C.c_uint64), ph.size,
                            _ptr(seeds, C.c_uint64), C.byref(cs), _ptr(end, C.c_int8), _ptr(best, C.c_int8),
                            C.byref(ms), _ptr(bv, C.c_int64), _ptr(fv, C.c_int64)))
    return AnnealResult(end, best, ms.value, bv, fv)


def anneal_gpu_tables(csr: AnnealCsr, max_weighted_deg: int, initial_spins, beta: BetaRunTables, phases, iters,
                      solver_seeds, cfg: LaunchConfig | None = None) -> AnnealResult:
    """annealGpu with explicit beta run tables (e.g. a linear schedule)."""
    init, seeds, ph, it, C_, (end, best, bv, fv) = _prep_common(csr, initial_spins, solver_seeds, phases, iters)
    bs = beta.struct()
    ms = C.c_float()
    cs = (cfg or LaunchConfig()).struct()
    check(lib().rfmf_anneal_tables(C.byref(csr.view()), int(max_weighted_deg), _ptr(init, C.c_int8), C_,
                                   C.byref(bs), _ptr(ph, C.c_uint64), _ptr(it, C.c_uint64), ph.size,
                                   _ptr(seeds, C.c_uint64), C.byref(cs), _ptr(end, C.c_int8),
                                   _ptr(best, C.c_int8), C.byref(ms), _ptr(bv, C.c_int64), _ptr(fv, C.c_int64)))
    return AnnealResult(end, best, ms.value, bv, fv)


def anneal_gpu_tables_noalloc(csr: AnnealCsr, max_weighted_deg: int, initial_spins, beta: BetaRunTables, phases,
                              iters, solver_seeds, cfg: LaunchConfig | None = None) -> AnnealResult:
    """anneal_gpu_tables through the allocation-free entry point: the same
    results, with every byte of device memory taken from the library's fixed
    per-device array (RfmfError status 2 when the call does not fit it)."""
    init, seeds, ph, it, C_, (end, best, bv, fv) = _prep_common(csr, initial_spins, solver_seeds, phases, iters)
    bs = beta.struct()
    ms = C.c_float()
    cs = (cfg or LaunchConfig()).struct()
    check(lib().rfmf_anneal_tables_noalloc(C.byref(csr.view()), int(max_weighted_deg), _ptr(init, C.c_int8), C_,
                                           C.byref(bs), _ptr(ph, C.c_uint64), _ptr(it, C.c_uint64), ph.size,
                                           _ptr(seeds, C.c_uint64), C.byref(cs), _ptr(end, C.c_int8),
                                           _ptr(best, C.c_int8), C.byref(ms), _ptr(bv, C.c_int64),
                                           _ptr(fv, C.c_int64)))
    return AnnealResult(end, best, ms.value, bv, fv)


def footprint_bytes(n: int, m2: int, chains: int, beta_runs: int, n_phases: int) -> int:
    return int(lib().rfmf_footprint_bytes(n, m2, chains, beta_runs, n_phases))


def flips_ladder(top_p: int, sweeps: int) -> tuple[list[int], list[int]]:
    """The group-count ladder top_p, top_p/2, ..., 1 with `sweeps` split into
    equal phases (rfmf_flips_ladder)."""
    n = C.c_int64()
    check(lib().rfmf_flips_ladder(int(top_p), int(sweeps), None, None, 0, C.byref(n)))
    ph = np.empty(n.value, dtype=np.uint64)
    it = np.empty(n.value, dtype=np.uint64)
    check(lib().rfmf_flips_ladder(int(top_p), int(sweeps), _ptr(ph, C.c_uint64), _ptr(it, C.c_uint64), n.value,
                                  C.byref(n)))
    return [int(x) for x in ph], [int(x) for x in it]


class Job:
    """Chains of one graph with their inputs and output buffers (rfmf_job)."""

    def __init__(self, csr: AnnealCsr, initial_spins, solver_seeds, beta: BetaRunTables, phases, iters):
        self.csr = csr
        self.init = np.ascontiguousarray(initial_spins, dtype=np.int8).reshape(-1, csr.n)
        self.seeds = np.ascontiguousarray(np.asarray(solver_seeds, dtype=np.uint64))
        self.beta = beta
        self.phases = np.ascontiguousarray(phases, dtype=np.uint64)
        self.iters = np.ascontiguousarray(iters, dtype=np.uint64)
        c = self.seeds.size
        if self.init.shape[0] != c:
            raise ValueError("initial_spins rows != chains")
        self.end_spins = np.zeros((c, csr.n), np.int8)
        self.best_spins = np.zeros((c, csr.n), np.int8)
        self.best_vals = np.zeros(c, np.int64)
        self.final_vals = np.zeros(c, np.int64)

    @property
    def chains(self) -> int:
        return int(self.seeds.size)

    def struct(self, graph_id: int = 0) -> _capi.Job:
        return _capi.Job(graph_id, 0, self.chains, _ptr(self.init, C.c_int8), _ptr(self.seeds, C.c_uint64),
                         self.beta.struct(), _ptr(self.phases, C.c_uint64), _ptr(self.iters, C.c_uint64),
                         self.phases.size, _ptr(self.end_spins, C.c_int8), _ptr(self.best_spins, C.c_int8),
                         _ptr(self.best_vals, C.c_int64), _ptr(self.final_vals, C.c_int64))


def _tuning(slice_sweeps=0, grid_cap=0, use_hip_graph=False, threads=0, autotune=False,
            score_table=0, state_in_hbm=False, rows_in_lds=0) -> _capi.Tuning:
    return _capi.Tuning(int(slice_sweeps), int(grid_cap), int(bool(use_hip_graph)), int(threads),
                        int(bool(autotune)), int(score_table), int(bool(state_in_hbm)), int(rows_in_lds))


class Context:
    """Resident device context: graphs stay in HBM between batch runs."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().rfmf_ctx_create(int(device), C.byref(self._h)))
        self._graphs: dict[int, AnnealCsr] = {}

    def close(self):
        if self._h:
            lib().rfmf_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self) -> dict:
        cu = C.c_int32()
        lds = C.c_int64()
        hbm = C.c_int64()
        name = C.create_string_buffer(64)
        check(lib().rfmf_ctx_device_info(self._h, C.byref(cu), C.byref(lds), C.byref(hbm), name, 64))
        return {"cu": cu.value, "lds_per_block": lds.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """Average ms of a plain device copy of `nbytes` (read + write)."""
        ms = C.c_float()
        check(lib().rfmf_ctx_copy_probe(self._h, int(nbytes), int(iters), C.byref(ms)))
        return ms.value

    def upload_graph(self, csr: AnnealCsr) -> int:
        gid = C.c_int32()
        check(lib().rfmf_ctx_upload_graph(self._h, C.byref(csr.view()), C.byref(gid)))
        self._graphs[gid.value] = csr
        return gid.value

    def release_graph(self, gid: int) -> None:
        check(lib().rfmf_ctx_release_graph(self._h, int(gid)))
        self._graphs.pop(gid, None)

    def batch_run(self, jobs: list[tuple[int, Job]], slice_sweeps: int = 0, grid_cap: int = 0,
                  use_hip_graph: bool = False, threads: int = 0, autotune: bool = False,
                  score_table: int = 0, state_in_hbm: bool = False, rows_in_lds: int = 0) -> float:
        """Runs (graph_id, Job) pairs to completion in packed persistent launches.
        score_table (8-bit state words): 0 auto, 1 the LDS score table whenever
        it fits, -1 never.  state_in_hbm: every chain works on its state in
        place in HBM (by default only chains too large for LDS do).
        rows_in_lds (graphs with +/-1 weights, n <= 32768): 0 auto, 1 the
        graph's rows in LDS whenever they fit, -1 never.  Results are
        identical either way."""
        arr = (_capi.Job * len(jobs))(*[j.struct(gid) for gid, j in jobs])
        tun = _tuning(slice_sweeps, grid_cap, use_hip_graph, threads, autotune, score_table, state_in_hbm,
                      rows_in_lds)
        ms = C.c_float()
        check(lib().rfmf_batch_run(self._h, arr, len(jobs), C.byref(tun), C.byref(ms)))
        return ms.value

    def last_launch(self) -> dict:
        """Geometry of the last batch launch and the autotuner's trial times."""
        t, g, b, o, m, h, r = (C.c_int32() for _ in range(7))
        trial = (C.c_float * 4)()
        check(lib().rfmf_ctx_last_launch(self._h, C.byref(t), C.byref(g), C.byref(b), C.byref(o), trial,
                                         C.byref(m), C.byref(h), C.byref(r)))
        return {"threads_per_chain": t.value, "grid": g.value, "word_bits": b.value, "wg_per_cu": o.value,
                "score_mode": ("int32", "int32-mul24", "int64", "mul24-table")[m.value],
                "state_in_hbm": bool(h.value), "rows_in_lds": bool(r.value),
                "trial_ms_per_sweep": {n: trial[i] for i, n in enumerate((128, 256, 512, 1024)) if trial[i] > 0}}

    def tune_flips(self, graph_id: int, sweeps: int, top_ps, beta=(0.1, 3.0), pilot_chains: int = 0,
                   seed_base: int = 0, quality_tol: float = 0.001, reference_top_p: int | None = None) -> dict:
        """Quality-gated choice of flips per sweep (rfmf_tune_flips): every
        candidate ladder top_p..1 anneals the same pilot chains over the full
        budget; among candidates whose median best cut is within quality_tol
        of the floor median (the reference_top_p ladder's, else the best), the
        fastest is chosen."""
        tops = np.ascontiguousarray(top_ps, dtype=np.uint64)
        ref = -1 if reference_top_p is None else [int(t) for t in tops].index(int(reference_top_p))
        spec = _capi.FlipsTune(int(sweeps), int(pilot_chains), int(seed_base), int(round(beta[0] * Q20)),
                               int(round(beta[1] * Q20)), _ptr(tops, C.c_uint64), tops.size, ref,
                               float(quality_tol))
        rep = _capi.FlipsReport()
        check(lib().rfmf_tune_flips(self._h, int(graph_id), C.byref(spec), C.byref(rep)))
        k = rep.chosen
        return {"chosen_top_p": int(tops[k]), "chosen": k,
                "median_best": {int(t): rep.median_best[i] for i, t in enumerate(tops)},
                "chain_steps_per_s": {int(t): rep.chain_steps_per_s[i] for i, t in enumerate(tops)}}

    def batch_run_windows(self, jobs: list[tuple[int, Job]], bounds: list[int], threads: int = 0,
                          autotune: bool = False) -> list[float]:
        """Runs jobs window by window ([bounds[i-1], bounds[i]) sweeps), then
        to completion; returns each window's kernel ms (all its launches)."""
        arr = (_capi.Job * len(jobs))(*[j.struct(gid) for gid, j in jobs])
        b = np.ascontiguousarray(bounds, dtype=np.int64)
        ms = (C.c_float * max(1, b.size))()
        tun = _tuning(0, 0, False, threads, autotune)
        check(lib().rfmf_batch_run_windows(self._h, arr, len(jobs), _ptr(b, C.c_int64), b.size, C.byref(tun), ms))
        return [ms[i] for i in range(b.size)]


class Run:
    """Step-wise batch: chains admitted once, advanced `sweeps` at a time."""

    def __init__(self, ctx: Context, jobs: list[tuple[int, Job]]):
        self.ctx = ctx
        self.jobs = jobs
        self._arr = (_capi.Job * len(jobs))(*[j.struct(gid) for gid, j in jobs])
        self._h = C.c_void_p()
        check(lib().rfmf_run_create(ctx.handle, self._arr, len(jobs), C.byref(self._h)))

    def advance(self, sweeps: int, use_hip_graph: bool = False, grid_cap: int = 0,
                threads: int = 0, score_table: int = 0, state_in_hbm: bool = False,
                rows_in_lds: int = 0) -> tuple[float, bool]:
        tun = _tuning(0, grid_cap, use_hip_graph, threads, False, score_table, state_in_hbm, rows_in_lds)
        ms = C.c_float()
        fin = C.c_int32()
        check(lib().rfmf_run_advance(self._h, int(sweeps), C.byref(tun), C.byref(ms), C.byref(fin)))
        return ms.value, bool(fin.value)

    def collect(self) -> None:
        check(lib().rfmf_run_collect(self._h))

    def close(self) -> None:
        if self._h:
            lib().rfmf_run_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching solve server over a Context."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, slice_sweeps: int):
        self.ctx = ctx
        self._h = C.c_void_p()
        check(lib().rfmf_server_create(ctx.handle, int(hbm_budget_bytes), int(slice_sweeps), C.byref(self._h)))
        self._keep: dict[int, Job] = {}

    def submit(self, job: Job) -> int:
        rid = C.c_int64()
        js = job.struct()
        check(lib().rfmf_server_submit(self._h, C.byref(job.csr.view()), C.byref(js), C.byref(rid)))
        self._keep[rid.value] = job
        return rid.value

    def poll(self, rid: int) -> tuple[bool, float, float, float]:
        done = C.c_int32()
        ts, ta, td = C.c_double(), C.c_double(), C.c_double()
        check(lib().rfmf_server_poll(self._h, int(rid), C.byref(done), C.byref(ts), C.byref(ta), C.byref(td)))
        return bool(done.value), ts.value, ta.value, td.value

    def wait(self, rid: int, timeout_s: float = 600.0) -> None:
        check(lib().rfmf_server_wait(self._h, int(rid), float(timeout_s)))

    def stats(self) -> dict:
        s = _capi.ServerStats()
        check(lib().rfmf_server_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _capi.ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            lib().rfmf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
