"""The allocation-free entry point (rfmf_anneal_tables_noalloc) against the
allocating one (rfmf_anneal_tables) on the same inputs, through the C ABI:
identical final and best spins and incremental cuts, call after call, with
the fixed device array reused by each call (a smaller call after a larger
one, word widths 8 and 16, int32 and int64 scores, and a refused call in
between)."""
import numpy as np
import pytest

import rfmf_amd as R
from rfmf_amd import workloads as W
from conftest import GSET

pytestmark = pytest.mark.gpu

LADDER = [128, 64, 32, 16, 8, 4, 2, 1]


def _inputs(csr, seeds, phases, iters, t0=None):
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    if t0 is None:
        beta = R.build_linear_beta_table(0.1, 3.0, sum(iters), len(seeds))
    else:
        beta = R.build_beta_run_tables(t0, sum(iters))
    return (csr, R.max_weighted_degree(csr), init, beta, phases, iters, solver)


def _same(args, label):
    print(f"[noalloc] {label}: n={args[0].n} chains={len(args[6])} sweeps={sum(args[5])}", flush=True)
    a = R.anneal_gpu_tables_noalloc(*args)
    b = R.anneal_gpu_tables(*args)
    assert np.array_equal(a.end_spins, b.end_spins), f"{label}: final spins differ"
    assert np.array_equal(a.best_spins, b.best_spins), f"{label}: best spins differ"
    assert np.array_equal(a.kernel_best_vals, b.kernel_best_vals), f"{label}: best cuts differ"
    assert np.array_equal(a.kernel_final_vals, b.kernel_final_vals), f"{label}: final cuts differ"
    assert a.kernel_ms > 0.0
    csr = args[0]
    for c in range(len(args[6])):
        assert R.cut_value_csr(csr, a.best_spins[c]) == a.kernel_best_vals[c]


def test_noalloc_equals_allocating_entry_point_call_after_call():
    torus = W.torus(40, 60, seed=2)                        # n = 2400, 8-bit words
    g11 = R.read_instance_from_file(str(GSET / "G11"))[0]  # n = 800, 8-bit words
    g1 = R.read_instance_from_file(str(GSET / "G1"))[0]    # n = 800, B = 67: 16-bit words
    _same(_inputs(torus, range(1, 201), LADDER, [100] * 8), "torus 40x60, 200 chains")
    # a smaller call into the same array
    _same(_inputs(g11, [42, 7, 42, 9], [8, 4, 2, 1], [150] * 4, t0=[105.0, 105.0, 90.0, 120.0]), "G11, 4 chains")
    _same(_inputs(g1, [3, 4, 5], [64, 16, 4, 1], [120] * 4, t0=[105.0, 90.0, 70.0]), "G1, 16-bit words")
    # int64 scores (the int32 gate fails at t0 = 0.01)
    _same(_inputs(g11, [15, 16], [8, 4, 2, 1], [60] * 4, t0=[0.005, 0.01]), "G11, int64 scores")


def test_noalloc_refuses_a_call_beyond_its_array_and_keeps_working():
    big = W.torus(100, 200, seed=0)
    args = _inputs(big, range(1, 4097), LADDER, [1] * 8)   # 4096 x 20000 B of spins alone exceed 64 MiB
    with pytest.raises(R.RfmfError) as ei:
        R.anneal_gpu_tables_noalloc(*args)
    assert ei.value.status == 2 and "scratch" in str(ei.value)
    # the headline shape fits: 1024 chains of the 100 x 200 torus
    _same(_inputs(big, range(1, 1025), LADDER, [250] * 8), "torus 100x200, 1024 chains")
    g11 = R.read_instance_from_file(str(GSET / "G11"))[0]
    _same(_inputs(g11, [1, 2, 3, 4, 5], [16, 1], [300, 300]), "G11 after the headline shape")
