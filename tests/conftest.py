import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"
GSET = GOLDEN / "gset"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU-only run (opt-in with RFMF_SLOW=1)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("RFMF_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow test; set RFMF_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def oracle_threads() -> int:
    """Host threads for the oracle: the CPUs this process may use, at most
    OMP_NUM_THREADS when set and at most 16."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, min(16, n))


def run_oracle_chains(csr, init, solver_seeds, beta, phases, iters):
    """CPU oracle over every chain (multi-threaded); returns (end, best, cur, best_vals)."""
    from oracle import oracle as O
    end, best, cv, bv, _ = O.anneal_many(csr.offsets, csr.edge_to, csr.edge_weight, init, solver_seeds,
                                         beta.run_end, beta.run_val, beta.chain_run_off, phases, iters,
                                         threads=oracle_threads())
    return end, best, cv, bv


@pytest.fixture(scope="session")
def gset():
    import rfmf_amd as R
    cache = {}

    def load(name):
        if name not in cache:
            cache[name] = R.read_instance_from_file(str(GSET / name))[0]
        return cache[name]
    return load
