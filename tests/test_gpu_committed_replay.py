"""Replays chains of the reference's committed 80M-sweep benchmark runs on the
GPU and requires the committed best and final partitions byte-for-byte.

tests/golden/committed_chains.json holds, per chain, the seed, t0, P ladder,
sweeps per phase, the cut values and both partition strings the reference
binary printed (results/V100-*/logs/*.out); committed_digests.json holds
every chain of the two 8e7-sweep runs (V100-SXM2-10M, V100-PCIE-10M: 12
instances x 80 chains each) with SHA-256 prefixes of the packed partitions.
A full replay is 8e7 sweeps per chain, minutes of GPU time, so the tests are
opt-in: RFMF_LONG=1 with RFMF_LONG_INSTANCES (default G62; "" for none)
replays one instance's chains of the fixture with full partitions, and with
RFMF_LONG_RUNS=V100-SXM2-10M (or V100-PCIE-10M) every chain of that run, all
12 instances in one batch."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import rfmf_amd as R
from conftest import GOLDEN, GSET

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("RFMF_LONG") != "1",
                                 reason="80M-sweep replay; opt in with RFMF_LONG=1")]


def unpack(hexstr: str, n: int) -> np.ndarray:
    bits = np.unpackbits(np.frombuffer(bytes.fromhex(hexstr), np.uint8))[:n]
    return np.where(bits > 0, 1, -1).astype(np.int8)


def _env_list(name, default):
    return [x for x in os.environ.get(name, default).split(",") if x]


@pytest.mark.parametrize("instance", _env_list("RFMF_LONG_INSTANCES", "G62"))
def test_replay_committed_chains(instance):
    recs = [c for c in json.loads((GOLDEN / "committed_chains.json").read_text())["chains"]
            if c["instance"] == instance and c["run"] == "V100-SXM2-10M"]
    assert recs, instance
    csr, _ = R.read_instance_from_file(str(GSET / instance))
    seeds = [c["seed"] for c in recs]
    t0s = [c["t0"] for c in recs]
    phases, iters = recs[0]["phases"], recs[0]["iters"]
    init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
    solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
    beta = R.build_beta_run_tables(t0s, sum(iters))
    job = R.Job(csr, init, solver, beta, phases, iters)
    total = sum(iters)
    slice_sweeps = 2_000_000
    print(f"[replay] {instance}: {len(recs)} chains x {total} sweeps, n={csr.n}", flush=True)
    t0 = time.time()
    kernel_ms = 0.0
    with R.Context(0) as ctx:
        gid = ctx.upload_graph(csr)
        with R.Run(ctx, [(gid, job)]) as run:
            done, at = False, 0
            while not done:
                ms, done = run.advance(slice_sweeps)
                kernel_ms += ms
                at = min(total, at + slice_sweeps)
                print(f"[replay] {instance}: {at}/{total} sweeps, {time.time() - t0:.0f} s", flush=True)
            run.collect()
    steps = len(recs) * total / (kernel_ms / 1e3)
    print(f"[replay] {instance}: kernel {kernel_ms / 1e3:.1f} s, {steps:.3e} chain-steps/s", flush=True)
    bad = []
    for i, c in enumerate(recs):
        ok = (np.array_equal(job.best_spins[i], unpack(c["best_partition"], c["n"])) and
              np.array_equal(job.end_spins[i], unpack(c["final_partition"], c["n"])) and
              job.best_vals[i] == c["best_cut"] and job.final_vals[i] == c["final_cut"])
        if not ok:
            bad.append(c["chain"])
    print(f"[replay] {instance}: {len(recs) - len(bad)}/{len(recs)} chains byte-identical to the committed "
          f"partitions; mismatching chains: {bad}", flush=True)
    assert not bad


def partition_digest(spins: np.ndarray) -> str:
    """First 16 hex digits of SHA-256 over the partition string's bits packed
    MSB first ('1' for spin +1), as the fixture stores them."""
    return hashlib.sha256(np.packbits((spins > 0).astype(np.uint8)).tobytes()).hexdigest()[:16]


@pytest.mark.parametrize("run_name", _env_list("RFMF_LONG_RUNS", ""))
def test_replay_every_committed_chain_of_a_run(run_name):
    """Every chain of one committed 8e7-sweep run (12 Gset instances x 80
    chains) annealed together in one batch, each required equal to the
    reference's printed partitions and cut values."""
    fx = json.loads((GOLDEN / "committed_digests.json").read_text())
    fields = fx["fields"]
    insts = fx["runs"][run_name]
    jobs, recs = [], []
    for inst, rec in insts.items():
        csr, _ = R.read_instance_from_file(str(GSET / inst))
        assert csr.n == rec["n"], inst
        chains = [dict(zip(fields, c)) for c in rec["chains"]]
        seeds = [c["seed"] for c in chains]
        init = np.stack([R.generate_random_spins(s, csr.n) for s in seeds])
        solver = [(s + R.K_SOLVER_SEED_OFFSET) % 2**64 for s in seeds]
        beta = R.build_beta_run_tables([c["t0"] for c in chains], sum(rec["iters"]))
        jobs.append((csr, R.Job(csr, init, solver, beta, rec["phases"], rec["iters"])))
        recs.append((inst, chains))
    total = max(int(sum(j.iters)) for _, j in jobs)
    nchains = sum(j.chains for _, j in jobs)
    slice_sweeps = 1_000_000
    print(f"[replay-run] {run_name}: {len(jobs)} instances, {nchains} chains x {total} sweeps in one batch",
          flush=True)
    t0 = time.time()
    kernel_ms = 0.0
    with R.Context(0) as ctx:
        gids = [ctx.upload_graph(csr) for csr, _ in jobs]
        with R.Run(ctx, [(gid, j) for gid, (_, j) in zip(gids, jobs)]) as run:
            done, at = False, 0
            while not done:
                ms, done = run.advance(slice_sweeps)
                kernel_ms += ms
                at = min(total, at + slice_sweeps)
                print(f"[replay-run] {run_name}: {at}/{total} sweeps, {time.time() - t0:.0f} s", flush=True)
            run.collect()
    print(f"[replay-run] {run_name}: kernel {kernel_ms / 1e3:.1f} s, "
          f"{nchains * total / (kernel_ms / 1e3):.3e} chain-steps/s", flush=True)
    bad = []
    for (inst, chains), (_, job) in zip(recs, jobs):
        for c, r in enumerate(chains):
            ok = (partition_digest(job.best_spins[c]) == r["best_partition_sha256_16"] and
                  partition_digest(job.end_spins[c]) == r["final_partition_sha256_16"] and
                  job.best_vals[c] == r["best_cut"] and job.final_vals[c] == r["final_cut"])
            if not ok:
                bad.append(f"{inst}:{c}")
        print(f"[replay-run] {inst}: best {int(job.best_vals.max())} (reference {max(r['best_cut'] for r in chains)})",
              flush=True)
    print(f"[replay-run] {run_name}: {nchains - len(bad)}/{nchains} chains byte-identical to the committed "
          f"partitions; mismatching: {bad[:40]}", flush=True)
    assert not bad


@pytest.mark.skipif(os.environ.get("RFMF_LONG_CLI") != "1", reason="80 chains x 8e7 sweeps; RFMF_LONG_CLI=1")
def test_cli_reproduces_the_committed_g62_log():
    """The reference's committed command for G62 (results/V100-SXM2-10M/
    reproduce.sh: 80 seeds and t0s, ladder 128..1, 10^7 sweeps per phase) run
    through the rfmf_main CLI: every stdout line equals the committed G62.out
    line (rebuilt from the fixtures: the device lines, both partitions of all
    80 chains, the summary) with time-sec masked, and the stderr banner and
    per-chain kernel values equal the committed G62.err lines."""
    import re
    import subprocess
    fx = json.loads((GOLDEN / "committed_digests.json").read_text())
    rec = fx["runs"]["V100-SXM2-10M"]["G62"]
    full = {c["chain"]: c for c in json.loads((GOLDEN / "committed_chains.json").read_text())["chains"]
            if c["instance"] == "G62" and c["run"] == "V100-SXM2-10M"}
    assert sorted(full) == list(range(80))
    chains = [full[c] for c in range(80)]
    phases = ",".join(str(p) for p in rec["phases"])
    iters = ",".join(str(i) for i in rec["iters"])
    expect = []
    for c in chains:
        expect.append(f"device: GPU instance: G62 --seed {c['seed']} --phases {phases} --t0 {c['t0']:.4f} "
                      f"--iters {iters} best-cut-value: {c['best_cut']} final-cut-value: {c['final_cut']} "
                      f"time-sec: <t>")
        for key, label in (("best_partition", "best"), ("final_partition", "final")):
            bits = np.unpackbits(np.frombuffer(bytes.fromhex(c[key]), np.uint8))[:c["n"]]
            expect.append(f"Solution-{label}-partition: " + "".join("1" if b else "0" for b in bits))
    expect.append(rec["log_lines"][3])
    cli = GOLDEN.parent.parent / "rfmf_amd" / "bin" / "rfmf_main"
    args = [str(cli), str(GSET / "G62"), "--seeds", ",".join(str(c["seed"]) for c in chains),
            "--t0s", ",".join(repr(c["t0"]) for c in chains), "--phases", phases, "--iters", iters]
    print("[cli-replay] G62: the committed command, 80 chains x 8e7 sweeps", flush=True)
    # RFMF_PROGRESS makes the library report every finished launch (one per
    # phase here, about 40 s apart) on stderr; those lines are echoed as they
    # come.  stdout (two 7000-character partitions per chain, about 1.1 MB)
    # goes to a file, so no pipe fills while stderr is read.
    import tempfile
    env = dict(os.environ, RFMF_PROGRESS="1")
    with tempfile.TemporaryFile("w+") as out:
        p = subprocess.Popen(args, stdout=out, stderr=subprocess.PIPE, text=True, env=env)
        err = []
        for line in p.stderr:
            err.append(line.rstrip("\n"))
            if line.startswith("rfmf: "):
                print(f"[cli-replay] {line.rstrip()}", flush=True)
        rc = p.wait(timeout=60)
        out.seek(0)
        stdout = out.read()
    assert rc == 0, "\n".join(err[-20:])
    mask = lambda s: re.sub(r"time-sec: [0-9.]+", "time-sec: <t>", s)
    got = [mask(l) for l in stdout.splitlines()]
    want = [mask(l) for l in expect]
    bad = [i for i, (a, b) in enumerate(zip(got, want)) if a != b]
    print(f"[cli-replay] G62: {len(got)} stdout lines, {len(want)} expected, {len(bad)} differ "
          f"(first: {bad[:5]})", flush=True)
    assert len(got) == len(want) and not bad
    for line in rec["log_lines"][:3]:
        assert line in err, line
    for c, ch in enumerate(chains):
        assert f"chain {c}: kernel-best-val: {ch['kernel_best']} kernel-final-val: {ch['kernel_final']}" in err
