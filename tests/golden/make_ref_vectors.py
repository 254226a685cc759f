"""Writes tests/golden/ref_vectors.json from the reference's own host code,
compiled from its sources by oracle/Makefile into oracle/_ref/ref_driver:
Philox4x32-10 outputs, per-vertex Gumbels, the 11-bit Gumbel map, packed
argmax keys, initial spins, beta run tables, int32-gate verdicts,
betaFixedForStep values, CSR digests of the Gset fixtures, and cut values of
committed partitions.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
from __future__ import annotations

import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DRIVER = HERE.parent.parent / "oracle" / "_ref" / "ref_driver"
OUT = HERE / "ref_vectors.json"


def run(*args) -> str:
    return subprocess.run([str(DRIVER), *map(str, args)], capture_output=True, text=True, check=True).stdout


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    rng = np.random.default_rng(20261203)
    out: dict = {}

    phil = []
    cases = [([0] * 4, [0, 0]), ([2**32 - 1] * 4, [2**32 - 1] * 2),
             ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])]
    for _ in range(29):
        cases.append((rng.integers(0, 2**32, 4).tolist(), rng.integers(0, 2**32, 2).tolist()))
    for ctr, key in cases:
        o = [int(x) for x in run("philox", *ctr, *key).split()]
        phil.append({"ctr": [int(x) for x in ctr], "key": [int(x) for x in key], "out": o})
    out["philox"] = phil

    gum = []
    for seed, step in [(0, 1), (123, 5), (4247410429 + 0xC2B2AE3D27D4EB4F, 1), (2**64 - 1, 2**32 + 7),
                       (987654321, 80_000_000)]:
        seed %= 2**64
        g = [int(x) for x in run("gumbels", seed, step, 64).split()]
        gum.append({"seed": seed, "step": step, "gumbels": g})
    out["gumbels"] = gum
    out["gumbel_bits"] = [int(x) for x in run("gumbelbits").split()]

    packs = []
    for s, i in [(0, 0), (-1, 0), (2**31 - 1, 5), (-(2**31), 7), (123456, 2**31 - 1), (-5, 3)]:
        packs.append({"score": s, "id": i, "key": int(run("pack", s, i))})
    out["pack"] = packs

    spins = []
    for seed in [0, 1, 42, 4247410429, 2**64 - 1, 987654321]:
        bits = run("spins", seed, 1000).strip()
        spins.append({"seed": seed, "n": 1000, "bits": bits})
    out["spins"] = spins

    betas = []
    for total, t0s in [(1000, [35.0, 0.5, 105.0, 35.0]), (100000, [35.0]), (0, [1.0]), (1, [2.0]),
                       (250000, [400.0, 0.01])]:
        lines = run("beta", total, *t0s).splitlines()
        nruns, ntab = map(int, lines[0].split())
        offs = [int(x) for x in lines[1].split()]
        ends = [int(l.split()[0]) for l in lines[2:2 + nruns]]
        vals = [int(l.split()[1]) for l in lines[2:2 + nruns]]
        betas.append({"total": total, "t0s": t0s, "n_tables": ntab, "chain_run_off": offs,
                      "run_end_sha": digest(np.array(ends, np.int64)),
                      "run_val_sha": digest(np.array(vals, np.int32)), "n_runs": nruns,
                      "head": [ends[:8], vals[:8]]})
    out["beta_tables"] = betas

    gates = []
    for total, t0, B in [(80_000_000, 35.0, 4), (80_000_000, 35.0, 70), (80_000_000, 0.01, 4),
                         (1_000_000, 1.0, 140), (1_000_000, 1.0, 147), (1_000_000, 1.0, 146), (0, 1.0, 10**9),
                         (10**12, 1e-3, 1), (150, 0.01, 67), (1200, 105.0, 67), (10, 1.0, 0)]:
        gates.append({"total": total, "t0": t0, "B": B, "safe": int(run("gate", total, t0, B))})
    out["gate"] = gates

    bf = []
    for step, t0 in [(1, 35.0), (2, 35.0), (1000, 35.0), (80_000_000, 35.0), (12345, 0.5), (2**40, 400.0)]:
        bf.append({"step": step, "t0": t0, "beta": int(run("betafix", step, t0))})
    out["beta_fixed"] = bf

    csrs = []
    for g in sorted((p.name for p in (HERE / "gset").iterdir()), key=lambda s: int(s[1:])):
        lines = run("csr", HERE / "gset" / g).splitlines()
        hdr = lines[0].split()
        off = np.array(lines[1].split(), np.uint32)
        to = np.array(lines[2].split(), np.int32)
        w = np.array(lines[3].split(), np.int32)
        csrs.append({"instance": g, "n": int(hdr[0]), "m2": int(hdr[1]), "B": int(hdr[2]),
                     "offsets_sha": digest(off), "to_sha": digest(to), "w_sha": digest(w)})
    out["csr"] = csrs

    chains = json.loads((HERE / "committed_chains.json").read_text())["chains"]
    cuts = []
    for c in chains[:3] + [c for c in chains if c["instance"] != "G62"][:6]:
        for which in ("best_partition", "final_partition"):
            bits = np.unpackbits(np.frombuffer(bytes.fromhex(c[which]), np.uint8))[:c["n"]]
            s = "".join("1" if b else "0" for b in bits)
            cuts.append({"instance": c["instance"], "run": c["run"], "chain": c["chain"], "which": which,
                         "cut": int(run("cut", HERE / "gset" / c["instance"], s))})
    out["committed_cuts"] = cuts

    OUT.write_text(json.dumps(out, indent=0))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
