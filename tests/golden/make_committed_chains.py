"""Writes tests/golden/committed_chains.json and committed_digests.json:
per-chain end states taken from the reference's committed benchmark logs
(results/<run>/logs/<G>.out, .err).

Each record holds the chain's resolved seed, t0, phases, iters, the best and
final cut values printed by the reference binary (host recompute), the
kernel's incremental values (stderr), and both partition strings packed as hex
(np.packbits of the 0/1 string, MSB first).  These are data, not code: the
reference binary printed them.

Run in the build container (reads /root/reference):
    python tests/golden/make_committed_chains.py
"""
from __future__ import annotations

import hashlib
import json
import re
from pathlib import Path

import numpy as np

REF = Path("/root/reference/results")
OUT = Path(__file__).resolve().parent / "committed_chains.json"
DIGESTS = Path(__file__).resolve().parent / "committed_digests.json"
DIGEST_RUNS = ["V100-SXM2-10M", "V100-PCIE-10M"]

# (run, instance, chain indices): every chain of G62 at 80M sweeps, and a few
# chains of the other families.
SELECT = [
    ("V100-SXM2-10M", "G62", list(range(80))),
    ("V100-SXM2-10M", "G70", [0, 1, 2, 35]),
    ("V100-SXM2-10M", "G61", [0, 1, 23]),
    ("V100-SXM2-10M", "G81", [0, 1, 71]),
    ("V100-PCIE-10M", "G62", [0, 1]),
]

DEV = re.compile(r"^device: GPU instance: (\S+) --seed (\d+) --phases (\S+) --t0 (\S+) --iters (\S+) "
                 r"best-cut-value: (-?\d+) final-cut-value: (-?\d+) time-sec: (\S+)$")
KER = re.compile(r"^chain (\d+): kernel-best-val: (-?\d+) kernel-final-val: (-?\d+)$")


def pack(bits: str) -> str:
    a = np.frombuffer(bits.encode(), dtype=np.uint8) - ord("0")
    return np.packbits(a).tobytes().hex()


def short_digest(packed_hex: str) -> str:
    return hashlib.sha256(bytes.fromhex(packed_hex)).hexdigest()[:16]


def parse(run: str, inst: str):
    out_lines = (REF / run / "logs" / f"{inst}.out").read_text().splitlines()
    err_lines = (REF / run / "logs" / f"{inst}.err").read_text().splitlines()
    kern = {}
    for l in err_lines:
        m = KER.match(l.strip())
        if m:
            kern[int(m.group(1))] = (int(m.group(2)), int(m.group(3)))
    chains = []
    i = 0
    while i < len(out_lines):
        m = DEV.match(out_lines[i].strip())
        if m:
            best = out_lines[i + 1].split(": ", 1)[1].strip()
            final = out_lines[i + 2].split(": ", 1)[1].strip()
            chains.append(dict(instance=m.group(1), seed=int(m.group(2)),
                               phases=[int(x) for x in m.group(3).split(",")], t0=float(m.group(4)),
                               iters=[int(x) for x in m.group(5).split(",")], best_cut=int(m.group(6)),
                               final_cut=int(m.group(7)), n=len(best), best_partition=pack(best),
                               final_partition=pack(final)))
            i += 3
        else:
            i += 1
    for c, rec in enumerate(chains):
        rec["chain"] = c
        rec["kernel_best"], rec["kernel_final"] = kern[c]
    return chains


def main():
    recs = []
    for run, inst, idx in SELECT:
        chains = parse(run, inst)
        for c in idx:
            r = dict(chains[c])
            r["run"] = run
            recs.append(r)
    OUT.write_text(json.dumps({"source": "reference results/<run>/logs/<instance>.{out,err}",
                               "chains": recs}, indent=0))
    print(f"wrote {len(recs)} chains to {OUT} ({OUT.stat().st_size} bytes)")
    runs = {}
    for run in DIGEST_RUNS:
        insts = sorted((p.stem for p in (REF / run / "logs").glob("*.out")), key=lambda s: int(s[1:]))
        runs[run] = {}
        for inst in insts:
            chains = parse(run, inst)
            err = (REF / run / "logs" / f"{inst}.err").read_text().splitlines()
            out = (REF / run / "logs" / f"{inst}.out").read_text().splitlines()
            heads = [l for l in err if l.startswith(("Loaded:", "Resolved seeds:", "Chains:"))]
            heads += [l for l in out if l.startswith("summary:")]
            runs[run][inst] = {
                "log_lines": heads,
                "phases": chains[0]["phases"], "iters": chains[0]["iters"], "n": chains[0]["n"],
                "chains": [[c["seed"], c["t0"], c["best_cut"], c["final_cut"], c["kernel_best"], c["kernel_final"],
                            short_digest(c["best_partition"]), short_digest(c["final_partition"])]
                           for c in chains]}
    DIGESTS.write_text(json.dumps({"source": "reference results/<run>/logs/<instance>.{out,err}",
                                   "fields": ["seed", "t0", "best_cut", "final_cut", "kernel_best",
                                              "kernel_final", "best_partition_sha256_16",
                                              "final_partition_sha256_16"],
                                   "runs": runs}, separators=(",", ":")))
    print(f"wrote {sum(len(v['chains']) for r in runs.values() for v in r.values())} chain digests to {DIGESTS} "
          f"({DIGESTS.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
