"""Writes tests/golden/reference_runs.json and copies the benchmark's Gset
instances into tests/golden/gset/.

reference_runs.json holds, per committed benchmark run of the reference
(results/<run>/results.csv: V100-SXM2 and V100-PCIE at 8e7 and 8e8 sweeps per
chain) and per instance: n, m, chains, best cut, hits at best, best chain and
its t0, kernel and wall seconds, and the 80 raw seeds, t0s, phases and iters
the run used.  These are data the reference's runs printed, not code.  The
Gset files are the instances those runs read (Gset/G60 ... G81).

Run in the build container (reads /root/reference):
    python tests/golden/make_reference_runs.py
"""
from __future__ import annotations

import csv
import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
OUT = HERE / "reference_runs.json"
RUNS = ["V100-SXM2-10M", "V100-PCIE-10M", "V100-SXM2-100M", "V100-PCIE-100M"]


def main():
    out = {"source": "reference results/<run>/results.csv"}
    names = set()
    for run in RUNS:
        rows = {}
        with open(REF / "results" / run / "results.csv", newline="") as f:
            for r in csv.DictReader(f):
                rows[r["instance"]] = {
                    "n": int(r["n"]), "m": int(r["m"]), "chains": int(r["chains"]),
                    "best_cut": int(r["best_cut"]), "hits_at_best": int(r["hits_at_best"]),
                    "best_chain": int(r["best_chain"]), "best_t0": float(r["best_t0"]),
                    "kernel_s": float(r["kernel_s"]), "wall_s": float(r["wall_s"]),
                    "t0s": [float(x) for x in r["t0s"].split(",")],
                    "seeds": [int(x) for x in r["seeds"].split(",")],
                    "phases": [int(x) for x in r["phases"].split(",")],
                    "iters": [int(x) for x in r["iters"].split(",")],
                }
                names.add(r["instance"])
        out[run] = rows
    OUT.write_text(json.dumps(out, indent=0))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")
    gdir = HERE / "gset"
    for name in sorted(names, key=lambda s: int(s[1:])):
        dst = gdir / name
        if not dst.exists():
            shutil.copyfile(REF / "Gset" / name, dst)
            print(f"copied Gset/{name} ({dst.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
