"""The reference's shell tests, run against the rfmf_main CLI.

* Cut consistency (tests/test_cut_consistency.sh): for every committed Gset
  instance, each chain's incremental best/final cut (stderr
  `kernel-best-val` / `kernel-final-val`) equals the cut the CLI recomputes
  from the returned partitions (stdout `best-cut-value` / `final-cut-value`),
  and the summary line names the right best chain, value and tie count.
* Chain isolation and determinism (tests/test_chain_isolation.sh): chain i's
  stdout block inside a C-chain launch is byte-identical to a C=1 run of the
  same (seed, t0), including the int64 score path (a t0 of 0.01 in the mix),
  and an identical command gives identical bytes.

Each case runs the binary as its own process, one at a time."""
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "rfmf_amd" / "bin" / "rfmf_main"
GSET = ROOT / "tests" / "golden" / "gset"


def run(instance, seeds, t0s, phases, iters, both=False):
    args = [str(CLI), str(GSET / instance), "--seeds", seeds, "--t0s", t0s, "--phases", phases, "--iters", iters]
    print("[cli]", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = re.sub(r"time-sec: [0-9.]+", "time-sec: <normalized>", r.stdout)
    return (out, r.stderr) if both else out


@pytest.mark.parametrize("inst", ["G1", "G11", "G61", "G62", "G70", "G81"])
def test_cut_consistency(inst):
    C = 3
    out, err = run(inst, "42,43,42", "105,105,90", "64,32,16,8,4,2,1", ",".join(["120"] * 7), both=True)
    kbest = [int(x) for x in re.findall(r"^chain \d+: kernel-best-val: (-?\d+)", err, re.M)]
    kfinal = [int(x) for x in re.findall(r"kernel-final-val: (-?\d+)$", err, re.M)]
    rbest = [int(x) for x in re.findall(r"^device: .*best-cut-value: (-?\d+)", out, re.M)]
    rfinal = [int(x) for x in re.findall(r"^device: .*final-cut-value: (-?\d+)", out, re.M)]
    assert len(kbest) == len(kfinal) == len(rbest) == len(rfinal) == C
    assert kbest == rbest, f"{inst}: incremental best != recomputed"
    assert kfinal == rfinal, f"{inst}: incremental final != recomputed"
    m = re.search(r"^summary: chains: (\d+) best-chain: (\d+) best-cut-value: (-?\d+).*chains-at-best: (\d+)",
                  out, re.M)
    assert m, "summary line missing"
    chains, best_chain, best_val, at_best = map(int, m.groups())
    mx = max(rbest)
    assert (chains, best_chain, best_val, at_best) == (C, rbest.index(mx), mx, rbest.count(mx))


def _blocks(out):
    """Per-chain stdout blocks: the `device:` line and its Solution- lines."""
    blocks, cur = [], None
    for line in out.splitlines():
        if line.startswith("device:"):
            cur = [line]
            blocks.append(cur)
        elif line.startswith("Solution-") and cur is not None:
            cur.append(line)
    return ["\n".join(b) for b in blocks]


ISOLATION = [
    ("G1-iso", "G1", "42,7,42,9", "105,105,90,120", "8,4,2,1", "150,150,150,150", [0, 1, 2, 3]),
    ("G11-iso33", "G11", ",".join(str(s) for s in range(501, 534)), ",".join(["105"] * 33), "8,1", "100,100",
     [0, 15, 32]),
    ("G81-iso", "G81", "3,4", "105,90", "64", "100", [0, 1]),
    ("G70-iso", "G70", "13,14", "105,105", "8,2,1", "150,150,150", [0, 1]),
    # t0 = 0.01 puts the whole launch on the int64 score path; the C=1 run of
    # the t0 = 105 chain takes the int32 path and must still match.
    ("G1-iso-int64", "G1", "42,42", "0.01,105", "8,4", "150,150", [0, 1]),
]


@pytest.mark.parametrize("label,inst,seeds,t0s,phases,iters,idx", ISOLATION, ids=[c[0] for c in ISOLATION])
def test_chain_isolation(label, inst, seeds, t0s, phases, iters, idx):
    multi = _blocks(run(inst, seeds, t0s, phases, iters))
    s, t = seeds.split(","), t0s.split(",")
    assert len(multi) == len(s)
    for i in idx:
        single = _blocks(run(inst, s[i], t[i], phases, iters))
        assert len(single) == 1 and single[0] == multi[i], f"{label}: chain {i} differs from its C=1 run"


@pytest.mark.parametrize("inst,seeds,t0s,phases,iters", [
    ("G1", "42,7,42", "105,105,90", "64,32,16,8,4,2,1", ",".join(["150"] * 7)),
    ("G81", "3,4", "105,90", "64", "100"),
])
def test_determinism(inst, seeds, t0s, phases, iters):
    assert run(inst, seeds, t0s, phases, iters) == run(inst, seeds, t0s, phases, iters)


# ------------------------------------------------------- committed benchmark logs

import json  # noqa: E402

GOLDEN = ROOT / "tests" / "golden"


def _committed(inst):
    fx = json.loads((GOLDEN / "committed_digests.json").read_text())
    rec = fx["runs"]["V100-SXM2-10M"][inst]
    return [dict(zip(fx["fields"], c)) for c in rec["chains"]], rec


@pytest.mark.parametrize("inst", ["G62", "G63", "G70", "G81"])
def test_cli_lines_match_committed_logs(inst):
    """The reference's own benchmark command for the instance (its 80 resolved
    seeds and t0s, ladder 128..1) with 100 sweeps per phase: the stderr
    banner lines equal the committed run's (results/V100-SXM2-10M/logs/*.err)
    except the sweep total, and every stdout line has the committed .out
    layout: the same device-line prefix per chain, partitions of n bits and
    the summary fields."""
    chains, rec = _committed(inst)
    seeds = ",".join(str(c["seed"]) for c in chains)
    t0s = ",".join(repr(c["t0"]) for c in chains)
    phases = ",".join(str(p) for p in rec["phases"])
    iters = ",".join(["100"] * len(rec["phases"]))
    args = [str(CLI), str(GSET / inst), "--seeds", seeds, "--t0s", t0s, "--phases", phases, "--iters", iters]
    print(f"[cli-log] {inst}: 80 committed chains x 800 sweeps", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    loaded, resolved, banner, summary = rec["log_lines"]
    err = r.stderr.splitlines()
    assert loaded in err and resolved in err
    total = sum(rec["iters"])
    assert banner.replace(f"total-sweeps-per-chain={total}", "total-sweeps-per-chain=800") in err
    out = r.stdout.splitlines()
    n = rec["n"]
    assert len(out) == 3 * len(chains) + 1
    for c, ch in enumerate(chains):
        head = (f"device: GPU instance: {inst} --seed {ch['seed']} --phases {phases} --t0 {ch['t0']:.4f} "
                f"--iters {iters} best-cut-value: ")
        line = out[3 * c]
        assert line.startswith(head), (line[:200], head)
        assert re.fullmatch(r"-?\d+ final-cut-value: -?\d+ time-sec: \d+\.\d{4}", line[len(head):]), line
        assert re.fullmatch(r"Solution-best-partition: [01]{%d}" % n, out[3 * c + 1])
        assert re.fullmatch(r"Solution-final-partition: [01]{%d}" % n, out[3 * c + 2])
    assert re.fullmatch(r"summary: chains: 80 best-chain: \d+ best-cut-value: -?\d+ time-sec: \d+\.\d{4} "
                        r"chains-at-best: \d+", out[-1]), out[-1]
    assert re.fullmatch(r"summary: chains: 80 best-chain: \d+ best-cut-value: -?\d+ time-sec: \d+\.\d{4} "
                        r"chains-at-best: \d+", summary)
