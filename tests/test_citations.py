"""Reference citations in the code are checkable: every `file:lines` cited
in include/, rfmf_amd/, oracle/, tests/ and bench.py names a file of the
reference checkout, every cited line range lies inside that file, and where
the comment names a code symbol next to the citation (an identifier with an
underscore, a `::` or an inner capital, on the citation's line or the lines
just before and after it) that the cited file also names, one of those
symbols occurs in the cited lines.  Skipped where the
reference checkout is absent (it is not shipped with this repository)."""
import re
from pathlib import Path

import pytest

from conftest import ROOT

REF = Path("/root/reference")
pytestmark = pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference checkout not present")

SOURCES = ["include", "rfmf_amd", "oracle", "tests", "bench.py", "__graft_entry__.py"]
SUFFIXES = {".h", ".hpp", ".cpp", ".hip", ".c", ".py"}
RANGES = r"\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*"
CITE = re.compile(r"(?<![\w/.])((?:src|tests|scripts)/[\w/]+\.\w+|[\w]+\.(?:cpp|h|hpp|cu|cuh|sh)|main\.cu|Makefile)"
                  r":(" + RANGES + r")")
# a continuation that cites more lines of the previous file: "B -- :314-325"
MORE = re.compile(r"(?<=\s):(" + RANGES + r")\b")
SYMBOL = re.compile(r"[A-Za-z_][A-Za-z0-9_]*(?:::[A-Za-z_][A-Za-z0-9_]*)*")


def _files():
    for s in SOURCES:
        p = ROOT / s
        if p.is_file():
            yield p
        else:
            for f in sorted(p.rglob("*")):
                if f.suffix in SUFFIXES and "_build" not in f.parts and "_ref" not in f.parts \
                        and f.name != "test_citations.py":
                    yield f


def _resolve(name: str):
    if "/" in name or name in ("main.cu", "Makefile"):
        p = REF / name
        return p if p.is_file() else None
    for d in ("src", "tests", "scripts"):
        p = REF / d / name
        if p.is_file():
            return p
    return None   # a file of this repository, not the reference


def _spans(ranges: str):
    for part in ranges.split(","):
        a, _, b = part.partition("-")
        yield int(a), int(b or a)


PATHS = re.compile(r"[\w/]+\.(?:cpp|hip|hpp|cuh|cu|sh|py|h|c)\b")


def _symbols(text: str):
    out = set()
    for tok in SYMBOL.findall(PATHS.sub(" ", text)):
        last = tok.split("::")[-1]
        if re.fullmatch(r"u?int\d+_t|size_t|ptrdiff_t|__\w+", last):
            continue
        if len(last) >= 4 and ("_" in tok or "::" in tok or re.search(r"[a-z][A-Z]", tok)):
            out.add(last)
    return out


def _citations():
    for f in _files():
        lines = f.read_text(errors="replace").splitlines()
        for i, line in enumerate(lines):
            prev = None
            for m in CITE.finditer(line):
                path = _resolve(m.group(1))
                if path is None:
                    prev = None
                    continue
                prev = path
                yield f, i, lines, path, m.group(2)
            if prev is not None:
                for m in MORE.finditer(line):
                    yield f, i, lines, prev, m.group(1)


def test_every_cited_range_lies_inside_its_reference_file():
    bad, n = [], 0
    for f, i, _, path, ranges in _citations():
        n += 1
        count = len(path.read_text(errors="replace").splitlines())
        for a, b in _spans(ranges):
            if not (1 <= a <= b <= count):
                bad.append(f"{f.relative_to(ROOT)}:{i + 1} cites {path.relative_to(REF)}:{a}-{b} "
                           f"({count} lines)")
    assert n > 50, n
    assert not bad, "\n".join(bad)


def test_cited_ranges_hold_the_symbols_named_beside_them():
    bad, checked = [], 0
    for f, i, lines, path, ranges in _citations():
        near = _symbols(" ".join(lines[max(0, i - 1):i + 2]))
        ref_lines = path.read_text(errors="replace").splitlines()
        body = "\n".join("\n".join(ref_lines[a - 1:b]) for a, b in _spans(ranges))
        whole = "\n".join(ref_lines)
        # symbols of the reference file only (names of this repository's own
        # functions beside a citation say what replaces it, not what it holds)
        named = {s for s in near if re.search(r"\b" + re.escape(s) + r"\b", whole)}
        if not named:
            continue
        checked += 1
        if not any(re.search(r"\b" + re.escape(s) + r"\b", body) for s in named):
            bad.append(f"{f.relative_to(ROOT)}:{i + 1} cites {path.relative_to(REF)}:{ranges} for "
                       f"{sorted(named)}")
    assert checked > 20, checked
    assert not bad, "\n".join(bad)
