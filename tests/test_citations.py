"""Every reference citation in the repository points at real code.

A citation is `[src/]<file>:<a>[-<b>][, <c>[-<d>] ...]` naming one of the reference's files, plus the
bare continuations `:<c>[-<d>]` that follow it on the same line.  Checks:
  - every range lies inside the cited file;
  - a citation written right after the name of a function, macro or typedef that the cited file
    defines overlaps that definition (a function: from its signature to its closing brace).
The reference tree is read as text; the test is skipped where it is absent.
"""
import os
import re
import subprocess

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the task's own records, written outside this repository's code, are not checked
SKIP = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree absent")

RANGE = r"\d+(?:-\d+)?"
RANGES = RANGE + r"(?:,\s*" + RANGE + r")*"


def _ref_files():
    out = {}
    for f in os.listdir(os.path.join(REF, "src")):
        with open(os.path.join(REF, "src", f), errors="replace") as fh:
            out[f] = fh.read().split("\n")
    for f in ("Makefile", "README", "README.GetHI"):
        p = os.path.join(REF, f)
        if os.path.exists(p):
            with open(p, errors="replace") as fh:
                out[f] = fh.read().split("\n")
    return out


def _definitions(lines):
    """name -> (first, last) line of each function, macro and typedef the file defines (1-based)."""
    defs = {}
    fn = re.compile(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;]*$")
    for i, line in enumerate(lines, 1):
        m = re.match(r"\s*#\s*define\s+([A-Za-z_]\w*)", line)
        if m:
            defs.setdefault(m.group(1), (i, i))
            continue
        m = re.match(r"^\}\s*([A-Za-z_]\w*)\s*;", line)
        if m:  # end of a typedef struct
            start = i
            while start > 1 and not lines[start - 1].startswith("typedef"):
                start -= 1
            defs.setdefault(m.group(1), (start, i))
            continue
        m = fn.match(line)
        if m and not line.startswith(("return", "else", "if", "for", "while")):
            end = i
            while end < len(lines) and lines[end - 1].rstrip() != "}":
                end += 1
            defs.setdefault(m.group(1), (i, end))
    return defs


def _citations(ref):
    names = "|".join(re.escape(k) for k in sorted(ref, key=len, reverse=True))
    tok = re.compile(r"(?<![\w/.])(?:src/)?(?P<file>" + names + r"):(?P<r>" + RANGES + r")"
                     r"|(?<![\w.])`?:(?P<r2>" + RANGES + r")")
    files = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True, text=True, check=True).stdout.split()
    for f in files:
        if f in SKIP or f.endswith((".json", ".npy", ".dat", ".txt", ".npz")):
            continue
        path = os.path.join(ROOT, f)
        if not os.path.isfile(path):
            continue
        with open(path, errors="replace") as fh:
            text = fh.read().split("\n")
        for ln, line in enumerate(text, 1):
            cur = None
            for m in tok.finditer(line):
                if m.group("file"):
                    cur = m.group("file")
                    rng, direct = m.group("r"), True
                elif cur is not None:
                    rng, direct = m.group("r2"), False
                else:
                    continue
                parts = []
                for part in re.split(r",\s*", rng):
                    a, _, b = part.partition("-")
                    parts.append((int(a), int(b) if b else int(a)))
                yield f, ln, line, cur, parts, line[:m.start()] if direct else None


def test_citations_in_range_and_on_target():
    ref = _ref_files()
    defs = {k: _definitions(v) for k, v in ref.items()}
    ident = re.compile(r"([A-Za-z_]\w*)(?:\(\w*\))?;?`?\s*\(?`?\s*$")
    bad, n = [], 0
    for f, ln, line, rf, parts, before in _citations(ref):
        n += len(parts)
        nlines = len(ref[rf]) - (1 if ref[rf] and ref[rf][-1] == "" else 0)
        out = [(a, b) for a, b in parts if not 1 <= a <= b <= nlines]
        for a, b in out:
            bad.append(f"{f}:{ln}: {rf}:{a}-{b} outside the file ({nlines} lines): {line.strip()[:100]}")
        if out or before is None:
            continue
        # a list of ranges after a list of names: the named definition overlaps one of them
        m = ident.search(before)
        if m and m.group(1) in defs[rf]:
            lo, hi = defs[rf][m.group(1)]
            if all(b < lo or a > hi for a, b in parts):
                bad.append(f"{f}:{ln}: {m.group(1)} is at {rf}:{lo}-{hi}, cited {parts}: {line.strip()[:100]}")
    assert n > 200
    assert not bad, "\n".join(bad)


def test_definitions_parser():
    """The definition finder on the reference's own get_inu and mk_T_maps."""
    ref = _ref_files()
    d = _definitions(ref["pixelize.c"])
    assert d["get_inu"] == (28, 55)
    assert d["mk_T_maps"] == (150, 286)
    assert d["N_SUBPART"] == (25, 25)
    assert _definitions(ref["common_gh.h"])["ParamGetHI"] == (58, 129)
