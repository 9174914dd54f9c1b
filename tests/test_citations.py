"""Every reference citation `file:line` (or `file:a-b`, `file:a,b-c`) in the documentation and the sources
names lines that exist in the cited reference file.

Cited names may be a path under the reference root (src/elastodynamics_cqbem_G.cu), a basename, or the
short forms used throughout (G.cu, H_collocation.cu, matrix_collection.cpp): a short form resolves to the
reference files whose basename ends with it, preferring the elastodynamics_cqbem_ file the hot path lives
in.  A citation passes when its lines lie inside one resolved file.  Runs only where the reference tree is
present (it never travels with the repository)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
EXTS = (".cu", ".cuh", ".cpp", ".hpp", ".h", ".json", ".txt", ".cmake")
SCANNED_DIRS = ("oracle", "include", "tests", "surfaceonlydeformables_amd", ".")
SCANNED_EXT = (".py", ".cpp", ".hpp", ".h", ".hip", ".md", ".sh")
NOT_OURS = ("VERDICT.md", "ADVICE.md", "SURVEY.md", "BASELINE.md")  # review and survey files written for this repo
PATTERN = re.compile(r"([A-Za-z_][\w./-]*\.(?:cu|cuh|cpp|hpp|h|json|txt|cmake)):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")


def reference_files():
    out = {}
    for dirpath, _, files in os.walk(REF):
        if "/.git" in dirpath or "/deps" in dirpath:
            continue
        for f in files:
            if f.endswith(EXTS):
                p = os.path.join(dirpath, f)
                with open(p, errors="replace") as fh:
                    out[os.path.relpath(p, REF)] = sum(1 for _ in fh)
    return out


def resolve(name, files):
    name = name.lstrip("./")
    if name.startswith("root/reference/"):
        name = name[len("root/reference/"):]
    if name in files:
        return [name]
    base = [p for p in files if os.path.basename(p) == name]
    if base:
        return base
    suffix = [p for p in files if p.endswith("_" + name) or p.endswith("/" + name)]
    preferred = [p for p in suffix if os.path.basename(p).startswith("elastodynamics_cqbem_")]
    return preferred or suffix


def scanned_files():
    seen = set()
    for d in SCANNED_DIRS:
        top = os.path.join(ROOT, d)
        walker = [(top, [], os.listdir(top))] if d == "." else os.walk(top)
        for dirpath, _, files in walker:
            if "_build" in dirpath or "__pycache__" in dirpath or "/_ref" in dirpath:
                continue
            for f in files:
                p = os.path.join(dirpath, f)
                if (f.endswith(SCANNED_EXT) and os.path.isfile(p) and p not in seen and f != "test_citations.py"
                        and f not in NOT_OURS):
                    seen.add(p)
                    yield p


def citations():
    for path in scanned_files():
        with open(path, errors="replace") as fh:
            for lineno, text in enumerate(fh, 1):
                for m in PATTERN.finditer(text):
                    name, spec = m.group(1), m.group(2)
                    if "/" not in name and name.count(".") > 1:  # e.g. a Python module path, not a file
                        continue
                    nums = [int(x) for part in spec.replace(" ", "").split(",") for x in part.split("-")]
                    yield os.path.relpath(path, ROOT), lineno, name, max(nums), min(nums)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_reference_citations_are_in_range():
    files = reference_files()
    bad, checked = [], 0
    for where, lineno, name, hi, lo in citations():
        cands = resolve(name, files)
        if not cands:
            continue  # not a reference file (one of this repository's own, or a library header)
        checked += 1
        if lo < 1 or not any(hi <= files[c] for c in cands):
            bad.append(f"{where}:{lineno}: {name}:{lo}..{hi} (file has {max(files[c] for c in cands)} lines)")
    assert checked > 100
    assert not bad, "\n".join(bad)
