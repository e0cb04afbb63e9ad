"""Every `file.rs:N` / `file.rs:N-M` citation of the reference in this
repository points inside the cited file (the reference's sources are read
from /root/reference, so the test is skipped where they are absent)."""
import os
import re

import pytest

from conftest import ROOT

REF = "/root/reference"
CITE = re.compile(r"((?:[\w.]+/)*[\w]+\.rs):(\d+)(?:-(\d+))?")
SKIP_DIRS = {".git", "_work", "gpurun_out", "golden", "__pycache__", "_build", "lib", ".pytest_cache"}
# Files written by others (the survey, the reviews) are not checked here.
SKIP_FILES = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md", "BASELINE.json"}
TEXT_EXT = (".py", ".cpp", ".h", ".hip", ".md", ".txt", "Makefile")


def reference_files():
    out = {}
    for dirpath, dirs, files in os.walk(REF):
        dirs[:] = [d for d in dirs if d not in (".git", "target")]
        for f in files:
            if f.endswith(".rs"):
                p = os.path.relpath(os.path.join(dirpath, f), REF)
                with open(os.path.join(dirpath, f), errors="replace") as fh:
                    out[p] = sum(1 for _ in fh)
    return out


def resolve(cited, files):
    """The reference file a citation names: a path suffix, or a bare file
    name (num.rs is num/src/num.rs, transform.rs is transform/src/...)."""
    hits = [p for p in files if p == cited or p.endswith("/" + cited)]
    if len(hits) > 1:
        preferred = [p for p in hits if p.startswith(cited.split(".")[0] + "/src/")]
        hits = preferred or hits
    return hits[0] if len(hits) == 1 else None


def citations():
    for dirpath, dirs, files in os.walk(ROOT):
        dirs[:] = [d for d in dirs if d not in SKIP_DIRS]
        for f in files:
            if f in SKIP_FILES or not f.endswith(TEXT_EXT):
                continue
            path = os.path.join(dirpath, f)
            with open(path, errors="replace") as fh:
                for lineno, line in enumerate(fh, 1):
                    for m in CITE.finditer(line):
                        yield os.path.relpath(path, ROOT), lineno, m.group(1), int(m.group(2)), int(m.group(3) or m.group(2))


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not on this machine")
def test_citations_point_inside_the_cited_files():
    files = reference_files()
    bad = []
    for where, lineno, cited, a, b in citations():
        ref = resolve(cited, files)
        if ref is None:
            bad.append(f"{where}:{lineno}: {cited} names no single reference file")
        elif not (1 <= a <= b <= files[ref]):
            bad.append(f"{where}:{lineno}: {cited}:{a}-{b} outside {ref} ({files[ref]} lines)")
    assert not bad, "\n".join(bad)
