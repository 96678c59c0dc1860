"""Every file:line citation of a reference file in the repository's sources
and documents points inside that file (CPU only; skipped where the
reference tree is not mounted)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SKIP_DIRS = {".git", "_work", "gpurun_out", "__pycache__", "_build", ".pytest_cache"}
# documents written about the build rather than by it
SKIP_FILES = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}
EXTS = (".py", ".hip", ".h", ".hpp", ".cpp", ".c", ".md", ".txt")
CITE = re.compile(r"([A-Za-z_][\w./-]*\.(?:cpp|hpp|h|c|pl4))((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")


def _ref_index():
    idx = {}
    for root, dirs, files in os.walk(REF):
        dirs[:] = [d for d in dirs if d != ".git"]
        for f in files:
            idx.setdefault(f, []).append(os.path.join(root, f))
    return idx


def _own_files():
    own = set()
    for root, dirs, files in os.walk(ROOT):
        dirs[:] = [d for d in dirs if d not in SKIP_DIRS]
        own.update(files)
    return own


def _sources():
    for root, dirs, files in os.walk(ROOT):
        dirs[:] = [d for d in dirs if d not in SKIP_DIRS]
        for f in files:
            if f.endswith(EXTS) and f not in SKIP_FILES and f != "test_citations.py":
                yield os.path.join(root, f)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_reference_citations_are_in_range():
    idx = _ref_index()
    own = _own_files()
    lengths = {}
    bad = []
    checked = 0
    for path in _sources():
        text = open(path, errors="replace").read()
        for m in CITE.finditer(text):
            name = os.path.basename(m.group(1))
            if name not in idx or name in own:
                continue
            cands = [p for p in idx[name] if p.endswith(m.group(1))] or idx[name]
            n = 0
            for p in cands:
                if p not in lengths:
                    lengths[p] = sum(1 for _ in open(p, errors="replace"))
                n = max(n, lengths[p])
            for part in re.findall(r"(\d+)(?:-(\d+))?", m.group(2)):
                hi = int(part[1] or part[0])
                lo = int(part[0])
                checked += 1
                if lo < 1 or hi > n or hi < lo:
                    line = text[:m.start()].count("\n") + 1
                    bad.append(f"{os.path.relpath(path, ROOT)}:{line}: {m.group(0)} ({name} has {n} lines)")
    assert checked > 100
    assert not bad, "\n".join(bad)
