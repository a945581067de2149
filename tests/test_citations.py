"""Every ``file:line`` citation of a reference source in this repository's code
points inside that file.

Runs where the isce2 sources are mounted at /root/reference (the build
container); skipped elsewhere.
"""
import collections
import os
import re
from pathlib import Path

import pytest

REF = Path("/root/reference")
REPO = Path(__file__).resolve().parents[1]
CITE = re.compile(r"([A-Za-z_][A-Za-z0-9_]*\.(?:cu|cpp|h|py|F|f90|f))"
                  r"((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")
SCAN = ("isce2_amd", "oracle", "include", "tests")


def _reference_lengths():
    lengths = collections.defaultdict(int)
    for dp, dn, fn in os.walk(REF):
        dn[:] = [d for d in dn if d != ".git"]
        for f in fn:
            try:
                with open(os.path.join(dp, f), "rb") as fh:
                    n = sum(1 for _ in fh)
            except OSError:
                continue
            lengths[f] = max(lengths[f], n)
    return lengths


@pytest.mark.skipif(not REF.is_dir(), reason="reference sources not mounted here")
def test_reference_citations_are_in_range():
    lengths = _reference_lengths()
    files = [REPO / "bench.py", REPO / "__graft_entry__.py"]
    for d in SCAN:
        files += [p for p in (REPO / d).rglob("*")
                  if p.suffix in (".py", ".h", ".hip", ".cpp") and "__pycache__" not in p.parts]
    bad, checked = [], 0
    for p in files:
        for ln, line in enumerate(p.read_text(errors="replace").splitlines(), 1):
            for m in CITE.finditer(line):
                name, ranges = m.group(1), m.group(2)
                if name not in lengths:
                    continue
                checked += 1
                nums = [int(x) for x in re.findall(r"\d+", ranges)]
                if max(nums) > lengths[name]:
                    bad.append(f"{p.relative_to(REPO)}:{ln}: {name}{ranges} (file has {lengths[name]} lines)")
                pairs = re.findall(r"(\d+)-(\d+)", ranges)
                bad += [f"{p.relative_to(REPO)}:{ln}: {name}{ranges} (empty range)"
                        for a, b in pairs if int(a) > int(b)]
    assert checked > 100
    assert not bad, "\n".join(bad)
