"""The strict oracle gate as a GPU run: ``bench.py --gate --require hip``.

Upstream's gate (`scripts/oracle_gate.py:34-108`) fails when a required backend is missing
or any technique disagrees with the oracle; it must exit 0 here, having checked a non-zero
number of (technique, case) pairs, every one matching, and written a schema-valid CSV.
"""

from __future__ import annotations

import csv
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.gpu


def test_gate_requires_hip_and_passes(tmp_path):
    out = tmp_path / "gate.csv"
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gate", "--require", "hip",
                          "--gate-csv", str(out), "--gate-inputs", "48",
                          "--gate-length", "200"],
                         capture_output=True, text=True, timeout=240, cwd=str(ROOT))
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    with open(out, newline="") as f:
        assert f.readline().startswith("# {")  # the environment capture
        rows = list(csv.DictReader(f))
    hip = [r for r in rows if r["backend"] == "hip"]
    assert hip, rows[:5]
    assert all(r["status"] in ("ok", "not-applicable") and int(r["mismatches"]) == 0
               for r in hip), [r for r in hip if r["status"] not in ("ok", "not-applicable")][:5]
    assert "PASS" in res.stdout
    techniques = {(r["kind"], r["technique"]) for r in hip if r["status"] == "ok"}
    for want in (("nfa", "worklist"), ("nfa", "bitpacked"), ("nfa", "chain"),
                 ("nfa", "register"), ("dfa", "gather"), ("dfa", "lds"), ("dfa", "split")):
        assert want in techniques, (want, sorted(techniques))
    assert sum(int(r["inputs"]) for r in hip) > 0
