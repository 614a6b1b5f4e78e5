"""bench.py's contract on the host: one GPU only, argument checks, and no GPU → a clear exit.

These run on the CPU (no device): the checks happen before any GPU work.
"""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench(*args, timeout=120):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=timeout, cwd=str(ROOT))


def test_more_than_one_gpu_is_refused_with_a_message():
    res = _bench("--gpus", "2")
    assert res.returncode == 2
    assert "one GPU" in res.stderr


def test_steps_must_be_positive():
    res = _bench("--steps", "0")
    assert res.returncode == 2
    assert "--steps" in res.stderr


def test_traffic_table_is_per_batch_and_near_the_input():
    """The committed counter summaries hold HBM bytes per batch: the headline walk reads its
    input about once (268,566,536 algorithmic bytes per batch)."""
    table = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    lds = table["lds:dfa256_4096x65536_nonlatching"]
    assert 1.0 <= lds / 268_566_536 < 1.05
    assert table["split:dfa256_4096x65536_nonlatching"] > lds
