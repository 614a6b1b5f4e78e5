"""Per-input result record and timing aggregates.

Interface parity: `src/gpufsm/core/result.py:11-94` (``Result``, ``batch_results``,
``BenchmarkStats``), field for field.
"""

from __future__ import annotations

import math
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any


@dataclass
class Result:
    """Verdict of one automaton run over one input, plus timings."""

    accepted: bool
    match_len: int
    kernel_ms: float = 0.0
    total_ms: float = 0.0
    transfer_ms: float = 0.0
    mem_bytes: int = 0

    def matches(self, other: Result) -> bool:
        """Verdict equality; timings are ignored."""
        return self.accepted == other.accepted and self.match_len == other.match_len


def batch_results(
    flags: Sequence[Any],
    lens: Sequence[Any],
    kernel_ms: float,
    transfer_ms: float = 0.0,
) -> list[Result]:
    """One :class:`Result` per input of a batched launch.

    The launch has one kernel time; it is carried by the first result only and the others
    hold zero, so ``results[0].kernel_ms`` is the batch time on every path
    (`src/gpufsm/core/result.py:27-50`).
    """
    out = []
    for i in range(len(flags)):
        first = i == 0
        out.append(
            Result(
                accepted=bool(flags[i]),
                match_len=int(lens[i]),
                kernel_ms=kernel_ms if first else 0.0,
                total_ms=(kernel_ms + transfer_ms) if first else 0.0,
                transfer_ms=transfer_ms if first else 0.0,
            )
        )
    return out


@dataclass
class BenchmarkStats:
    """Timings of ``n`` repeated runs after warmup, with mean / std / normal CI95."""

    backend: str
    technique: str
    accepted: bool
    match_len: int
    n: int
    raw_ms: list[float] = field(default_factory=list)

    @property
    def mean_ms(self) -> float:
        return sum(self.raw_ms) / len(self.raw_ms) if self.raw_ms else 0.0

    @property
    def std_ms(self) -> float:
        if len(self.raw_ms) < 2:
            return 0.0
        m = self.mean_ms
        return math.sqrt(sum((x - m) ** 2 for x in self.raw_ms) / (len(self.raw_ms) - 1))

    @property
    def ci95_ms(self) -> float:
        if len(self.raw_ms) < 2:
            return 0.0
        return 1.96 * self.std_ms / math.sqrt(len(self.raw_ms))

    def as_row(self) -> dict[str, float | str | int | bool]:
        return {
            "backend": self.backend,
            "technique": self.technique,
            "accepted": self.accepted,
            "match_len": self.match_len,
            "n": self.n,
            "mean_ms": round(self.mean_ms, 6),
            "std_ms": round(self.std_ms, 6),
            "ci95_ms": round(self.ci95_ms, 6),
        }
