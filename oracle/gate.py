"""Oracle gate: no number is reported before the backend agrees with the oracle.

TEST INFRASTRUCTURE. Mirrors `src/gpufsm/bench/oracle.py:14-80` (``DEFAULT_CHECK``,
``OracleMismatch``, ``compare``, ``matches``, ``require``), including its refusal to pass
vacuously on an empty sample or a non-positive limit. The expected verdicts come from the
C oracle (``oracle.native``) rather than one Python ``run`` per input, so the gate can
check whole benchmark batches; the C oracle is itself pinned to the Python restatement.
"""

from __future__ import annotations

import numpy as np

from gpufsm.api import run_batch
from gpufsm.core.dfa import DFA
from gpufsm.core.packing import pack_bytes
from gpufsm.core.registry import Backend

from oracle import native

DEFAULT_CHECK = 64


class OracleMismatch(AssertionError):
    """A backend disagreed with the oracle; the run must not report timings."""


def expected(automaton, inputs: list[bytes]) -> list[tuple[bool, int]]:
    data, offsets = pack_bytes(inputs)
    if isinstance(automaton, DFA):
        flags, lens, _ = native.dfa_batch(automaton, data, offsets)
    else:
        flags, lens, _ = native.nfa_batch(automaton, data, offsets)
    return [(bool(f), int(n)) for f, n in zip(flags, lens)]


def compare(automaton, inputs: list[bytes], backend: Backend | str, technique: str | None = None,
            limit: int = DEFAULT_CHECK) -> list[tuple[int, tuple[bool, int], tuple[bool, int]]]:
    """``(index, got, expected)`` for every input of the first ``limit`` that differs."""
    if limit <= 0:
        raise ValueError(f"limit must be >= 1 to check anything, got {limit}")
    sample = inputs[:limit]
    if not sample:
        raise ValueError("no inputs to compare: the oracle gate would pass vacuously")
    got = run_batch(automaton, sample, backend=backend, technique=technique)
    want = expected(automaton, sample)
    bad = []
    for i, (res, ref) in enumerate(zip(got, want, strict=True)):
        if (res.accepted, res.match_len) != ref:
            bad.append((i, (res.accepted, res.match_len), ref))
    return bad


def matches(automaton, inputs, backend, technique=None, limit: int = DEFAULT_CHECK) -> bool:
    return not compare(automaton, inputs, backend, technique, limit)


def require(automaton, inputs, backend, technique=None, limit: int = DEFAULT_CHECK) -> None:
    bad = compare(automaton, inputs, backend, technique, limit)
    if not bad:
        return
    i, got, want = bad[0]
    name = backend.value if isinstance(backend, Backend) else backend
    raise OracleMismatch(
        f"{name}/{technique or 'default'} disagrees with the CPU reference on "
        f"{len(bad)}/{min(len(inputs), limit)} inputs; first at index {i}: "
        f"got {got}, expected {want}"
    )


def require_arrays(flags: np.ndarray, lens: np.ndarray, want_flags: np.ndarray,
                   want_lens: np.ndarray, what: str) -> None:
    """Full-batch array form of :func:`require`, for device-resident benchmark outputs."""
    if flags.size == 0:
        raise ValueError("no outputs to compare: the oracle gate would pass vacuously")
    bad = np.flatnonzero((flags != want_flags) | (lens != want_lens))
    if bad.size:
        i = int(bad[0])
        raise OracleMismatch(
            f"{what} disagrees with the oracle on {bad.size}/{flags.size} inputs; first at "
            f"index {i}: got ({bool(flags[i])}, {int(lens[i])}), expected "
            f"({bool(want_flags[i])}, {int(want_lens[i])})"
        )
