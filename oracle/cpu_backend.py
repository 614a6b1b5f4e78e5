"""Registers the CPU oracle as ``Backend.CPU``. TEST INFRASTRUCTURE ONLY.

Mirrors `src/gpufsm/backends/cpu.py:28-67`: techniques ``reference`` (default, both kinds)
and ``bitmap`` (NFA), each compiled once per executor, each ``run`` timed with
``time.perf_counter`` into ``kernel_ms`` and ``total_ms``.
"""

from __future__ import annotations

import time

from gpufsm.core.registry import Backend, Kind, register, register_availability
from gpufsm.core.result import Result

from oracle.bitmap import BitmapProgram
from oracle.reference import simulate, simulate_dfa

_COMPILERS = {
    (Kind.NFA, "reference"): lambda a: (lambda data: simulate(a, data)),
    (Kind.NFA, "bitmap"): lambda a: BitmapProgram(a).run,
    (Kind.DFA, "reference"): lambda a: (lambda data: simulate_dfa(a, data)),
}


class CPUExecutor:
    def __init__(self, automaton, technique: str = "reference") -> None:
        self.automaton = automaton
        self.technique = technique
        self._run = _COMPILERS[(Kind.of(automaton), technique)](automaton)

    def run(self, input_bytes: bytes) -> Result:
        t0 = time.perf_counter()
        accepted, match_len = self._run(input_bytes)
        ms = (time.perf_counter() - t0) * 1000.0
        return Result(accepted=accepted, match_len=match_len, kernel_ms=ms, total_ms=ms)


def _make(automaton, technique: str) -> CPUExecutor:
    return CPUExecutor(automaton, technique)


_installed = False


def install() -> None:
    """Register the CPU techniques (idempotent)."""
    global _installed
    if _installed:
        return
    for kind, tech in _COMPILERS:
        register(kind, Backend.CPU, tech, default=(tech == "reference"))(_make)
    register_availability(Backend.CPU, lambda: True)
    _installed = True
