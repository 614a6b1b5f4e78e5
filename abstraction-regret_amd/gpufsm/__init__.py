"""gpufsm — latch-first-match NFA/DFA simulation with an MI355X-native HIP backend.

The public surface follows the upstream ``gpufsm`` package (`src/gpufsm/__init__.py`):
build an automaton, then ``run`` / ``run_batch`` / ``benchmark`` it on a backend::

    from gpufsm import NFABuilder, run_batch

    b = NFABuilder()
    s0 = b.add_state(); s1 = b.add_state(accept=True)
    b.set_start(s0); b.add_transition(s0, "a", s1)
    run_batch(b.build(), [b"a", b"b"], backend="hip")

The HIP kernels are reached through ``libgpufsm_hip.so`` (a C ABI loaded with ctypes).
The CPU oracle (``simulate`` / ``simulate_dfa`` in upstream) is test infrastructure here
and is not exported.
"""

from __future__ import annotations

from .api import benchmark, run, run_batch
from .core.dfa import DFA, DFABuilder, random_dfa
from .core.nfa import ANY_SYMBOL, NFA, NFABuilder
from .core.registry import (
    Backend,
    Kind,
    available_backends,
    is_available,
    list_kinds,
    list_techniques,
)
from .core.result import BenchmarkStats, Result

__version__ = "0.1.0"

__all__ = [
    "ANY_SYMBOL",
    "NFA",
    "NFABuilder",
    "DFA",
    "DFABuilder",
    "random_dfa",
    "Backend",
    "Kind",
    "Result",
    "BenchmarkStats",
    "run",
    "run_batch",
    "benchmark",
    "available_backends",
    "is_available",
    "list_kinds",
    "list_techniques",
]
