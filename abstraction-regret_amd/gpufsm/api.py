"""Public entry points: ``run``, ``run_batch``, ``benchmark``.

Interface parity: `src/gpufsm/api.py:15-97`. The kind is taken from the automaton, the
technique from the registry, and the executor protocol is the reference's: ``run(bytes)``
is required, ``run_batch(list[bytes])`` is used when present. On the loop fallback the
batch's summed time goes on the first result, as in the reference.

Deviation: the default ``backend`` is :attr:`Backend.HIP`. The CPU oracle is test
infrastructure in this project and is not part of the installed package, so
``Backend.CPU`` has no technique unless a caller registers one.
"""

from __future__ import annotations

from . import backends as _backends  # noqa: F401  (registers the backends)
from .core.registry import Automaton, Backend, Kind, get_factory
from .core.result import BenchmarkStats, Result, batch_results


def run(
    automaton: Automaton,
    input_bytes: bytes,
    backend: Backend | str = Backend.HIP,
    technique: str | None = None,
) -> Result:
    """Run ``automaton`` over one input."""
    backend = Backend(backend)
    technique, factory = get_factory(Kind.of(automaton), backend, technique)
    return factory(automaton, technique).run(input_bytes)


def run_batch(
    automaton: Automaton,
    inputs: list[bytes],
    backend: Backend | str = Backend.HIP,
    technique: str | None = None,
) -> list[Result]:
    """Run ``automaton`` over each input; ``results[0].kernel_ms`` is the batch time.

    As in the reference, a fresh executor is built for every call, so this is the plain
    per-call path: the automaton is uploaded, the batch packed and copied, one launch,
    results copied back. :class:`gpufsm.batch.BatchEngine` keeps automata and inputs
    resident instead.
    """
    backend = Backend(backend)
    technique, factory = get_factory(Kind.of(automaton), backend, technique)
    executor = factory(automaton, technique)
    batch = getattr(executor, "run_batch", None)
    if callable(batch):
        return batch(inputs)
    per_input = [executor.run(b) for b in inputs]
    if not per_input:
        return []
    return batch_results(
        [r.accepted for r in per_input],
        [r.match_len for r in per_input],
        kernel_ms=sum(r.kernel_ms for r in per_input),
        transfer_ms=sum(r.transfer_ms for r in per_input),
    )


def benchmark(
    automaton: Automaton,
    input_bytes: bytes,
    backend: Backend | str = Backend.HIP,
    technique: str | None = None,
    repeats: int = 10,
    warmup: int = 3,
) -> BenchmarkStats:
    """Time ``repeats`` runs of one executor after ``warmup`` untimed runs."""
    if repeats < 1:
        raise ValueError("repeats must be >= 1")
    backend = Backend(backend)
    technique, factory = get_factory(Kind.of(automaton), backend, technique)
    executor = factory(automaton, technique)
    last: Result | None = None
    for _ in range(max(0, warmup)):
        last = executor.run(input_bytes)
    raw: list[float] = []
    for _ in range(repeats):
        last = executor.run(input_bytes)
        raw.append(last.kernel_ms)
    assert last is not None
    return BenchmarkStats(
        backend=backend.value,
        technique=technique,
        accepted=last.accepted,
        match_len=last.match_len,
        n=repeats,
        raw_ms=raw,
    )
