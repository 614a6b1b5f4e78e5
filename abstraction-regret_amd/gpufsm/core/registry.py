"""(Kind, Backend, technique) -> executor-factory registry: the one extension point.

Interface parity: `src/gpufsm/core/registry.py:27-162`. Names, signatures, the
one-default-per-(kind, backend) rule, availability probes with a recorded reason, and the
``KeyError`` texts are the reference's, because callers and tests match on them. The
``Backend`` enum gains ``HIP``; the NVIDIA-only members are kept so that code naming them
still parses, and they simply have nothing registered.
"""

from __future__ import annotations

from collections.abc import Callable
from enum import Enum
from typing import Protocol, runtime_checkable

from .dfa import DFA
from .nfa import NFA
from .result import Result

Automaton = NFA | DFA


class Kind(str, Enum):
    """Automaton model a technique simulates."""

    NFA = "nfa"
    DFA = "dfa"

    @classmethod
    def of(cls, automaton: Automaton) -> Kind:
        if isinstance(automaton, NFA):
            return cls.NFA
        if isinstance(automaton, DFA):
            return cls.DFA
        raise TypeError(f"expected an NFA or a DFA, got {type(automaton).__name__}")


class Backend(str, Enum):
    CPU = "cpu"
    TRITON = "triton"
    CUDA = "cuda"
    WARP = "warp"
    HIP = "hip"  # hand-written gfx950 kernels behind a C ABI (this package)


@runtime_checkable
class Executor(Protocol):
    def run(self, input_bytes: bytes) -> Result: ...


ExecutorFactory = Callable[[Automaton, str], Executor]

_REGISTRY: dict[tuple[Kind, Backend, str], ExecutorFactory] = {}
_AVAILABILITY: dict[Backend, Callable[[], bool]] = {}
_DEFAULTS: dict[tuple[Kind, Backend], str] = {}
_UNAVAILABLE: dict[Backend, str] = {}


def register(
    kind: Kind, backend: Backend, technique: str, default: bool = False
) -> Callable[[ExecutorFactory], ExecutorFactory]:
    """Decorator registering ``factory(automaton, technique) -> Executor``.

    ``default=True`` declares the technique used when a caller names none; a second,
    different default for the same (kind, backend) raises ``RuntimeError``.
    """

    def deco(factory: ExecutorFactory) -> ExecutorFactory:
        _REGISTRY[(kind, backend, technique)] = factory
        if default:
            previous = _DEFAULTS.get((kind, backend))
            if previous is not None and previous != technique:
                raise RuntimeError(
                    f"two defaults declared for {backend.value}/{kind.value}: "
                    f"{previous!r} and {technique!r}"
                )
            _DEFAULTS[(kind, backend)] = technique
        return factory

    return deco


def unregister_backend(backend: Backend) -> None:
    """Drop every registration of ``backend`` (used by tests that install a stand-in)."""
    for key in [k for k in _REGISTRY if k[1] == backend]:
        del _REGISTRY[key]
    for key in [k for k in _DEFAULTS if k[1] == backend]:
        del _DEFAULTS[key]
    _AVAILABILITY.pop(backend, None)


def register_availability(backend: Backend, probe: Callable[[], bool]) -> None:
    _AVAILABILITY[backend] = probe


def set_unavailable_reason(backend: Backend, reason: str) -> None:
    """Record why ``backend``'s probe said no (shown by ``unavailable_reason``)."""
    _UNAVAILABLE[backend] = reason


def unavailable_reason(backend: Backend) -> str | None:
    return _UNAVAILABLE.get(backend)


def is_available(backend: Backend) -> bool:
    probe = _AVAILABILITY.get(backend)
    try:
        return bool(probe()) if probe else any(b == backend for _, b, _ in _REGISTRY)
    except Exception:
        return False


def get_factory(kind: Kind, backend: Backend, technique: str | None) -> tuple[str, ExecutorFactory]:
    """Resolve ``(kind, backend, technique)``; ``technique=None`` means the declared default."""
    techs = list_techniques(backend, kind)
    if not techs:
        reason = (
            "the backend is not available on this machine (import or build failed); "
            "run `gpufsm env` to see why"
            if not is_available(backend)
            else f"it registers no {kind.value} technique"
        )
        raise KeyError(
            f"backend {backend.value!r} has no {kind.value} technique: {reason}. "
            f"Available backends: {[b.value for b in available_backends()]}"
        )
    if technique is None:
        technique = _DEFAULTS.get((kind, backend))
        if technique is None:
            raise KeyError(
                f"backend {backend.value!r} declares no default {kind.value} technique; "
                f"name one explicitly (available: {techs}) or mark one with "
                f"register(..., default=True)"
            )
    if (kind, backend, technique) not in _REGISTRY:
        raise KeyError(
            f"{kind.value} technique {technique!r} not registered for backend "
            f"{backend.value!r}; available: {techs}"
        )
    return technique, _REGISTRY[(kind, backend, technique)]


def list_techniques(backend: Backend, kind: Kind = Kind.NFA) -> list[str]:
    return [t for (k, b, t) in _REGISTRY if b == backend and k == kind]


def list_kinds(backend: Backend) -> list[Kind]:
    return [k for k in Kind if list_techniques(backend, k)]


def available_backends() -> list[Backend]:
    return [b for b in Backend if list_kinds(b) and is_available(b)]
