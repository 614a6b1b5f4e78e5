"""HIP backend: hand-written gfx950 kernels behind the C ABI of ``libgpufsm_hip.so``.

Registers ``Backend.HIP`` techniques with the unchanged registry
(`src/gpufsm/core/registry.py:62-117`), the way the upstream CUDA backend registers its own
(`src/gpufsm/backends/cuda/__init__.py:19-28`):

* NFA: ``auto`` (default; ``chain`` when the automaton is chain-structured, else
  ``register`` for at most 512 states over batches of 1024 inputs or more, else
  ``worklist``), ``worklist``, ``bitpacked``, ``chain``, ``register``;
* DFA: ``auto`` (default; ``split`` when inputs are long, else ``lds`` for tables of at
  most 320 states, else ``gather``), ``gather``, ``lds``, ``split``.

Techniques are registered whenever the library loads. The availability probe, registered
unconditionally, additionally needs a visible gfx950 device; when it says no, the reason is
recorded for :func:`gpufsm.core.registry.unavailable_reason`.
"""

from __future__ import annotations

from ...core.registry import Backend, Kind, register, register_availability, set_unavailable_reason
from . import _lib


def available() -> bool:
    try:
        _lib.load()
    except _lib.HIPUnavailable as exc:
        set_unavailable_reason(Backend.HIP, str(exc))
        return False
    try:
        arch = _lib.device_arch(0)
    except _lib.HIPError as exc:
        set_unavailable_reason(Backend.HIP, f"no usable HIP device: {exc}")
        return False
    if not arch.startswith("gfx950"):
        set_unavailable_reason(Backend.HIP, f"device 0 is {arch}, the kernels are built for gfx950")
        return False
    return True


def _make_nfa(automaton, technique: str):
    from .executors import HIPNFAExecutor

    return HIPNFAExecutor(automaton, technique)


def _make_dfa(automaton, technique: str):
    from .executors import HIPDFAExecutor

    return HIPDFAExecutor(automaton, technique)


try:
    _lib.load()
    _LOADED = True
except _lib.HIPUnavailable as _exc:  # pragma: no cover - depends on the build
    set_unavailable_reason(Backend.HIP, str(_exc))
    _LOADED = False

if _LOADED:
    for _tech in _lib.NFA_TECHNIQUES:
        register(Kind.NFA, Backend.HIP, _tech, default=(_tech == "auto"))(_make_nfa)
    for _tech in _lib.DFA_TECHNIQUES:
        register(Kind.DFA, Backend.HIP, _tech, default=(_tech == "auto"))(_make_dfa)

register_availability(Backend.HIP, available)
