"""HIP executors behind the registry: the plain per-call drop-in.

Mirror the upstream CUDA adapters (`src/gpufsm/backends/cuda/nfa.py:55-139`,
`src/gpufsm/backends/cuda/dfa.py:21-53`): construction uploads the automaton, ``run_batch``
packs the batch, makes one native call, and returns ``batch_results`` with the batch's
kernel time on the first result and the host packing time as ``transfer_ms``. ``run`` is a
batch of one. An empty batch returns ``[]``.
"""

from __future__ import annotations

import time

from ...core.dfa import DFA
from ...core.nfa import NFA
from ...core.packing import pack_bytes
from ...core.result import Result, batch_results
from . import _lib
from .handles import DeviceDFA, DeviceNFA


class HIPNFAExecutor:
    def __init__(self, nfa: NFA, technique: str = "auto") -> None:
        if technique not in _lib.NFA_TECHNIQUES:
            raise ValueError(f"unknown HIP NFA technique {technique!r}")
        self.nfa = nfa
        self.technique = technique
        self._tech = _lib.NFA_TECHNIQUES[technique]
        self._dev = DeviceNFA(nfa)
        props = self._dev.props
        if technique != "auto":
            # unsupported sizes are ValueError at construction, as upstream's executors do
            # (`triton/worklist.py:118-122`); the library gives the reason
            try:
                self._dev.resolve(1, 0, self._tech)
            except _lib.HIPError as exc:
                self._dev.close()
                raise ValueError(f"HIP technique {technique!r} cannot run this NFA: {exc}"
                                 f" (chain-structured: {bool(props.chain_ok)}, "
                                 f"{nfa.num_states} states)") from exc

    def run(self, input_bytes: bytes) -> Result:
        return self.run_batch([input_bytes])[0]

    def run_batch(self, inputs: list[bytes]) -> list[Result]:
        if not inputs:
            return []
        t0 = time.perf_counter()
        data, offsets = pack_bytes(inputs)
        transfer_ms = (time.perf_counter() - t0) * 1000.0
        flags, lens, kernel_ms = self._dev.run_host(data, offsets, self._tech)
        return batch_results(flags, lens, kernel_ms, transfer_ms)

    def close(self) -> None:
        self._dev.close()


class HIPDFAExecutor:
    def __init__(self, dfa: DFA, technique: str = "auto") -> None:
        if technique not in _lib.DFA_TECHNIQUES:
            raise ValueError(f"unknown HIP DFA technique {technique!r}")
        self.dfa = dfa
        self.technique = technique
        self._tech = _lib.DFA_TECHNIQUES[technique]
        self._dev = DeviceDFA(dfa)
        if technique == "lds" and not self._dev.props.lds_ok:
            self._dev.close()
            raise ValueError("HIP technique 'lds' needs a DFA of at most 320 states")

    def run(self, input_bytes: bytes) -> Result:
        return self.run_batch([input_bytes])[0]

    def run_batch(self, inputs: list[bytes]) -> list[Result]:
        if not inputs:
            return []
        t0 = time.perf_counter()
        data, offsets = pack_bytes(inputs)
        transfer_ms = (time.perf_counter() - t0) * 1000.0
        flags, lens, kernel_ms = self._dev.run_host(data, offsets, self._tech)
        return batch_results(flags, lens, kernel_ms, transfer_ms)

    def close(self) -> None:
        self._dev.close()
