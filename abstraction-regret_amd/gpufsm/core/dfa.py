"""Dense-table DFA, its random generator, and a builder with an absorbing dead state.

Interface parity: `src/gpufsm/core/dfa.py:22-110` (``ALPHABET``, ``DFA``, ``random_dfa``,
``DFABuilder``). The table is row-major by state: the successor of ``s`` on byte ``c`` is
``trans[s * 256 + c]``. Every entry is a valid state index (the builders guarantee it and
the device kernels rely on it).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ALPHABET = 256


@dataclass(frozen=True, eq=False)
class DFA:
    """Total deterministic automaton over bytes (``trans`` int32 ``[num_states * 256]``).

    Equality and hashing are by identity: the HIP backend keys its device-resident tables
    on the object. (The reference's generated ``__eq__`` compares arrays, which makes the
    type unhashable; identity is the only key that works for a cache.)
    """

    num_states: int
    start_state: int
    accept: np.ndarray  # bool [num_states]
    trans: np.ndarray  # int32 [num_states * 256]

    @property
    def table_bytes(self) -> int:
        return int(self.trans.size) * self.trans.itemsize


def random_dfa(num_states: int, *, accept_prob: float = 0.05, seed: int = 0) -> DFA:
    """Uniform random total DFA (`src/gpufsm/core/dfa.py:44-57`).

    Draw order from one ``numpy.random.default_rng(seed)``: first the ``num_states * 256``
    targets as int32 in ``[0, num_states)``, then ``num_states`` uniform floats whose value
    below ``accept_prob`` marks a state accepting. State 0 is the start and is forced
    non-accepting.
    """
    if num_states < 1:
        raise ValueError("num_states must be >= 1")
    rng = np.random.default_rng(seed)
    trans = rng.integers(0, num_states, size=num_states * ALPHABET, dtype=np.int32)
    accept = rng.random(num_states) < accept_prob
    accept[0] = False
    return DFA(num_states=num_states, start_state=0, accept=accept, trans=trans)


class DFABuilder:
    """Incremental DFA; unset edges go to an extra non-accepting absorbing dead state."""

    def __init__(self) -> None:
        self._accepting: list[bool] = []
        self._edges: list[dict[int, int]] = []
        self._start = 0

    def add_state(self, accept: bool = False) -> int:
        self._accepting.append(bool(accept))
        self._edges.append({})
        return len(self._accepting) - 1

    def _check(self, state: int) -> None:
        if not 0 <= state < len(self._accepting):
            raise IndexError(f"state {state} out of range (0..{len(self._accepting) - 1})")

    def set_start(self, state: int) -> None:
        self._check(state)
        self._start = state

    def add_transition(self, src: int, symbol: int, dst: int) -> None:
        self._check(src)
        self._check(dst)
        if not 0 <= symbol < ALPHABET:
            raise ValueError(f"symbol must be 0..255, got {symbol}")
        self._edges[src][symbol] = dst

    def build(self) -> DFA:
        n = len(self._accepting)
        if n == 0:
            raise ValueError("cannot build a DFA with zero states")
        dead = n
        table = np.full((n + 1, ALPHABET), dead, dtype=np.int32)
        for s, row in enumerate(self._edges):
            for c, d in row.items():
                table[s, c] = d
        accept = np.zeros(n + 1, dtype=bool)
        accept[:n] = self._accepting
        return DFA(num_states=n + 1, start_state=self._start, accept=accept, trans=table.reshape(-1))
