"""NFA in compressed-sparse-row form and the builder that produces it.

Interface parity: `src/gpufsm/core/nfa.py:16-157` (``ANY_SYMBOL``, ``NFA``, ``NFABuilder``).
Every field, its dtype and the builder's validation rules are the reference's, because the
HIP backend has to consume automata built by the reference's own code unchanged.

Layout (per state ``s``):

* symbol edges ``sym_targets[sym_row_ptr[s]:sym_row_ptr[s+1]]`` with labels
  ``sym_symbols[...]`` in ``0..255`` or :data:`ANY_SYMBOL` (a wildcard edge);
* epsilon edges ``eps_targets[eps_row_ptr[s]:eps_row_ptr[s+1]]``.

The row order inside a state is insertion order. Duplicate edges and self-loops are legal.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ANY_SYMBOL = 256
"""Label of a wildcard edge: it is taken on every input byte."""


def _coerce_symbol(symbol: int | str) -> int:
    """Validate one edge label and return it as an int.

    Rules (`src/gpufsm/core/nfa.py:20-40`): a one-character string is its code point and
    must lie in ``0..255``; an int must be ``0..255`` or :data:`ANY_SYMBOL`. ``chr(256)`` is
    therefore rejected instead of silently becoming the wildcard.
    """
    if isinstance(symbol, str):
        if len(symbol) != 1:
            raise ValueError(f"symbol string must be a single char, got {symbol!r}")
        code = ord(symbol)
        if code > 255:
            raise ValueError(
                f"symbol {symbol!r} has code point {code}; transitions are over bytes "
                f"(0..255). Encode the text first, or pass ANY_SYMBOL for a wildcard."
            )
        return code
    value = int(symbol)
    if value == ANY_SYMBOL or 0 <= value <= 255:
        return value
    raise ValueError(f"symbol must be 0..255 or ANY_SYMBOL, got {symbol}")


@dataclass(frozen=True, eq=False)
class NFA:
    """Immutable CSR automaton (build one with :class:`NFABuilder`).

    Equality and hashing are by identity (``eq=False``), as in the reference: the HIP
    backend keys its device-resident copies on the object, and numpy fields cannot be
    compared with ``==`` anyway.
    """

    num_states: int
    start_state: int
    accept: np.ndarray  # bool [num_states]
    sym_row_ptr: np.ndarray  # int32 [num_states + 1]
    sym_targets: np.ndarray  # int32 [nnz_sym]
    sym_symbols: np.ndarray  # int32 [nnz_sym], 0..255 or ANY_SYMBOL
    eps_row_ptr: np.ndarray  # int32 [num_states + 1]
    eps_targets: np.ndarray  # int32 [nnz_eps]

    @property
    def uses_any_symbol(self) -> bool:
        return bool(np.any(self.sym_symbols == ANY_SYMBOL))

    @property
    def num_sym_transitions(self) -> int:
        return int(self.sym_targets.size)

    @property
    def num_eps_transitions(self) -> int:
        return int(self.eps_targets.size)

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return (
            f"NFA(states={self.num_states}, start={self.start_state}, "
            f"accept={int(np.count_nonzero(self.accept))}, "
            f"sym={self.num_sym_transitions}, eps={self.num_eps_transitions})"
        )


class NFABuilder:
    """Collects states and edges, then freezes them into an :class:`NFA`."""

    def __init__(self) -> None:
        self._accepting: list[bool] = []
        self._edges: list[list[tuple[int, int]]] = []  # per state: (label, target)
        self._eps: list[list[int]] = []
        self._start = 0

    def add_state(self, accept: bool = False) -> int:
        self._accepting.append(bool(accept))
        self._edges.append([])
        self._eps.append([])
        return len(self._accepting) - 1

    def set_start(self, state: int) -> None:
        self._check(state)
        self._start = state

    def set_accept(self, state: int, accept: bool = True) -> None:
        self._check(state)
        self._accepting[state] = bool(accept)

    def add_transition(self, src: int, symbol: int | str, dst: int) -> None:
        self._check(src)
        self._check(dst)
        self._edges[src].append((_coerce_symbol(symbol), dst))

    def add_epsilon(self, src: int, dst: int) -> None:
        self._check(src)
        self._check(dst)
        self._eps[src].append(dst)

    def build(self) -> NFA:
        n = len(self._accepting)
        if n == 0:
            raise ValueError("cannot build an NFA with zero states")
        sym_deg = np.fromiter((len(e) for e in self._edges), dtype=np.int64, count=n)
        eps_deg = np.fromiter((len(e) for e in self._eps), dtype=np.int64, count=n)
        if int(sym_deg.sum()) > np.iinfo(np.int32).max or int(eps_deg.sum()) > np.iinfo(
            np.int32
        ).max:
            raise ValueError("edge count exceeds the int32 CSR offset range")
        sym_row_ptr = np.zeros(n + 1, dtype=np.int32)
        eps_row_ptr = np.zeros(n + 1, dtype=np.int32)
        sym_row_ptr[1:] = np.cumsum(sym_deg)
        eps_row_ptr[1:] = np.cumsum(eps_deg)
        flat = [edge for row in self._edges for edge in row]
        sym_symbols = np.array([e[0] for e in flat], dtype=np.int32)
        sym_targets = np.array([e[1] for e in flat], dtype=np.int32)
        eps_targets = np.array([t for row in self._eps for t in row], dtype=np.int32)
        return NFA(
            num_states=n,
            start_state=self._start,
            accept=np.array(self._accepting, dtype=bool),
            sym_row_ptr=sym_row_ptr,
            sym_targets=sym_targets,
            sym_symbols=sym_symbols,
            eps_row_ptr=eps_row_ptr,
            eps_targets=eps_targets,
        )

    def _check(self, state: int) -> None:
        if not 0 <= state < len(self._accepting):
            raise IndexError(f"state {state} out of range (0..{len(self._accepting) - 1})")
