"""Packed-bitset NFA spec, Python restatement. TEST INFRASTRUCTURE ONLY.

Restates `src/gpufsm/core/bitmap.py:20-120` (``_build_masks``, ``_epsilon_closure``,
``BitmapProgram``, ``simulate_bitmap``): the active set is one Python int used as a
bitset; a step ORs per-state target masks for the byte and for ANY edges, then closes
over epsilon masks by frontier. Verdicts are the same as :func:`oracle.reference.simulate`
(pinned by the golden corpus replay in ``tests/test_oracle.py``).
"""

from __future__ import annotations

from oracle.reference import ANY_SYMBOL, NFA


def _masks(nfa: NFA) -> tuple[list[dict[int, int]], list[int], list[int], int]:
    n = nfa.num_states
    by_byte: list[dict[int, int]] = [{} for _ in range(n)]
    wild = [0] * n
    eps = [0] * n
    row, tgt, lab = nfa.sym_row_ptr, nfa.sym_targets, nfa.sym_symbols
    for s in range(n):
        for k in range(int(row[s]), int(row[s + 1])):
            bit = 1 << int(tgt[k])
            label = int(lab[k])
            if label == ANY_SYMBOL:
                wild[s] |= bit
            else:
                by_byte[s][label] = by_byte[s].get(label, 0) | bit
    erow, etgt = nfa.eps_row_ptr, nfa.eps_targets
    for s in range(n):
        for k in range(int(erow[s]), int(erow[s + 1])):
            eps[s] |= 1 << int(etgt[k])
    acc = 0
    for s in range(n):
        if nfa.accept[s]:
            acc |= 1 << s
    return by_byte, wild, eps, acc


def _bits(mask: int):
    while mask:
        low = mask & -mask
        yield low.bit_length() - 1
        mask ^= low


def _close(active: int, eps: list[int]) -> int:
    frontier = active
    while frontier:
        grown = 0
        for s in _bits(frontier):
            grown |= eps[s]
        grown &= ~active
        active |= grown
        frontier = grown
    return active


class BitmapProgram:
    """Masks built once per automaton; :meth:`run` simulates one input."""

    def __init__(self, nfa: NFA) -> None:
        self.by_byte, self.wild, self.eps, self.accept = _masks(nfa)
        self.start = _close(1 << nfa.start_state, self.eps)

    def run(self, input_bytes: bytes) -> tuple[bool, int]:
        active = self.start
        if active & self.accept:
            return True, 0
        for i, byte in enumerate(bytes(input_bytes)):
            nxt = 0
            for s in _bits(active):
                nxt |= self.wild[s] | self.by_byte[s].get(byte, 0)
            if not nxt:
                break
            active = _close(nxt, self.eps)
            if active & self.accept:
                return True, i + 1
        return False, 0


def simulate_bitmap(nfa: NFA, input_bytes: bytes) -> tuple[bool, int]:
    return BitmapProgram(nfa).run(input_bytes)
