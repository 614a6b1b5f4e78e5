"""Test generators and (de)serialisation of the golden corpus.

The example automata live in ``gpufsm.examples``; the JSON layout is the one of the vendored
``tests/golden/golden.json`` (`tests/generate_golden.py:50-91`).
"""

from __future__ import annotations

import numpy as np

from gpufsm.core.dfa import DFA
from gpufsm.core.nfa import NFA, NFABuilder
from gpufsm.examples import EXAMPLES  # noqa: F401  (re-exported for the tests)




def nfa_from_json(d: dict) -> NFA:
    return NFA(
        num_states=int(d["num_states"]),
        start_state=int(d["start_state"]),
        accept=np.array(d["accept"], dtype=bool),
        sym_row_ptr=np.array(d["sym_row_ptr"], dtype=np.int32),
        sym_targets=np.array(d["sym_targets"], dtype=np.int32),
        sym_symbols=np.array(d["sym_symbols"], dtype=np.int32),
        eps_row_ptr=np.array(d["eps_row_ptr"], dtype=np.int32),
        eps_targets=np.array(d["eps_targets"], dtype=np.int32),
    )


def dfa_from_json(d: dict) -> DFA:
    return DFA(
        num_states=int(d["num_states"]),
        start_state=int(d["start_state"]),
        accept=np.array(d["accept"], dtype=bool),
        trans=np.array(d["trans"], dtype=np.int32),
    )


def random_nfa(rng, n: int, alphabet: bytes, *, max_sym: int = 2, max_eps: int = 1,
               p_accept: float = 0.2, p_any: float = 0.1) -> NFA:
    """Small random NFA with wildcard and epsilon edges (test generator)."""
    b = NFABuilder()
    for _ in range(n):
        b.add_state(accept=rng.random() < p_accept)
    b.set_start(rng.randrange(n))
    for s in range(n):
        for _ in range(rng.randint(0, max_sym)):
            sym = 256 if rng.random() < p_any else rng.choice(alphabet)
            b.add_transition(s, sym, rng.randrange(n))
        for _ in range(rng.randint(0, max_eps)):
            b.add_epsilon(s, rng.randrange(n))
    return b.build()


def chain_nfa(rng, num_chains: int, len_range=(1, 12), width_range=(1, 8), *,
              self_loop: str = "any", p_accept: float = 0.05, p_last: float = 0.5,
              p_shared: float = 0.1, alphabet: bytes | None = None):
    """Random chain-structured NFA (test generator) and the bytes its classes use.

    Every non-start state lies on one chain hanging off the start state; state ids are
    shuffled and the start is not state 0. ``self_loop`` is the start state's self-loop:
    ``"any"`` (one wildcard edge), ``"all"`` (256 literal edges), ``"partial"`` (a random
    subset) or ``"none"``. A fraction ``p_shared`` of inner chain positions is also entered
    straight from the start on the same class, so a state can be reached both ways.
    """
    alphabet = bytes(range(256)) if alphabet is None else alphabet
    lens = [rng.randint(*len_range) for _ in range(num_chains)]
    n = 1 + sum(lens)
    ids = list(range(n))
    rng.shuffle(ids)
    b = NFABuilder()
    for _ in range(n):
        b.add_state()
    start = ids[0]
    b.set_start(start)
    if self_loop == "any":
        b.add_transition(start, 256, start)
    elif self_loop == "all":
        for c in range(256):
            b.add_transition(start, c, start)
    elif self_loop == "partial":
        for c in rng.sample(range(256), rng.randint(1, 255)):
            b.add_transition(start, c, start)
    used = set()
    k = 1
    for length in lens:
        prev = start
        for pos in range(length):
            s = ids[k]
            k += 1
            width = min(rng.randint(*width_range), len(alphabet))
            cls = rng.sample(list(alphabet), width)
            used.update(cls)
            shared = prev != start and rng.random() < p_shared
            for c in cls:
                b.add_transition(prev, c, s)
                if shared:
                    b.add_transition(start, c, s)
            if rng.random() < p_accept or pos == length - 1 and rng.random() < p_last:
                b.set_accept(s)
            prev = s
    return b.build(), bytes(sorted(used))
