"""Two small hand-built NFAs with labelled inputs, for smoke checks (``gpufsm verify``).

Restates `src/gpufsm/examples.py:15-65`: the same two automata (edges, epsilon edges,
accepting states) and the same labelled inputs.
"""

from __future__ import annotations

from .core.nfa import NFABuilder


def ab_star_c_plus_d():
    """Roughly ``(a b* c+)+ d`` with an epsilon back edge."""
    b = NFABuilder()
    s0, s1, s2, s3 = b.add_state(), b.add_state(), b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, "a", s1)
    b.add_transition(s1, "b", s1)
    b.add_transition(s1, "c", s2)
    b.add_epsilon(s2, s0)
    b.add_transition(s2, "c", s2)
    b.add_transition(s2, "d", s3)
    labelled = [(b"abcd", True), (b"acd", True), (b"abcabcabcd", True), (b"acacd", True),
                (b"cd", False), (b"d", False), (b"", False), (b"abc", False)]
    return b.build(), labelled


def a_star_b_c_opt_d():
    """``a* b (c|epsilon) d``."""
    b = NFABuilder()
    s0, s1, s2, s3 = b.add_state(), b.add_state(), b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, "a", s0)
    b.add_transition(s0, "b", s1)
    b.add_transition(s1, "c", s2)
    b.add_epsilon(s1, s2)
    b.add_transition(s2, "d", s3)
    labelled = [(b"abd", True), (b"bcd", True), (b"bd", True), (b"aaaaaaabd", True),
                (b"aaabcd", True), (b"ac", False), (b"bb", False), (b"", False)]
    return b.build(), labelled


EXAMPLES = {"ab_star_c_plus_d": ab_star_c_plus_d, "a_star_b_c_opt_d": a_star_b_c_opt_d}
