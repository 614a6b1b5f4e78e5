"""Pins the CPU oracle before anything is compared against it.

Each oracle function is checked against data or mathematics that does not come from the
oracle itself:

* the upstream golden corpus (516 NFA + 36 DFA verdicts produced by the upstream oracle);
* the upstream closed-form cases (`tests/test_reference.py:21-46`, `tests/test_dfa.py:23-29`);
* Python's ``re`` engine on the two example languages (shortest accepted prefix);
* a linear-algebra formulation of subset construction (boolean matrices for moves and a
  transitive-closure matrix for epsilon edges), on random NFAs;
* the DFA-as-NFA embedding, and the C restatement in both of its modes.

A plausible slip (closure skipped after a step, a wildcard ignored, ``match_len`` off by
one, no stop on an empty set) fails at least one of these.
"""

from __future__ import annotations

import random
import re

import numpy as np
import pytest

from fixtures import EXAMPLES, dfa_from_json, nfa_from_json, random_nfa
from gpufsm.core.dfa import DFABuilder, random_dfa
from gpufsm.core.nfa import ANY_SYMBOL, NFABuilder
from gpufsm.core.packing import pack_bytes
from oracle import native
from oracle.bitmap import simulate_bitmap
from oracle.reference import dfa_as_nfa, examined_nfa, simulate, simulate_dfa


def test_golden_corpus_is_non_trivial(golden):
    assert len(golden["nfa_cases"]) == 516
    assert len(golden["dfa_cases"]) == 36
    assert any(c["accepted"] for c in golden["nfa_cases"])
    assert any(not c["accepted"] for c in golden["nfa_cases"])


def test_golden_nfa_python_and_bitmap(golden):
    for case in golden["nfa_cases"]:
        nfa = nfa_from_json(case["nfa"])
        data = bytes.fromhex(case["input_hex"])
        want = (case["accepted"], case["match_len"])
        assert simulate(nfa, data) == want, case["id"]
        assert simulate_bitmap(nfa, data) == want, case["id"]


def test_golden_dfa_python(golden):
    for case in golden["dfa_cases"]:
        dfa = dfa_from_json(case["dfa"])
        data = bytes.fromhex(case["input_hex"])
        assert simulate_dfa(dfa, data) == (case["accepted"], case["match_len"]), case["id"]


@pytest.mark.parametrize("indexed", [False, True])
def test_golden_nfa_c_oracle(golden, indexed):
    for case in golden["nfa_cases"]:
        nfa = nfa_from_json(case["nfa"])
        data = bytes.fromhex(case["input_hex"])
        buf, off = pack_bytes([data])
        f, n, ex = native.nfa_batch(nfa, buf, off, indexed=indexed, threads=1)
        assert (bool(f[0]), int(n[0])) == (case["accepted"], case["match_len"]), case["id"]
        assert int(ex[0]) == examined_nfa(nfa, data), case["id"]


def test_golden_dfa_c_oracle(golden):
    cases = golden["dfa_cases"]
    for case in cases:
        dfa = dfa_from_json(case["dfa"])
        buf, off = pack_bytes([bytes.fromhex(case["input_hex"])])
        f, n, _ = native.dfa_batch(dfa, buf, off, threads=2)
        assert (bool(f[0]), int(n[0])) == (case["accepted"], case["match_len"]), case["id"]


def test_closed_forms():
    b = NFABuilder()
    b.set_start(b.add_state(accept=True))
    assert simulate(b.build(), b"anything") == (True, 0)
    b = NFABuilder()
    s0, s1, s2 = b.add_state(), b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, "x", s1)
    b.add_epsilon(s1, s2)
    assert simulate(b.build(), b"x") == (True, 1)
    d = DFABuilder()
    q0, q1, q2 = d.add_state(), d.add_state(), d.add_state(accept=True)
    d.set_start(q0)
    d.add_transition(q0, ord("a"), q1)
    d.add_transition(q1, ord("b"), q2)
    dfa = d.build()
    assert simulate_dfa(dfa, b"ab") == (True, 2)
    assert simulate_dfa(dfa, b"abc") == (True, 2)
    assert simulate_dfa(dfa, b"ax") == (False, 0)
    assert simulate_dfa(dfa, b"a") == (False, 0)
    assert simulate_dfa(dfa, b"") == (False, 0)


def test_examined_closed_forms():
    """Bytes examined, from the automaton's shape alone (Python and C oracles):

    * an anchored literal ``abc``: an input whose set dies on byte k (0-based) examined k + 1;
      a match examined its match length; an input that ends first examined all of it;
    * an accepting start: 0 for every input;
    * a start with an ANY self-loop and nothing accepting: every byte of every input;
    * an epsilon edge into a state looping on ``a``: ``a`` * m then ``b`` dies on byte m.
    """
    b = NFABuilder()
    s = [b.add_state(accept=(i == 3)) for i in range(4)]
    b.set_start(s[0])
    for i, c in enumerate("abc"):
        b.add_transition(s[i], c, s[i + 1])
    lit = b.build()
    cases = {b"xyz": 1, b"ayz": 2, b"abz": 3, b"abcz": 3, b"abc": 3, b"ab": 2, b"": 0,
             b"ab" + b"q" * 50: 3}
    b = NFABuilder()
    b.set_start(b.add_state(accept=True))
    acc = b.build()
    b = NFABuilder()
    b.set_start(b.add_state())
    b.add_transition(0, ANY_SYMBOL, 0)
    b.add_transition(0, "a", b.add_state())
    loop = b.build()
    b = NFABuilder()
    t0, t1 = b.add_state(), b.add_state()
    b.set_start(t0)
    b.add_epsilon(t0, t1)
    b.add_transition(t1, "a", t1)
    eps = b.build()
    rng = random.Random(5)
    for m in (0, 1, 7, 300):
        cases_eps = b"a" * m + b"b" + b"a" * 9
        assert examined_nfa(eps, cases_eps) == m + 1
    for nfa, want in ((lit, cases),
                      (acc, {x: 0 for x in cases}),
                      (loop, {bytes(rng.randrange(256) for _ in range(n)): n
                              for n in (0, 1, 5, 999)})):
        batch = list(want)
        assert [examined_nfa(nfa, x) for x in batch] == [want[x] for x in batch]
        data, off = pack_bytes(batch)
        for indexed in (False, True):
            _, _, ex = native.nfa_batch(nfa, data, off, indexed=indexed)
            assert list(ex) == [want[x] for x in batch], indexed
    data, off = pack_bytes([b"a" * m + b"b" for m in (0, 3, 64)])
    _, _, ex = native.nfa_batch(eps, data, off)
    assert list(ex) == [1, 4, 65]


_EXAMPLE_REGEX = {"ab_star_c_plus_d": rb"(?:ab*c+)+d", "a_star_b_c_opt_d": rb"a*bc?d"}


@pytest.mark.parametrize("name", list(EXAMPLES))
def test_examples_against_python_re(name):
    """Latch-first-match = the shortest prefix in the language, found here by ``re``."""
    nfa, labelled = EXAMPLES[name]()
    pattern = re.compile(_EXAMPLE_REGEX[name])
    rng = random.Random(5)
    inputs = [d for d, _ in labelled] + [
        bytes(rng.choice(b"abcd") for _ in range(rng.randint(0, 12))) for _ in range(400)
    ]
    for data, want_acc in labelled:
        assert simulate(nfa, data)[0] == want_acc
    for data in inputs:
        want = next(((True, k) for k in range(len(data) + 1) if pattern.fullmatch(data[:k])),
                    (False, 0))
        assert simulate(nfa, data) == want, data


def _matrix_oracle(nfa, data: bytes) -> tuple[bool, int]:
    """Subset construction in boolean linear algebra.

    ``E*`` is the reflexive-transitive closure of the epsilon adjacency matrix (repeated
    squaring); a step is ``S' = E*ᵀ (M_bᵀ S)`` with ``M_b`` the adjacency of edges labelled
    ``b`` or ANY. The empty-set stop is implied: an empty set never accepts again.
    """
    n = nfa.num_states
    eps = np.eye(n, dtype=np.int64)
    for s in range(n):
        for k in range(nfa.eps_row_ptr[s], nfa.eps_row_ptr[s + 1]):
            eps[s, nfa.eps_targets[k]] = 1
    for _ in range(max(1, int(np.ceil(np.log2(max(n, 2)))) + 1)):
        eps = np.minimum(eps @ eps, 1)
    acc = np.asarray(nfa.accept, dtype=np.int64)
    state = np.zeros(n, dtype=np.int64)
    state[nfa.start_state] = 1
    state = np.minimum(eps.T @ state, 1)
    if state @ acc:
        return True, 0
    for i, byte in enumerate(data):
        move = np.zeros((n, n), dtype=np.int64)
        for s in range(n):
            for k in range(nfa.sym_row_ptr[s], nfa.sym_row_ptr[s + 1]):
                if nfa.sym_symbols[k] in (byte, ANY_SYMBOL):
                    move[s, nfa.sym_targets[k]] = 1
        state = np.minimum(eps.T @ np.minimum(move.T @ state, 1), 1)
        if state @ acc:
            return True, i + 1
        if not state.any():
            return False, 0
    return False, 0


def test_random_nfas_against_matrix_subset_construction():
    rng = random.Random(1234)
    for _ in range(300):
        nfa = random_nfa(rng, rng.randint(1, 9), b"abc", max_sym=3, max_eps=2)
        data = bytes(rng.choice(b"abcd") for _ in range(rng.randint(0, 14)))
        assert simulate(nfa, data) == _matrix_oracle(nfa, data)


def test_dfa_equals_its_nfa_embedding():
    rng = random.Random(3)
    for n in (1, 2, 7, 40):
        dfa = random_dfa(n, accept_prob=0.1, seed=n)
        nfa = dfa_as_nfa(dfa)
        for _ in range(60):
            data = bytes(rng.randrange(256) for _ in range(rng.randint(0, 30)))
            assert simulate_dfa(dfa, data) == simulate(nfa, data)


def test_c_oracle_matches_python_on_random_batches():
    rng = random.Random(99)
    for trial in range(40):
        nfa = random_nfa(rng, rng.randint(1, 70), b"abcd", max_sym=4, max_eps=2,
                         p_accept=0.05)
        batch = [bytes(rng.choice(b"abcde") for _ in range(rng.randint(0, 40)))
                 for _ in range(rng.randint(0, 30))]
        data, off = pack_bytes(batch)
        for indexed in (False, True):
            f, n, ex = native.nfa_batch(nfa, data, off, indexed=indexed, threads=3)
            got = [(bool(a), int(b)) for a, b in zip(f, n)]
            assert got == [simulate(nfa, d) for d in batch], trial
            assert list(ex) == [examined_nfa(nfa, d) for d in batch], trial


def test_c_dfa_oracle_examined_counts():
    dfa = random_dfa(16, accept_prob=0.1, seed=4)
    rng = random.Random(8)
    batch = [bytes(rng.randrange(256) for _ in range(rng.randint(0, 50))) for _ in range(100)]
    data, off = pack_bytes(batch)
    f, n, ex = native.dfa_batch(dfa, data, off, threads=4)
    for i, d in enumerate(batch):
        acc, ml = simulate_dfa(dfa, d)
        assert (bool(f[i]), int(n[i])) == (acc, ml)
        assert int(ex[i]) == (ml if acc else len(d))
