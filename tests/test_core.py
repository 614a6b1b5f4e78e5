"""Core data model, packing and builders (CPU). Mirrors the upstream CPU suites
`tests/test_nfa.py`, `tests/test_packing.py:33-96` and `tests/test_validation.py:29-101`."""

from __future__ import annotations

import numpy as np
import pytest

from gpufsm import ANY_SYMBOL, DFABuilder, NFABuilder, random_dfa
from gpufsm.core.dfa import ALPHABET
from gpufsm.core.packing import (
    WORD_BITS,
    accept_bitmask,
    pack_accept,
    pack_bytes,
    pack_inputs,
    symbols,
    words_for,
)


def _nfa(accepting: set[int], n: int = 8):
    b = NFABuilder()
    for i in range(n):
        b.add_state(accept=i in accepting)
    b.set_start(0)
    return b.build()


def test_builder_csr_shapes():
    b = NFABuilder()
    s0, s1 = b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, "a", s1)
    b.add_transition(s0, ANY_SYMBOL, s1)
    b.add_epsilon(s1, s0)
    nfa = b.build()
    assert nfa.num_states == 2 and nfa.start_state == s0
    assert nfa.accept.tolist() == [False, True]
    assert nfa.sym_row_ptr.tolist() == [0, 2, 2]
    assert nfa.eps_row_ptr.tolist() == [0, 0, 1]
    assert nfa.sym_targets.tolist() == [1, 1]
    assert nfa.sym_symbols.tolist() == [ord("a"), ANY_SYMBOL]
    assert nfa.uses_any_symbol is True
    assert nfa.num_sym_transitions == 2 and nfa.num_eps_transitions == 1
    assert nfa.sym_row_ptr.dtype == np.int32


def test_builder_keeps_insertion_order_per_state():
    b = NFABuilder()
    s = [b.add_state() for _ in range(3)]
    b.add_transition(s[1], "b", s[2])
    b.add_transition(s[0], "a", s[1])
    b.add_transition(s[1], "c", s[0])
    nfa = b.build()
    assert nfa.sym_row_ptr.tolist() == [0, 1, 3, 3]
    assert nfa.sym_symbols.tolist() == [ord("a"), ord("b"), ord("c")]
    assert nfa.sym_targets.tolist() == [1, 2, 0]


def test_nfa_identity_semantics():
    a, b = _nfa({1}), _nfa({1})
    assert a != b and a == a
    assert len({a, b}) == 2


def test_empty_builder_rejected():
    with pytest.raises(ValueError):
        NFABuilder().build()
    with pytest.raises(ValueError):
        DFABuilder().build()


def test_symbol_coercion_rules():
    b = NFABuilder()
    s0 = b.add_state()
    with pytest.raises(ValueError, match="code point 256"):
        b.add_transition(s0, chr(256), s0)
    with pytest.raises(ValueError, match="over bytes"):
        b.add_transition(s0, "€", s0)
    with pytest.raises(ValueError):
        b.add_transition(s0, "ab", s0)
    with pytest.raises(ValueError):
        b.add_transition(s0, 257, s0)
    with pytest.raises(ValueError):
        b.add_transition(s0, -1, s0)
    b.add_transition(s0, chr(255), s0)
    b.add_transition(s0, ANY_SYMBOL, s0)
    assert b.build().sym_symbols.tolist() == [255, ANY_SYMBOL]


def test_builder_bounds():
    b = NFABuilder()
    b.add_state()
    with pytest.raises(IndexError):
        b.add_transition(0, "a", 5)
    with pytest.raises(IndexError):
        b.set_start(3)
    d = DFABuilder()
    d.add_state()
    d.add_state()
    with pytest.raises(IndexError):
        d.set_start(7)
    with pytest.raises(IndexError):
        d.add_transition(0, ord("a"), 99)
    with pytest.raises(IndexError):
        d.add_transition(-1, ord("a"), 0)
    with pytest.raises(ValueError, match="0..255"):
        d.add_transition(0, ALPHABET, 0)


def test_dfa_builder_dead_state():
    d = DFABuilder()
    s0, s1 = d.add_state(), d.add_state(accept=True)
    d.set_start(s0)
    d.add_transition(s0, ord("a"), s1)
    dfa = d.build()
    assert dfa.num_states == 3
    assert dfa.trans.size == 3 * 256 and dfa.trans.dtype == np.int32
    assert dfa.trans[0 * 256 + ord("a")] == 1
    assert dfa.trans[0 * 256 + ord("b")] == 2
    assert (dfa.trans[2 * 256:3 * 256] == 2).all()
    assert dfa.accept.tolist() == [False, True, False]


def test_random_dfa_recipe():
    """Draw order of `src/gpufsm/core/dfa.py:44-57`: targets, then accept floats."""
    dfa = random_dfa(64, accept_prob=0.3, seed=1)
    rng = np.random.default_rng(1)
    trans = rng.integers(0, 64, size=64 * 256, dtype=np.int32)
    acc = rng.random(64) < 0.3
    acc[0] = False
    assert (dfa.trans == trans).all() and (dfa.accept == acc).all()
    assert dfa.start_state == 0 and not dfa.accept[0]
    with pytest.raises(ValueError):
        random_dfa(0)


@pytest.mark.parametrize(("n", "words"), [(1, 1), (63, 1), (64, 1), (65, 2), (128, 2),
                                          (129, 3), (500, 8)])
def test_words_for(n, words):
    assert words_for(n) == words


def test_symbols_int32():
    out = symbols(b"\x00A\xff")
    assert out.dtype == np.int32 and out.tolist() == [0, 65, 255]


def test_pack_inputs_and_bytes():
    inputs = [b"ab", b"", b"xyz"]
    data, offsets = pack_inputs(inputs)
    assert offsets.tolist() == [0, 2, 2, 5] and data.dtype == np.int32
    data8, off8 = pack_bytes(inputs)
    assert data8.dtype == np.uint8 and off8.dtype == np.int64
    assert off8.tolist() == [0, 2, 2, 5]
    for i, s in enumerate(inputs):
        assert bytes(data8[off8[i]:off8[i + 1]]) == s
    d, o = pack_inputs([b"", b""])
    assert d.size == 0 and o.tolist() == [0, 0, 0]
    d, o = pack_bytes([])
    assert d.size == 0 and o.tolist() == [0]


def test_accept_packing():
    assert accept_bitmask(_nfa({0, 3, 7})) == 0b10001001
    n = 200
    acc = {0, 31, 32, 63, 64, 127, 128, 199}
    words = pack_accept(_nfa(acc, n=n))
    assert words.size == words_for(n)
    assert {s for s in range(n) if int(words[s // WORD_BITS]) >> (s % WORD_BITS) & 1} == acc
    w8 = pack_accept(_nfa({0}), nwords=8, dtype=np.int64)
    assert w8.dtype == np.int64 and w8.size == 8 and int(w8[0]) == 1 and not w8[1:].any()
