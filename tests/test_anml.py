"""ANML loader: symbol-set grammar, the two start modes, and refusal of files it cannot
represent. Verdicts are checked with the oracle. Mirrors the cases of upstream
`tests/test_anml.py:23-180` with fixtures written for this suite.
"""

from __future__ import annotations

import pytest

from gpufsm.core.nfa import ANY_SYMBOL
from gpufsm.io import load_anml, parse_symbol_set
from oracle.reference import simulate

_PAIR = """<?xml version="1.0"?>
<automata-network id="pair">
  <state-transition-element id="p" symbol-set="[0x78]" start="start-of-data">
    <activate-on-match element="q"/>
  </state-transition-element>
  <state-transition-element id="q" symbol-set="[yz]">
    <report-on-match reportcode="7"/>
  </state-transition-element>
</automata-network>
"""

_ANYWHERE = """<?xml version="1.0"?>
<automata-network id="any">
  <state-transition-element id="k" symbol-set="[0x6B]" start="all-input">
    <report-on-match/>
  </state-transition-element>
</automata-network>
"""


def _write(tmp_path, text, name="a.anml"):
    p = tmp_path / name
    p.write_text(text)
    return p


def _wrap(body: str) -> str:
    return f"<anml><automata-network>{body}</automata-network></anml>"


def test_symbol_set_forms():
    assert parse_symbol_set("*") == {ANY_SYMBOL}
    assert parse_symbol_set("q") == {ord("q")}
    assert parse_symbol_set("[0x30-0x32]") == {0x30, 0x31, 0x32}
    assert parse_symbol_set("[xyz]") == {ord("x"), ord("y"), ord("z")}
    assert parse_symbol_set("[\\x41]") == parse_symbol_set("[0X41]") == {0x41}
    assert parse_symbol_set("[a-c0x7a]") == {ord("a"), ord("b"), ord("c"), 0x7A}
    neg = parse_symbol_set("[^0x00-0x0f]")
    assert len(neg) == 240 and 0x10 in neg and 0x0F not in neg
    assert parse_symbol_set("[^]") == set(range(256))


@pytest.mark.parametrize("text", ["[]", "[0x4]", "[0xG1]", "[z-a]", "ab"])
def test_malformed_symbol_sets_raise(text):
    with pytest.raises(ValueError):
        parse_symbol_set(text)


def test_start_of_data_pair(tmp_path):
    nfa = load_anml(_write(tmp_path, _PAIR))
    assert nfa.num_states == 3 + 2 and nfa.start_state == 0
    assert simulate(nfa, b"xy") == (True, 2)
    assert simulate(nfa, b"xz..") == (True, 2)
    assert simulate(nfa, b"x") == (False, 0)
    assert simulate(nfa, b"ax") == (False, 0)
    assert simulate(nfa, b"axy") == (False, 0)   # start-of-data: position 0 only
    assert simulate(nfa, b"") == (False, 0)


def test_all_input_matches_anywhere(tmp_path):
    for fold in (False, True):
        nfa = load_anml(_write(tmp_path, _ANYWHERE), fold_starts=fold)
        assert simulate(nfa, b"k") == (True, 1)
        assert simulate(nfa, b"aak") == (True, 3)
        assert simulate(nfa, b"aaa") == (False, 0)
        assert simulate(nfa, b"") == (False, 0)
    assert load_anml(_write(tmp_path, _ANYWHERE), fold_starts=True).num_states == 2


def test_start_of_data_blocks_folding(tmp_path):
    nfa = load_anml(_write(tmp_path, _PAIR), fold_starts=True)
    assert nfa.num_states == 5 and nfa.eps_row_ptr[-1] == 2   # kept the synthetic states
    assert simulate(nfa, b"axy") == (False, 0)


def test_port_suffix_and_chain(tmp_path):
    body = ('<state-transition-element id="a" symbol-set="a" start="all-input">'
            '<activate-on-match element="b:in"/></state-transition-element>'
            '<state-transition-element id="b" symbol-set="[b-d]">'
            '<activate-on-match element="c"/></state-transition-element>'
            '<state-transition-element id="c" symbol-set="*"><report-on-match/>'
            '</state-transition-element>')
    for fold in (False, True):
        nfa = load_anml(_write(tmp_path, _wrap(body)), fold_starts=fold)
        assert simulate(nfa, b"zzac!") == (True, 5)
        assert simulate(nfa, b"zzaa") == (False, 0)
        assert simulate(nfa, b"abab") == (True, 3)
        assert simulate(nfa, b"aeae") == (False, 0)
        assert simulate(nfa, b"xadd") == (True, 4)


@pytest.mark.parametrize("body, msg", [
    ('<state-transition-element id="s" symbol-set="a" start="all-input"/>'
     '<or id="g"><report-on-match/></or>', "unsupported"),
    ('<state-transition-element id="s" symbol-set="a"/>'
     '<state-transition-element id="s" symbol-set="b"/>', "duplicate"),
    ('<state-transition-element symbol-set="a" start="all-input"/>', "without an 'id'"),
    ('<state-transition-element id="s" symbol-set="a" start="all-input">'
     '<activate-on-match element="nope"/></state-transition-element>',
     "activates unknown element"),
    ('<state-transition-element id="s" symbol-set="a"><activate-on-match/>'
     '</state-transition-element>', "without 'element'"),
    ('<state-transition-element id="s" symbol-set="[0x4]"/>', "STE 's'"),
    ('', "no state-transition-element"),
])
def test_refuses_what_it_cannot_represent(tmp_path, body, msg):
    with pytest.raises(ValueError, match=msg):
        load_anml(_write(tmp_path, _wrap(body)))
