"""The general NFA techniques on the GPU: worklist (ballot compaction), bitpacked (the
bit-parallel step on non-chain automata) and register (one lane per input).

Every verdict is compared with the C oracle (itself pinned against the upstream golden
corpus in the CPU suite). The cases cover the epsilon fold (cycles, chains, closures of the
start state), wildcards, one to 313 words of state set, LDS-staged and global edge lists,
global worklist scratch, ragged and unaligned batches, dense active sets, and automata
whose closure is too large to fold (they keep the unfolded worklist), and the bit-parallel
step's first and diagonal masks (chosen by the layout, forced, switched off; forward and
reversed chains).
"""

from __future__ import annotations

import dataclasses
import os
import random

import numpy as np
import pytest

from fixtures import random_nfa
from gpufsm import run_batch
from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.backends.hip.handles import DeviceNFA
from gpufsm.core.nfa import NFABuilder
from oracle import native

pytestmark = pytest.mark.gpu

GENERAL = ["worklist", "bitpacked", "register"]


def _check(nfa, data, off, tech):
    with DeviceNFA(nfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.NFA_TECHNIQUES[tech])
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    np.testing.assert_array_equal(f, wf, err_msg=tech)
    np.testing.assert_array_equal(n, wn, err_msg=tech)
    return wf, wn


def _ragged(rng, count, max_len, alphabet=None):
    lens = rng.integers(0, max_len + 1, size=count)
    total = int(lens.sum())
    if alphabet is None:
        data = rng.integers(0, 256, size=total, dtype=np.uint8)
    else:
        data = rng.choice(np.frombuffer(alphabet, dtype=np.uint8), size=total)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return data, off


@pytest.mark.parametrize("tech", GENERAL + ["auto"])
def test_plain_config_general_techniques(tech):
    """The plain drop-in NFA (1024 states, 4 random edges per state, 64 x 4 KiB) through
    every general technique that takes 1024 states."""
    nfa, (data, off) = W.workload_plain()
    if tech == "register":
        with pytest.raises(_lib.HIPError, match="512"):
            _check(nfa, data, off, tech)
        return
    _check(nfa, data, off, tech)


@pytest.mark.parametrize("n", [1, 40, 64, 65, 128, 200, 300, 400, 450, 512])
def test_register_every_word_count(n):
    """One to eight 64-bit words per lane, 4096 ragged inputs (unaligned starts, empty
    inputs), epsilon edges and wildcards; the same batch through the wave techniques."""
    rng = random.Random(n)
    nfa = random_nfa(rng, n, b"abcdefgh", max_sym=4, max_eps=2, p_accept=0.02, p_any=0.05)
    data, off = _ragged(np.random.default_rng(n), 4096, 300, b"abcdefghij")
    for tech in GENERAL:
        _check(nfa, data, off, tech)


def test_epsilon_cycles_and_long_epsilon_chains():
    """A 300-state epsilon chain (closure of the start is every state up to the first
    labelled edge), epsilon cycles, and an epsilon edge back to the start."""
    b = NFABuilder()
    s = [b.add_state(accept=(i == 299)) for i in range(300)]
    b.set_start(s[0])
    for i in range(0, 299):
        if i % 50 == 49:
            b.add_transition(s[i], ord("x"), s[i + 1])
        else:
            b.add_epsilon(s[i], s[i + 1])
    for i in range(10, 300, 37):
        b.add_epsilon(s[i], s[i - 9])  # cycles
    b.add_epsilon(s[150], s[0])
    b.add_transition(s[200], 256, s[120])
    nfa = b.build()
    rng = np.random.default_rng(4)
    data, off = _ragged(rng, 2000, 40, b"xyx")
    for tech in GENERAL:
        wf, wn = _check(nfa, data, off, tech)
    assert wf.any() and not wf.all()


def test_closure_too_large_to_fold_keeps_the_unfolded_worklist():
    """Every state reaches every other by epsilon edges (3000 states: 9 M closure entries,
    above the fold's cap): no folded layout, so bitpacked and register refuse the automaton
    and the worklist walks it with its own epsilon closure."""
    n = 3000
    b = NFABuilder()
    for i in range(n):
        b.add_state(accept=(i == n - 1))
    b.set_start(0)
    for i in range(n):
        b.add_epsilon(i, (i + 1) % n)
        b.add_transition(i, ord("a") + (i % 3), (i * 7) % n)
    nfa = b.build()
    with DeviceNFA(nfa) as dev:
        assert not dev.props.general_ok
    data, off = _ragged(np.random.default_rng(5), 64, 30, b"abcd")
    _check(nfa, data, off, "worklist")
    for tech in ("bitpacked", "register"):
        with pytest.raises(ValueError, match="cannot run"):
            run_batch(nfa, [b"ab"], backend="hip", technique=tech)


@pytest.mark.parametrize("n", [4000, 20000, 65536])
def test_gather_and_worklist_large_random(n):
    """Edge lists too large to stage (global reads), bitsets of 63 to 1024 words, and for
    the worklist the global list scratch."""
    rng = random.Random(n + 7)
    nfa = random_nfa(rng, n, bytes(range(32)), max_sym=6, max_eps=1, p_accept=0.002,
                     p_any=0.01)
    data, off = _ragged(np.random.default_rng(n), 96, 600, bytes(range(40)))
    for tech in ("worklist", "bitpacked"):
        _check(nfa, data, off, tech)


def test_gather_wave_per_input_on_large_batches():
    """Batches with four or more inputs per SIMD take the one-wave-per-input bit-parallel
    step (smaller batches take the four-waves-per-input block step, as every other case
    here): 5000 ragged inputs, ε edges and wildcards."""
    rng = random.Random(41)
    nfa = random_nfa(rng, 3000, bytes(range(24)), max_sym=5, max_eps=1, p_accept=0.003,
                     p_any=0.01)
    data, off = _ragged(np.random.default_rng(41), 5000, 120, bytes(range(30)))
    wf, _ = _check(nfa, data, off, "bitpacked")
    assert wf.any() and not wf.all()


def test_dense_active_non_chain_control():
    """A pattern union with optional positions (epsilon edges: not chain-structured), no
    accepting state, inputs of 95% pattern bytes: active sets in the hundreds on every byte,
    walked to the end by every technique."""
    pu = W.pattern_union(400, (5, 12), (16, 64), seed=21, accepting=False)
    nfa = W.optional_positions(pu, 0.3, seed=22)
    lens = np.full(48, 1500, dtype=np.int64)
    lens[::5] = 7
    data, off = W.pattern_inputs(pu, lens, 0.95, seed=23)
    with DeviceNFA(nfa) as dev:
        assert not dev.props.chain_ok and dev.props.general_ok
    for tech in ("worklist", "bitpacked"):
        _check(nfa, data, off, tech)


def test_optional_union_latching_matches_oracle():
    """The same construction with accepting last positions: first matches at every depth,
    ragged lengths, through the wave techniques and auto."""
    pu = W.pattern_union(250, (4, 14), (1, 24), seed=31)
    nfa = W.optional_positions(pu, 0.2, seed=32)
    lens = W.lognormal_lengths(500, 300, 1.0, 0, 5000, seed=33)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=34)
    for tech in ("worklist", "bitpacked", "auto"):
        wf, wn = _check(nfa, data, off, tech)
    assert wf.any() and (wn > 50).any()


def test_start_closure_accepts_gives_true_zero():
    b = NFABuilder()
    s0, s1, s2 = b.add_state(), b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_epsilon(s0, s1)
    b.add_epsilon(s1, s2)
    b.add_transition(s0, ord("a"), s0)
    nfa = b.build()
    for tech in GENERAL:
        got = [(r.accepted, r.match_len)
               for r in run_batch(nfa, [b"", b"abc"], backend="hip", technique=tech)]
        assert got == [(True, 0), (True, 0)]


def test_auto_resolution_for_general_automata():
    """auto: register for small general automata over large batches, the masked
    bit-parallel step for general automata whose start loops on every byte with many
    successors, worklist otherwise, chain for chain-structured ones."""
    rng = random.Random(3)
    small = random_nfa(rng, 100, b"ab", max_sym=2, max_eps=1)
    big = random_nfa(rng, 2000, b"ab", max_sym=2, max_eps=1)
    with DeviceNFA(small) as dev:
        assert dev.resolve(4096, 1 << 20) == _lib.NFA_REGISTER
        assert dev.resolve(16, 1 << 10) == _lib.NFA_WORKLIST
    with DeviceNFA(big) as dev:
        assert dev.resolve(4096, 1 << 20) == _lib.NFA_WORKLIST
        assert dev.resolve(1, 1, _lib.NFA_BITPACKED) == _lib.NFA_BITPACKED
    with DeviceNFA(W.pattern_union(50, seed=1).nfa) as dev:
        assert dev.resolve(4096, 1 << 20) == _lib.NFA_CHAIN
    # an unanchored union with optional positions: the start loops on every byte with
    # first masks, so the masked bit-parallel step
    pu = W.pattern_union(300, (4, 12), (8, 64), seed=51)
    with DeviceNFA(W.optional_positions(pu, 0.3, seed=52)) as dev:
        assert dev.resolve(16, 1 << 20) == _lib.NFA_BITPACKED
        assert dev.resolve(4096, 1 << 20) == _lib.NFA_BITPACKED


def _renumber_reversed(nfa):
    """The same automaton with every state but the start renumbered s -> n - s (start 0
    stays 0): successor edges s -> s + 1 of a pattern chain become s -> s - 1."""
    n = nfa.num_states
    assert nfa.start_state == 0
    new = np.arange(n)
    new[1:] = n - np.arange(1, n)
    inv = np.argsort(new)
    b = NFABuilder()
    for s in range(n):
        b.add_state(accept=bool(nfa.accept[inv[s]]))
    b.set_start(0)
    rp, sy, tg = nfa.sym_row_ptr, nfa.sym_symbols, nfa.sym_targets
    erp, et = nfa.eps_row_ptr, nfa.eps_targets
    for old in range(n):
        for k in range(rp[old], rp[old + 1]):
            b.add_transition(int(new[old]), int(sy[k]), int(new[tg[k]]))
        for k in range(erp[old], erp[old + 1]):
            b.add_epsilon(int(new[old]), int(new[et[k]]))
    return b.build()


@pytest.fixture
def diag_mode():
    """Sets GFSM_GATHER_DIAG for the handles created inside the test (read when a handle's
    layout is built), restoring the previous value afterwards."""
    old = os.environ.get("GFSM_GATHER_DIAG")

    def set_mode(v):
        os.environ["GFSM_GATHER_DIAG"] = v

    yield set_mode
    if old is None:
        os.environ.pop("GFSM_GATHER_DIAG", None)
    else:
        os.environ["GFSM_GATHER_DIAG"] = old


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("count", [48, 4200])
def test_gather_masks_match_oracle_every_mode(diag_mode, reverse, count):
    """Pattern unions with optional positions through the bit-parallel step with the masks
    the layout chooses (first masks for the looping start, diagonals +1/+2 or -1/-2), forced
    masks and no masks: equal verdicts, equal to the oracle. 48 inputs take the block step,
    4200 the wave-per-input step; reversed chains give negative diagonals."""
    pu = W.pattern_union(300, (4, 12), (8, 64), seed=51)
    nfa = W.optional_positions(pu, 0.3, seed=52)
    # 9 accepting pattern ends of 300: first matches hundreds of bytes in, sets stay dense
    acc = np.asarray(nfa.accept).copy()
    ends = np.flatnonzero(acc)
    acc[ends[np.random.default_rng(55).random(ends.size) >= 0.03]] = False
    nfa = dataclasses.replace(nfa, accept=acc)
    if reverse:
        nfa = _renumber_reversed(nfa)
    lens = W.lognormal_lengths(count, 200 if count > 100 else 1500, 1.0, 0, 4000, seed=53)
    data, off = W.pattern_inputs(pu, lens, 0.9, seed=54)
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    assert wf.any() and not wf.all()
    sizes = {}
    for mode in ("0", "1", "2"):
        diag_mode(mode)
        with DeviceNFA(nfa) as dev:
            sizes[mode] = dev.props.device_bytes
            f, n, _ = dev.run_host(data, off, _lib.NFA_BITPACKED)
        np.testing.assert_array_equal(f, wf, err_msg=f"mode {mode}")
        np.testing.assert_array_equal(n, wn, err_msg=f"mode {mode}")
    # masks were built (here smaller than the list edges they replace)
    assert sizes["1"] != sizes["0"] and sizes["2"] != sizes["0"]


def test_gather_forced_masks_on_random_automata(diag_mode):
    """Forced masks on random automata with epsilon edges and wildcards (offsets of either
    sign, starts without a loop), one to 313 words; the dense non-chain control prefix with
    the layout's masks."""
    diag_mode("2")
    for n in (3, 64, 65, 200, 1000, 20000):
        rng = random.Random(n + 100)
        nfa = random_nfa(rng, n, b"abcdefgh", max_sym=4, max_eps=2, p_accept=0.01,
                         p_any=0.05)
        data, off = _ragged(np.random.default_rng(n), 64, 400, b"abcdefghij")
        _check(nfa, data, off, "bitpacked")
    diag_mode("1")
    pu = W.pattern_union(8000, (5, 12), (16, 64), seed=0, total_states=65536, accepting=False)
    nfa = W.optional_positions(pu, 0.25, seed=1)
    data, off = W.pattern_inputs(pu, np.full(8, 2048, dtype=np.int64), 0.95, seed=3)
    _check(nfa, data, off, "bitpacked")
