"""HIP backend parity with the CPU oracle: every verdict bit-exact.

Run on an MI355X (``-m gpu``). The HIP backend must be available there: a missing or
broken library fails these tests instead of skipping them. Cases mirror the upstream GPU
suites (`tests/test_gpu_backends.py:36-137`, `tests/test_dfa_gpu.py:34-63`) and add the
golden corpus, ragged batches, unaligned offsets and the pattern-union layouts.
"""

from __future__ import annotations

import random

import numpy as np
import pytest

from fixtures import EXAMPLES, chain_nfa, dfa_from_json, nfa_from_json, random_nfa
from gpufsm import Backend, is_available, run, run_batch
from gpufsm import workloads as W
from gpufsm.core.dfa import random_dfa
from gpufsm.core.nfa import NFABuilder
from gpufsm.core.packing import pack_bytes
from gpufsm.core.registry import unavailable_reason
from gpufsm.backends.hip.handles import DeviceDFA, DeviceNFA, make_tuning
from gpufsm.backends.hip import _lib
from oracle import native
from oracle.reference import simulate, simulate_dfa

pytestmark = pytest.mark.gpu

NFA_TECHS = ["auto", "worklist", "bitpacked", "register"]  # every one runs any small NFA
DFA_TECHS = ["auto", "gather", "lds", "split"]


@pytest.fixture(scope="module", autouse=True)
def _require_hip():
    assert is_available(Backend.HIP), f"HIP backend unavailable: {unavailable_reason(Backend.HIP)}"


def _verdicts(results):
    return [(r.accepted, r.match_len) for r in results]


def _oracle_nfa(nfa, batch):
    data, off = pack_bytes(batch)
    f, n, _ = native.nfa_batch(nfa, data, off)
    return [(bool(a), int(b)) for a, b in zip(f, n)]


def _oracle_dfa(dfa, batch):
    data, off = pack_bytes(batch)
    f, n, _ = native.dfa_batch(dfa, data, off)
    return [(bool(a), int(b)) for a, b in zip(f, n)]


def test_device_is_gfx950():
    assert _lib.device_arch(0).startswith("gfx950")


@pytest.mark.parametrize("tech", NFA_TECHS)
def test_golden_nfa_corpus(golden, tech):
    for case in golden["nfa_cases"]:
        nfa = nfa_from_json(case["nfa"])
        data = bytes.fromhex(case["input_hex"])
        res = run(nfa, data, backend="hip", technique=tech)
        assert (res.accepted, res.match_len) == (case["accepted"], case["match_len"]), case["id"]


@pytest.mark.parametrize("tech", ["auto", "gather", "split"])
def test_golden_dfa_corpus(golden, tech):
    for case in golden["dfa_cases"]:
        dfa = dfa_from_json(case["dfa"])
        data = bytes.fromhex(case["input_hex"])
        res = run(dfa, data, backend="hip", technique=tech)
        assert (res.accepted, res.match_len) == (case["accepted"], case["match_len"]), case["id"]


def test_golden_dfa_corpus_lds(golden):
    """All golden DFAs have at most 8 states, so the LDS table applies to every one."""
    for case in golden["dfa_cases"]:
        dfa = dfa_from_json(case["dfa"])
        data = bytes.fromhex(case["input_hex"])
        res = run(dfa, data, backend="hip", technique="lds")
        assert (res.accepted, res.match_len) == (case["accepted"], case["match_len"]), case["id"]


@pytest.mark.parametrize("tech", NFA_TECHS)
def test_examples(tech):
    for name, make in EXAMPLES.items():
        nfa, labelled = make()
        batch = [d for d, _ in labelled]
        got = _verdicts(run_batch(nfa, batch, backend="hip", technique=tech))
        assert got == [simulate(nfa, d) for d in batch], name


@pytest.mark.parametrize("tech", NFA_TECHS)
def test_fuzz_batches(tech):
    rng = random.Random(11)
    for trial in range(30):
        nfa = random_nfa(rng, rng.randint(1, 64), b"abcd", max_sym=3, max_eps=2, p_any=0.05)
        batch = [bytes(rng.choice(b"abcd") for _ in range(rng.randint(0, 24)))
                 for _ in range(rng.randint(1, 40))]
        got = _verdicts(run_batch(nfa, batch, backend="hip", technique=tech))
        assert got == [simulate(nfa, d) for d in batch], trial


@pytest.mark.parametrize("tech", ["worklist", "bitpacked", "register"])
def test_multiword_and_large_automata(tech):
    """Random epsilon-NFAs of 65 to 20,000 states (multi-word sets: one to 313 words; the
    register technique runs up to 512 states)."""
    rng = random.Random(23)
    for n in (65, 130, 300, 500, 1000, 5000, 20000):
        nfa = random_nfa(rng, n, b"abcde", max_sym=3, max_eps=1, p_accept=0.01, p_any=0.02)
        batch = [bytes(rng.choice(b"abcdef") for _ in range(rng.randint(0, 200)))
                 for _ in range(50)]
        if tech == "register" and n > 512:
            with pytest.raises(ValueError, match="512"):
                run_batch(nfa, batch, backend="hip", technique=tech)
            continue
        got = _verdicts(run_batch(nfa, batch, backend="hip", technique=tech))
        assert got == _oracle_nfa(nfa, batch), n


def test_edge_cases():
    b = NFABuilder()
    b.set_start(b.add_state(accept=True))
    acc = b.build()
    assert _verdicts(run_batch(acc, [b"", b"xyz"], backend="hip")) == [(True, 0), (True, 0)]
    assert run_batch(acc, [], backend="hip") == []
    nfa, _ = EXAMPLES["a_star_b_c_opt_d"]()
    assert _verdicts(run_batch(nfa, [b"", b"", b""], backend="hip")) == [(False, 0)] * 3
    dfa = random_dfa(8, accept_prob=0.3, seed=2)
    assert run_batch(dfa, [], backend="hip") == []
    assert _verdicts(run_batch(dfa, [b""], backend="hip")) == [simulate_dfa(dfa, b"")]


def test_inputs_longer_than_int32_are_refused():
    """match_len is int32, so the host form refuses an input of 2^31 bytes or more with
    EINVAL. The offsets are checked before the data is read, so a 1-byte buffer will do."""
    import ctypes

    nfa, _ = EXAMPLES["ab_star_c_plus_d"]()
    data = np.zeros(1, dtype=np.uint8)
    off = np.array([0, 1, 1 + (1 << 31)], dtype=np.int64)
    flags = np.zeros(2, dtype=np.int32)
    lens = np.zeros(2, dtype=np.int32)
    lib = _lib.load()
    with DeviceNFA(nfa) as dev:
        rc = lib.gfsm_nfa_run_host(
            dev.handle, _lib.NFA_AUTO, data.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 2,
            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None)
    assert rc == _lib.EINVAL
    assert b"2^31" in lib.gfsm_last_error()


@pytest.mark.parametrize("n", [16, 256, 300, 4096, 40000])
@pytest.mark.parametrize("tech", DFA_TECHS)
def test_random_dfa_ragged_unaligned(n, tech):
    dfa = random_dfa(n, accept_prob=0.02, seed=n)
    if tech == "lds" and n > 320:
        with pytest.raises(ValueError):
            run_batch(dfa, [b"x"], backend="hip", technique=tech)
        return
    rng = random.Random(n)
    batch = [bytes(rng.randrange(256) for _ in range(rng.randint(0, 300))) for _ in range(257)]
    got = _verdicts(run_batch(dfa, batch, backend="hip", technique=tech))
    assert got == _oracle_dfa(dfa, batch)


@pytest.mark.parametrize("count", [131072, 262144])
def test_dfa_lds_many_inputs_per_cu(count):
    """Batches with 512 and 1024 inputs per CU (on 256 CUs) run the per-input walk in
    larger blocks; ragged short inputs, some empty, some latching."""
    dfa = random_dfa(256, accept_prob=0.02, seed=count)
    rng = np.random.default_rng(count)
    lens = rng.integers(0, 97, size=count)
    data = rng.integers(0, 256, size=int(lens.sum()), dtype=np.uint8)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    with DeviceDFA(dfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.DFA_TECHNIQUES["lds"])
    wf, wn, _ = native.dfa_batch(dfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    assert wf.any() and not wf.all()


def test_dfa_auto_resolution_by_batch_shape():
    """auto segments a batch of few long inputs, and walks one lane per input when the
    inputs alone fill every CU (far more than 256 per CU here) and the longest input is
    known to be short against the batch, or when inputs are short."""
    lds, split = _lib.DFA_TECHNIQUES["lds"], _lib.DFA_TECHNIQUES["split"]

    def longest(n):
        return make_tuning(max_input_bytes=n)

    with DeviceDFA(random_dfa(256, accept_prob=0.0, seed=5)) as dev:
        assert dev.resolve(4096, 4096 << 16) == split        # 4096 x 64 KiB
        # 1 M x 16 KiB: one lane per input once the longest input is known
        assert dev.resolve(1 << 20, (1 << 20) << 14, tuning=longest(1 << 14)) == lds
        assert dev.resolve(1 << 20, (1 << 20) << 14) == split  # longest not known
        assert dev.resolve(4096, 4096 << 12) == lds          # 4096 x 4 KiB
        # 65,535 inputs of 8 KiB plus one of 512 MiB: the long one would be walked by one
        # lane alone, so the batch is segmented even though the inputs fill the CUs
        total = 65535 * 8192 + (512 << 20)
        assert dev.resolve(65536, total, tuning=longest(512 << 20)) == split
        # 262,144 x 4 KiB, longest known: one lane per input
        assert dev.resolve(262144, 262144 << 12, tuning=longest(4096)) == lds
        # the host form computes the longest input itself (a skewed batch: split)
        lens = np.full(65536, 64, dtype=np.int64)
        lens[7] = 3 << 20
        off = np.zeros(lens.size + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        data = np.random.default_rng(1).integers(0, 256, size=int(off[-1]), dtype=np.uint8)
        f, n, _ = dev.run_host(data, off, _lib.DFA_AUTO)
        wf, wn, _ = native.dfa_batch(dev.dfa, data, off)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
    gather = _lib.DFA_TECHNIQUES["gather"]
    with DeviceDFA(random_dfa(4096, accept_prob=0.0, seed=5)) as dev:  # table in L2
        assert dev.resolve(1 << 20, (1 << 20) << 14, tuning=longest(1 << 14)) == gather
        assert dev.resolve(4096, 4096 << 16) == split        # one batch of 64 KiB inputs
        # 64 such batches in one launch: one lane per input once the longest is known
        assert dev.resolve(262144, 262144 << 16, tuning=longest(1 << 16)) == gather
        assert dev.resolve(262144, 262144 << 16) == split  # longest not known


@pytest.mark.parametrize("count", [131072, 262144])
def test_dfa_lds_many_lanes_long_inputs(count):
    """The 512- and 1024-lane per-input walks on inputs long enough to run whole 128-byte
    rounds: lengths up to 700 bytes, unaligned starts, and first accepts after byte 128 and
    inside the read-ahead round (one accepting state entered from 8 of the 65,536 (state,
    byte) pairs, so first matches land at every depth)."""
    dfa = random_dfa(256, accept_prob=0.0, seed=count)
    trans = np.asarray(dfa.trans).reshape(256, 256)
    trans[trans == 9] = 10
    trans[[1, 40, 80, 120, 160, 200, 240, 255], [5, 6, 7, 8, 9, 10, 11, 12]] = 9
    dfa.trans[:] = trans.reshape(-1)
    dfa.accept[9] = True
    rng = np.random.default_rng(count + 1)
    lens = rng.integers(0, 701, size=count)
    data = rng.integers(0, 256, size=int(lens.sum()), dtype=np.uint8)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    with DeviceDFA(dfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.DFA_TECHNIQUES["lds"])
    wf, wn, _ = native.dfa_batch(dfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    assert (wn[wf.astype(bool)] > 128).sum() > 1000   # accepts past the first round
    assert ((wn > 128) & (wn < 256)).any() and (wn > 512).any()


@pytest.mark.parametrize("tech", DFA_TECHS)
def test_dfa_non_latching_long_inputs(tech):
    """No accepting state: every byte is walked, across 16-byte vector boundaries."""
    dfa = random_dfa(256, accept_prob=0.0, seed=5)
    dfa.accept[17] = True  # a rare accepting state so some inputs latch late
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 5000, size=300)
    data = rng.integers(0, 256, size=int(lens.sum()), dtype=np.uint8)
    off = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    with DeviceDFA(dfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.DFA_TECHNIQUES[tech])
    wf, wn, _ = native.dfa_batch(dfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)


@pytest.mark.parametrize("tech", ["auto", "bitpacked", "worklist", "chain"])
def test_pattern_union_ragged(tech):
    """Ragged log-normal lengths (capped at 3000 to keep the oracle quick) over a union of
    short patterns, so that some inputs match and some do not."""
    pu = W.pattern_union(300, (3, 12), (1, 32), seed=3)
    lens = np.minimum(W.lognormal_lengths(300, 1024, 1.0, 0, 1 << 20, seed=4), 3000)
    data2, off2 = W.pattern_inputs(pu, lens, 0.7, seed=5)
    with DeviceNFA(pu.nfa) as dev:
        assert dev.props.chain_ok == 1
        f, n, _ = dev.run_host(data2, off2, _lib.NFA_TECHNIQUES[tech])
    wf, wn, _ = native.nfa_batch(pu.nfa, data2, off2)
    assert wf.any() and not wf.all()
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)


def test_pattern_union_many_tiles_non_latching_prefix():
    """65536-state union: 16 tiles; a few accepting states planted so some inputs latch."""
    pu, (data, off) = W.workload_divergent(count=16, length=1500, seed=4, accepting=False)
    last = np.cumsum(pu.lengths)  # state id of the last position of each pattern
    pu.nfa.accept[last[:400]] = True
    with DeviceNFA(pu.nfa) as dev:
        assert dev.props.chain_ok == 1 and dev.props.num_tiles >= 16
        f, n, _ = dev.run_host(data, off, _lib.NFA_BITPACKED)
    wf, wn, _ = native.nfa_batch(pu.nfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)


def test_pattern_union_latching_bitpacked_vs_worklist():
    pu, (data, off) = W.workload_divergent(count=64, length=4096, seed=6)
    with DeviceNFA(pu.nfa) as dev:
        fb, nb, _ = dev.run_host(data, off, _lib.NFA_BITPACKED)
        fw, nw, _ = dev.run_host(data, off, _lib.NFA_WORKLIST)
        fc, nc, _ = dev.run_host(data, off, _lib.NFA_CHAIN)
    wf, wn, _ = native.nfa_batch(pu.nfa, data, off)
    for f, n in ((fb, nb), (fw, nw), (fc, nc)):
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)


def _check_nfa(nfa, data, off, tech):
    with DeviceNFA(nfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.NFA_TECHNIQUES[tech])
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    return wf, wn


def _class_batch(rng, used, count, max_len, p_class=0.8, empty_frac=0.1):
    lens = rng.integers(0, max_len, size=count)
    lens[rng.random(count) < empty_frac] = 0
    total = int(lens.sum())
    pool = np.frombuffer(used, dtype=np.uint8)
    data = np.where(rng.random(total) < p_class, rng.choice(pool, size=total),
                    rng.integers(0, 256, size=total)).astype(np.uint8)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return data, off


def _chain_mode(monkeypatch, mode):
    """pairs: byte-pair lists, two items per wave in the register tier (the default);
    wave64: byte-pair lists, one item per wave; first: per-byte lists of first positions;
    triple: the lane tier over byte-triple lists (for automata the lists do not apply to,
    the pairs mode)."""
    monkeypatch.setenv("GFSM_CHAIN_FIRST", "1" if mode == "first" else "0")
    monkeypatch.setenv("GFSM_CHAIN_WAVE64", "1" if mode == "wave64" else "0")
    monkeypatch.setenv("GFSM_CHAIN_TRIPLE", "1" if mode == "triple" else "0")


@pytest.mark.parametrize("mode", ["pairs", "wave64", "first", "triple"])
@pytest.mark.parametrize("self_loop", ["any", "all", "partial", "none"])
@pytest.mark.parametrize("chunk", ["64", "333", "4096"])
def test_chain_random_automata(self_loop, chunk, mode, monkeypatch):
    """Shuffled ids, shared first positions, accepting inner states, every start self-loop
    kind; chunked (wildcard start) and unchunked inputs, chunk sizes not multiples of 4;
    second positions from byte-pair lists or first positions from per-byte lists."""
    monkeypatch.setenv("GFSM_CHAIN_CHUNK", chunk)
    _chain_mode(monkeypatch, mode)
    rng = random.Random(f"{self_loop}-{chunk}")
    nrng = np.random.default_rng(rng.randrange(1 << 30))
    for trial in range(4):
        nfa, used = chain_nfa(rng, rng.randint(1, 200), (3, 30), (1, 12), self_loop=self_loop,
                              p_accept=0.0005, p_last=0.03,
                              alphabet=bytes(rng.sample(range(256), 24)))
        data, off = _class_batch(nrng, used, 97, 3000)
        _check_nfa(nfa, data, off, "chain")


@pytest.mark.parametrize("mode", ["pairs", "wave64", "first", "triple"])
@pytest.mark.parametrize("chunk", ["64", "1000"])
def test_chain_long_inputs_match_positions(chunk, mode, monkeypatch):
    """Wildcard-start union with a few accepting patterns: first matches fall deep into
    64 KiB inputs, at every phase relative to the chunk boundaries."""
    monkeypatch.setenv("GFSM_CHAIN_CHUNK", chunk)
    _chain_mode(monkeypatch, mode)
    pu = W.pattern_union(400, (3, 8), (4, 32), seed=8, accepting=False)
    last = np.cumsum(pu.lengths)
    pu.nfa.accept[last[:3]] = True
    lens = np.full(48, 1 << 16, dtype=np.int64)
    data, off = W.pattern_inputs(pu, lens, 0.9, seed=9)
    wf, wn = _check_nfa(pu.nfa, data, off, "chain")
    assert wf.any() and not wf.all()
    assert len(set((wn[wf.astype(bool)] % int(chunk)).tolist())) > 5


@pytest.mark.parametrize("mode", ["pairs", "wave64", "first", "triple"])
def test_chain_overflow_falls_back_to_tiles(mode, monkeypatch):
    """Active sets of more than 1024 slots: 40 chains of wildcard-like classes, each 40 long,
    so every input overflows the LDS list and is rerun by the tiled pass."""
    _chain_mode(monkeypatch, mode)
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    b.add_transition(s0, 256, s0)
    for chain in range(40):
        prev = s0
        for pos in range(40):
            s = b.add_state(accept=(pos == 39 and chain < 4))
            for c in range(255 if pos == 39 else 256):
                b.add_transition(prev, c if pos < 39 else (c + chain) % 256, s)
            prev = s
    nfa = b.build()
    nrng = np.random.default_rng(12)
    data, off = _ragged(nrng, 40, 2000)
    wf, wn = _check_nfa(nfa, data, off, "chain")
    assert wf.any()


@pytest.mark.parametrize("mode", ["pairs", "wave64", "triple"])
def test_chain_medium_active_sets(mode, monkeypatch):
    """About 40 listed states at once: 3 chains of 15 positions whose classes hold every
    byte but the last position's. Two items per wave overflow their 32 lanes (the LDS tier
    reruns them), one item per wave holds them."""
    _chain_mode(monkeypatch, mode)
    monkeypatch.setenv("GFSM_CHAIN_CHUNK", "512")
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    b.add_transition(s0, 256, s0)
    for chain in range(3):
        prev = s0
        for pos in range(15):
            s = b.add_state(accept=(pos == 14))
            if pos < 14:
                b.add_transition(prev, 256, s)
            else:
                b.add_transition(prev, 7 + chain, s)
            prev = s
    nfa = b.build()
    nrng = np.random.default_rng(14)
    lens = nrng.integers(0, 3000, size=64)
    data = np.where(nrng.random(int(lens.sum())) < 0.998, 0x41, 7).astype(np.uint8)
    off = np.zeros(65, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    wf, wn = _check_nfa(nfa, data, off, "chain")
    assert wf.any() and not wf.all()


def _wildcard_chains(chains: int = 6, depth: int = 15):
    """Start state on every byte, then `chains` chains of `depth` positions; every position
    takes any byte except the last, which takes byte 7 + chain. After depth - 1 bytes about
    chains * (depth - 1) states are live on every byte."""
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    b.add_transition(s0, 256, s0)
    for chain in range(chains):
        prev = s0
        for pos in range(depth):
            s = b.add_state(accept=(pos == depth - 1))
            b.add_transition(prev, 256 if pos < depth - 1 else 7 + chain, s)
            prev = s
    return b.build()


def _wildcard_chains_verdicts(data, off, chains: int = 6, depth: int = 15):
    """Closed form for `_wildcard_chains`: the first accept ends on the first byte of
    7 .. 7 + chains - 1 at an input position p >= depth - 1, with match length p + 1."""
    idx = np.flatnonzero((data >= 7) & (data < 7 + chains))
    count = off.size - 1
    flags = np.zeros(count, np.int32)
    lens = np.zeros(count, np.int32)
    for i in range(count):
        k = np.searchsorted(idx, off[i] + depth - 1)
        if k < idx.size and idx[k] < off[i + 1]:
            flags[i], lens[i] = 1, idx[k] - off[i] + 1
    return flags, lens


def _wildcard_chain_batch(rng, count, length, late_hits):
    """`count` inputs of `length` bytes of 0x41; every other input gets one final-position
    byte in its last `late_hits` bytes, so its earlier chunks all have to run."""
    data = np.full(count * length, 0x41, dtype=np.uint8)
    off = np.arange(count + 1, dtype=np.int64) * length
    for i in range(0, count, 2):
        data[off[i] + length - 1 - rng.integers(0, late_hits)] = 7 + rng.integers(0, 6)
    return data, off


def test_wildcard_chain_closed_form_matches_the_oracle():
    """Pins the closed form used at full size below against the oracle, on inputs with the
    final-position bytes anywhere, including before position depth - 1."""
    nfa = _wildcard_chains()
    rng = np.random.default_rng(21)
    lens = rng.integers(0, 600, size=48)
    data = np.where(rng.random(int(lens.sum())) < 0.995, 0x41,
                    rng.integers(5, 15, size=int(lens.sum()))).astype(np.uint8)
    off = np.zeros(49, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    f, n = _wildcard_chains_verdicts(data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    assert wf.any() and not wf.all()


@pytest.mark.parametrize("mode", ["pairs", "wave64", "triple"])
def test_chain_many_deferred_items_per_wave(mode, monkeypatch):
    """About 84 live states on every byte overflow the register tier in both modes, so every
    later chunk of an input goes to the LDS tier unrun. 64 inputs of 1 MiB in 64-byte chunks
    are a million such items, over a hundred per wave of the persistent grid: the register
    tier's staged appends fill and flush their 64-entry buffers many times, and the LDS tier
    takes its items in blocks. Checked against the closed form pinned above."""
    _chain_mode(monkeypatch, mode)
    monkeypatch.setenv("GFSM_CHAIN_CHUNK", "64")
    nfa = _wildcard_chains()
    data, off = _wildcard_chain_batch(np.random.default_rng(22), 64, 1 << 20, 4096)
    with DeviceNFA(nfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.NFA_TECHNIQUES["chain"])
    wf, wn = _wildcard_chains_verdicts(data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    assert int(wf.sum()) == 32


@pytest.mark.parametrize("mode", ["pairs", "wave64", "triple"])
def test_chain_additions_into_lanes_freed_on_the_same_byte(mode, monkeypatch):
    """After A B...B, 32 states are listed (32 chains A B B ... B); on byte X all of them die
    while 20 second positions are added (chains B X Y). Two items per wave have no lane
    free before X, so every addition must go to a lane freed by the deaths on X, or the
    match on the following Y is lost."""
    _chain_mode(monkeypatch, mode)
    A, B, X, Y = 0x41, 0x42, 0x58, 0x59
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    b.add_transition(s0, 256, s0)
    for _ in range(32):
        prev = s0
        for pos in range(30):
            s = b.add_state()
            b.add_transition(prev, A if pos == 0 else B, s)
            prev = s
    for k in range(20):
        s1, s2, s3 = b.add_state(), b.add_state(), b.add_state(accept=k in (0, 19))
        b.add_transition(s0, B, s1)
        b.add_transition(s1, X, s2)
        b.add_transition(s2, Y, s3)
    nfa = b.build()
    rng = np.random.default_rng(15)
    inputs = []
    for _ in range(96):
        parts = []
        for _ in range(rng.integers(1, 6)):
            # noise from bytes no class holds: every segment starts from an empty list
            parts.append(bytes(rng.integers(0, 64, size=rng.integers(1, 40)).astype(np.uint8)))
            parts.append(bytes([A]) + bytes([B]) * int(rng.integers(1, 26)) +
                         (bytes([X, Y]) if rng.random() < 0.3 else bytes([X])))
        inputs.append(b"".join(parts))
    data, off = pack_bytes(inputs)
    wf, wn = _check_nfa(nfa, data, off, "chain")
    assert wf.any() and not wf.all()


@pytest.mark.parametrize("mode", ["pairs", "wave64", "first", "triple"])
def test_chain_many_tiles(mode, monkeypatch):
    """90,206 states in 23 tiles: above 65,536 slots the lists hold 32-bit slots and the
    accept words are read from global memory instead of LDS. Accepting third positions of
    50 patterns make about half the inputs match."""
    _chain_mode(monkeypatch, mode)
    pu = W.pattern_union(4000, (5, 40), (1, 8), seed=17)
    first = 1 + np.concatenate(([0], np.cumsum(pu.lengths)[:-1]))
    pu.nfa.accept[:] = False
    pu.nfa.accept[first[:50] + 2] = True
    lens = np.random.default_rng(17).integers(0, 6000, size=96)
    data, off = W.pattern_inputs(pu, lens, 0.8, seed=18)
    with DeviceNFA(pu.nfa) as dev:
        assert dev.props.num_tiles > 16
    wf, wn = _check_nfa(pu.nfa, data, off, "chain")
    assert wf.any() and not wf.all()


def test_chain_run_batch_api():
    pu = W.pattern_union(120, seed=13)
    lens = np.random.default_rng(13).integers(0, 9000, size=60)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=14)
    batch = [data[off[i]:off[i + 1]].tobytes() for i in range(60)]
    got = _verdicts(run_batch(pu.nfa, batch, backend="hip", technique="chain"))
    assert got == _oracle_nfa(pu.nfa, batch)


def test_chain_rejects_non_chain_nfa():
    nfa, _ = EXAMPLES["ab_star_c_plus_d"]()
    with pytest.raises(Exception, match="chain-structured"):
        run_batch(nfa, [b"abcd"], backend="hip", technique="chain")


def test_plain_nfa_config():
    nfa, (data, off) = W.workload_plain()
    with DeviceNFA(nfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.NFA_AUTO)
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)


def _ragged(rng, count, max_len, empty_frac=0.1):
    lens = rng.integers(0, max_len, size=count)
    lens[rng.random(count) < empty_frac] = 0
    data = rng.integers(0, 256, size=int(lens.sum()), dtype=np.uint8)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return data, off


def _check_dfa(dfa, data, off, tech):
    with DeviceDFA(dfa) as dev:
        f, n, _ = dev.run_host(data, off, _lib.DFA_TECHNIQUES[tech])
    wf, wn, _ = native.dfa_batch(dfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    return wf


@pytest.mark.parametrize("n", [16, 256, 4096, 40000])
def test_split_rare_accepts_in_continuation_segments(n):
    """Accepting states entered rarely, so matches land deep inside long inputs."""
    dfa = random_dfa(n, accept_prob=0.0, seed=n + 1)
    dfa.accept[1] = True  # one accepting state: first visit after about n bytes
    rng = np.random.default_rng(n)
    data, off = _ragged(rng, 96, 12 * max(n, 1024))
    wf = _check_dfa(dfa, data, off, "split")
    assert wf.any()


def test_split_permutation_dfa_never_merges():
    """Every byte permutes a 16-state component, so speculative paths never merge and pass 3
    re-walks every continuation segment. A second component counts consecutive 0xFF bytes
    and accepts at two in a row: matches land deep inside the inputs."""
    from gpufsm.core.dfa import DFA

    m = 16
    rng = np.random.default_rng(9)
    perms = np.stack([rng.permutation(m) for _ in range(256)], axis=1)  # [p, byte]
    n = 3 * m  # state = 3 * p + k, k = 0xFF run length capped at 2
    trans = np.zeros((n, 256), dtype=np.int32)
    for p in range(m):
        for k in range(3):
            for b in range(256):
                k2 = min(k + 1, 2) if b == 0xFF else 0
                trans[3 * p + k, b] = 3 * perms[p, b] + k2
    for accepting in (False, True):
        accept = np.zeros(n, dtype=bool)
        if accepting:
            accept[2::3] = True
        dfa = DFA(num_states=n, start_state=0, accept=accept, trans=trans.reshape(-1))
        data, off = _ragged(rng, 48, 60000)
        wf = _check_dfa(dfa, data, off, "split")
        if accepting:
            assert wf.any() and not wf.all()


def test_split_latching_and_many_short_inputs():
    dfa = random_dfa(256, accept_prob=0.10, seed=3)
    rng = np.random.default_rng(3)
    data, off = _ragged(rng, 5000, 300)
    _check_dfa(dfa, data, off, "split")
    data, off = W.uniform_batch(64, 65536, seed=1)
    _check_dfa(dfa, data, off, "split")


def _union_as_anml(pu) -> str:
    """A pattern union written as ANML: one all-input STE chain per pattern."""
    parts = ["<anml><automata-network>"]
    pos = 0
    for p, length in enumerate(pu.lengths):
        for i in range(int(length)):
            cls = pu.class_bytes[pu.class_start[pos]:pu.class_start[pos + 1]]
            sym = "[" + "".join(f"0x{int(c):02x}" for c in cls) + "]"
            start = ' start="all-input"' if i == 0 else ""
            inner = (f'<activate-on-match element="p{p}_{i + 1}"/>' if i + 1 < length
                     else "<report-on-match/>")
            parts.append(f'<state-transition-element id="p{p}_{i}" symbol-set="{sym}"{start}>'
                         f"{inner}</state-transition-element>")
            pos += 1
    parts.append("</automata-network></anml>")
    return "".join(parts)


@pytest.mark.parametrize("fold", [False, True])
def test_anml_pattern_set(tmp_path, fold):
    """ANML pattern sets: with the synthetic start states (epsilon edges, general list walk)
    and folded (chain-structured, chain technique), both equal to the oracle."""
    from gpufsm.io import load_anml

    pu = W.pattern_union(150, (3, 10), (1, 24), seed=17)
    path = tmp_path / "set.anml"
    path.write_text(_union_as_anml(pu))
    nfa = load_anml(path, fold_starts=fold)
    lens = np.random.default_rng(17).integers(0, 5000, size=80)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=18)
    with DeviceNFA(nfa) as dev:
        assert dev.props.chain_ok == (1 if fold else 0)
        f, n, _ = dev.run_host(data, off, _lib.NFA_AUTO)
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    assert wf.any() and not wf.all()
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
