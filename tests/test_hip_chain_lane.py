"""The chain technique's lane tier on the GPU: one (input, chunk) item per lane over the
byte-triple lists of states entered at depth 3 (tuning flag GFSM_TUNE_CHAIN_TRIPLE).

Every verdict is compared with the C oracle. The cases cover accepts at every depth the tier
treats apart (1, 2 and 3 from the flags, the pair accept bits and the list header; deeper
ones from held states), lists of one entry (inline in the header), of two to four and of
more than four entries, held sets that overflow the lane's eight slots (the LDS tier reruns
those items and the later items of their inputs), chunk sizes that are not multiples of 16,
batches whose device pointer is not 16-byte aligned (the tier's 16-byte group loads fall
back to byte loads at the batch's ends), empty and one- or two-byte inputs, the ragged
pattern union of the benchmark, and automata the lists do not apply to (the flag is then
ignored).
"""

from __future__ import annotations

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.backends.hip.handles import DeviceNFA, make_tuning
from gpufsm.core.nfa import NFABuilder
from oracle import native

pytestmark = pytest.mark.gpu

CHAIN = _lib.NFA_TECHNIQUES["chain"]


def _run_device(nfa, data, off, chunk, shift=0, triple=True):
    """The batch through gfsm_nfa_run_device_ws on device buffers, the data placed `shift`
    bytes into its allocation."""
    import torch

    count = off.size - 1
    total = int(off[-1])
    buf = torch.zeros(total + shift + 64, dtype=torch.uint8, device="cuda")
    if total:
        buf[shift:shift + total] = torch.from_numpy(data).cuda()
    d_off = torch.from_numpy(off).cuda()
    flags = torch.zeros(max(count, 1), dtype=torch.int32, device="cuda")
    lens = torch.zeros_like(flags)
    tuning = make_tuning(chunk_bytes=chunk, flags=_lib.TUNE_CHAIN_TRIPLE if triple else 0)
    with DeviceNFA(nfa) as dev:
        ws = dev.workspace_size(count, total, CHAIN, tuning)
        w = torch.empty(max(ws, 1), dtype=torch.uint8, device="cuda")
        dev.run_device_ws(buf.data_ptr() + shift, d_off.data_ptr(), count, total,
                          flags.data_ptr(), lens.data_ptr(), w.data_ptr(), ws,
                          torch.cuda.current_stream().cuda_stream, CHAIN, tuning)
        torch.cuda.synchronize()
    return flags.cpu().numpy()[:count], lens.cpu().numpy()[:count]


def _check(nfa, data, off, chunk=1024, shift=0):
    f, n = _run_device(nfa, data, off, chunk, shift)
    wf, wn, _ = native.nfa_batch(nfa, data, off)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    return wf, wn


def _union(patterns, accepting):
    """Start state with a wildcard self-loop; pattern k is a chain over classes
    patterns[k] (a list of byte sets); its last position accepts when accepting[k]."""
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    b.add_transition(s0, 256, s0)
    for cls, acc in zip(patterns, accepting):
        prev = s0
        for pos, c in enumerate(cls):
            s = b.add_state(accept=acc and pos == len(cls) - 1)
            for byte in c:
                b.add_transition(prev, byte, s)
            prev = s
    return b.build()


def _batch(rng, count, max_len, alphabet, empty_frac=0.1):
    lens = rng.integers(0, max_len + 1, size=count)
    lens[rng.random(count) < empty_frac] = 0
    lens[:3] = [1, 2, 3]
    data = rng.choice(np.frombuffer(alphabet, dtype=np.uint8), size=int(lens.sum()))
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return data.astype(np.uint8), off


@pytest.mark.parametrize("chunk", [64, 333, 4096])
@pytest.mark.parametrize("depth", [1, 2, 3, 4, 6])
def test_accepts_at_every_depth(depth, chunk):
    """Patterns of lengths 1 to 6 over a 12-byte alphabet, only those of length `depth`
    accepting, so that every match is found by one mechanism: the flags (depth 1), the pair
    bits (2), the list header (3) or a held state (4 and more), in every chunk phase."""
    rng = np.random.default_rng(10 * depth + chunk)
    pats, acc = [], []
    for k in range(150):
        length = 1 + k % 6
        pats.append([set(rng.choice(12, size=rng.integers(1, 4), replace=False).tolist())
                     for _ in range(length)])
        acc.append(length == depth and k < 60)
    nfa = _union(pats, acc)
    data, off = _batch(rng, 700, 900, bytes(range(14)))
    wf, wn = _check(nfa, data, off, chunk)
    assert wf.any() and not wf.all()
    assert int(wn[wf.astype(bool)].min()) >= depth


def test_long_lists_and_held_overflow():
    """Patterns sharing their first three classes: the list of (A, B, C) holds 6 entries
    (more than four: the slow path of the entry loads); (D, E, F) holds 12, more than a lane
    holds, so those items overflow to the LDS tier, with the later items of their inputs."""
    A, B, C, D, E, F = 65, 66, 67, 68, 69, 70
    pats = [[{A}, {B}, {C}, {48 + k}, {90}] for k in range(6)]
    pats += [[{D}, {E}, {F}, {48 + k}, {91}] for k in range(12)]
    pats += [[{A, D}, {48}, {49, 50}]]
    acc = [k % 2 == 0 for k in range(len(pats))]
    nfa = _union(pats, acc)
    rng = np.random.default_rng(5)
    pieces = [b"ABC0Z", b"ABC2Z", b"DEF3[", b"DEF4[", b"A01", b"D02", b"ABC", b"DEF"]
    inputs = []
    for _ in range(400):
        parts = [bytes(rng.integers(97, 123, size=rng.integers(0, 300)).astype(np.uint8))]
        for _ in range(rng.integers(0, 6)):
            parts.append(pieces[rng.integers(0, len(pieces))])
            parts.append(bytes(rng.integers(97, 123, size=rng.integers(0, 200)).astype(np.uint8)))
        inputs.append(b"".join(parts))
    lens = np.array([len(x) for x in inputs], dtype=np.int64)
    off = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    data = np.frombuffer(b"".join(inputs), dtype=np.uint8).copy()
    for chunk in (64, 512):
        wf, wn = _check(nfa, data, off, chunk)
    assert wf.any() and not wf.all()


@pytest.mark.parametrize("shift", [1, 7, 15])
def test_unaligned_batch_pointer(shift):
    """The batch starts `shift` bytes past a 16-byte boundary: every item's groups are
    aligned in absolute address terms, and the groups that straddle the batch's first and
    last bytes are read byte by byte."""
    pu = W.pattern_union(300, (2, 9), (1, 24), seed=shift)
    lens = np.random.default_rng(shift).integers(0, 700, size=257)
    data, off = W.pattern_inputs(pu, lens, 0.8, seed=shift + 1)
    wf, wn = _check(pu.nfa, data, off, 256, shift)
    assert wf.any() and not wf.all()


def test_ragged_union_against_the_oracle():
    """The benchmark's 2000-pattern union (45,912 states, 12 tiles) on 2048 ragged inputs."""
    pu = W.pattern_union(2000, (5, 40), (1, 32), seed=0)
    lens = W.lognormal_lengths(2048, 4096, 1.0, 64, 1 << 18, seed=1)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=2)
    with DeviceNFA(pu.nfa) as dev:
        assert dev.props.start_always and dev.props.num_tiles <= 16
    for chunk in (512, 2048):
        wf, wn = _check(pu.nfa, data, off, chunk)
    assert wf.any() and not wf.all()


def test_flag_ignored_where_the_lists_do_not_apply():
    """No wildcard self-loop on the start state (no chunks, no lists), and more than 16
    tiles (slots past 16 bits): the flag changes nothing, the verdicts stay exact."""
    b = NFABuilder()
    s0 = b.add_state()
    b.set_start(s0)
    for c in range(0, 256, 2):
        b.add_transition(s0, c, s0)
    prev = s0
    for c in (10, 12, 14):
        s = b.add_state(accept=c == 14)
        b.add_transition(prev, c, s)
        prev = s
    nfa = b.build()
    rng = np.random.default_rng(9)
    data, off = _batch(rng, 200, 400, bytes([10, 12, 14, 3]))
    _check(nfa, data, off, 256)
    big = W.pattern_union(4000, (5, 40), (1, 8), seed=17)
    first = 1 + np.concatenate(([0], np.cumsum(big.lengths)[:-1]))
    big.nfa.accept[:] = False
    big.nfa.accept[first[:50] + 2] = True
    lens = np.random.default_rng(17).integers(0, 3000, size=64)
    data, off = W.pattern_inputs(big, lens, 0.8, seed=18)
    wf, _ = _check(big.nfa, data, off, 512)
    assert wf.any()
