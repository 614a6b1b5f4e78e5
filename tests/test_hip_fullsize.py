"""Parity at the configured full sizes, through properties that need no oracle run.

The oracle gate in ``bench.py`` compares full batches bit for bit; these tests check, at the
same sizes, what must hold whatever the verdicts are:

* techniques agree with each other (split, gather and lds on 256 MiB; chain and bitpacked on
  the 109 MB ragged mix and on the 1 GiB high-divergence batch; the masked bit-parallel step
  and the same step over gather lists only on the general 65,536-state control);
* verdicts do not depend on batch order (split segments and chain chunks then fall on other
  bytes of every input), nor on the chain chunk length;
* latch-first-match: ``match_len <= len``, a rejected input reports 0, an accepted input's
  prefix of ``match_len`` bytes accepts with the same length, and its prefix one byte
  shorter does not accept.
"""

from __future__ import annotations

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.backends.hip.handles import DeviceDFA, DeviceNFA, make_tuning

pytestmark = pytest.mark.gpu


def _repack(data, off, order):
    """The batch with its inputs in ``order``."""
    lens = np.diff(off)[order]
    out = np.zeros(order.size + 1, dtype=np.int64)
    np.cumsum(lens, out=out[1:])
    parts = [data[off[i]:off[i + 1]] for i in order]
    return (np.concatenate(parts) if parts else np.zeros(0, np.uint8)), out


def _prefixes(data, off, ends):
    """Inputs i cut to ``ends[i]`` bytes."""
    parts = [data[off[i]:off[i] + int(e)] for i, e in enumerate(ends)]
    out = np.zeros(len(parts) + 1, dtype=np.int64)
    np.cumsum([p.size for p in parts], out=out[1:])
    return (np.concatenate(parts) if parts else np.zeros(0, np.uint8)), out


def _check_latch_invariants(run, data, off, f, n):
    """``run(data, off) -> (flags, lens)`` on prefixes of the batch's inputs."""
    lens = np.diff(off)
    assert (n <= lens).all() and (n >= 0).all()
    assert (n[f == 0] == 0).all()
    acc = np.flatnonzero(f)
    if not acc.size:
        return
    # every input cut to its match length (rejected ones to 0 bytes) accepts at that length
    pf, pn = run(*_prefixes(data, off, np.where(f != 0, n, 0)))
    np.testing.assert_array_equal(pf[acc], 1)
    np.testing.assert_array_equal(pn[acc], n[acc])
    # one byte shorter, nothing accepts (a match of length 0 is the start state accepting)
    pos = acc[n[acc] > 0]
    sf, _ = run(*_prefixes(data, off, np.where(f != 0, np.maximum(n - 1, 0), 0)))
    assert not sf[pos].any()


@pytest.fixture(scope="module")
def dfa_batch():
    return W.uniform_batch(4096, 65536, seed=0)  # 256 MiB, the configured DFA batch


@pytest.mark.parametrize("n,accept", [(256, 0.0), (256, 0.10), (4096, 0.0)])
def test_dfa_full_batch_techniques_agree(dfa_batch, n, accept):
    data, off = dfa_batch
    dfa = W.dense_dfa(n, accept, seed=0)
    if accept == 0.0:
        dfa.accept[7] = True  # one accepting state: matches deep inside most inputs
    with DeviceDFA(dfa) as dev:
        techs = ["split", "gather"] + (["lds"] if dev.props.lds_ok else [])
        res = {t: dev.run_host(data, off, _lib.DFA_TECHNIQUES[t])[:2] for t in techs}
        f, m = res["split"]
        for t in techs[1:]:
            np.testing.assert_array_equal(res[t][0], f, err_msg=t)
            np.testing.assert_array_equal(res[t][1], m, err_msg=t)
        # reversed batch order: split segments now start at other bytes of every input
        order = np.arange(off.size - 2, -1, -1)
        rd, ro = _repack(data, off, order)
        rf, rm = dev.run_host(rd, ro, _lib.DFA_SPLIT)[:2]
        np.testing.assert_array_equal(rf, f[order])
        np.testing.assert_array_equal(rm, m[order])
        _check_latch_invariants(lambda d, o: dev.run_host(d, o, _lib.DFA_SPLIT)[:2],
                                data, off, f, m)


def _run_chain(dev, data, off, chunk):
    """Chain technique through the workspace form with a fixed chunk length."""
    import torch

    count, total = off.size - 1, int(off[-1])
    tu = make_tuning(chunk, 0, 0)
    need = dev.workspace_size(count, total, _lib.NFA_CHAIN, tu)
    ws = torch.empty(max(need, 256), dtype=torch.uint8, device="cuda")
    d = torch.from_numpy(data).cuda()
    o = torch.from_numpy(off).cuda()
    f = torch.empty(count, dtype=torch.int32, device="cuda")
    n = torch.empty(count, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    dev.run_device_ws(d.data_ptr(), o.data_ptr(), count, total, f.data_ptr(), n.data_ptr(),
                      ws.data_ptr(), ws.numel(), s.cuda_stream, _lib.NFA_CHAIN, tu)
    s.synchronize()
    return f.cpu().numpy(), n.cpu().numpy()


def test_ragged_mix_full_batch_chain_agrees_with_tiles():
    """The ragged NFA mix at full size: 16,384 inputs, 109 MB."""
    pu, (data, off) = W.workload_ragged(count=16384)
    with DeviceNFA(pu.nfa) as dev:
        f, n = _run_chain(dev, data, off, 2048)
        bf, bn = dev.run_host(data, off, _lib.NFA_BITPACKED)[:2]
        np.testing.assert_array_equal(bf, f)
        np.testing.assert_array_equal(bn, n)
        f2, n2 = _run_chain(dev, data, off, 8192)  # other chunk boundaries
        np.testing.assert_array_equal(f2, f)
        np.testing.assert_array_equal(n2, n)
        order = np.random.default_rng(5).permutation(off.size - 1)
        rd, ro = _repack(data, off, order)
        rf, rn = _run_chain(dev, rd, ro, 2048)
        np.testing.assert_array_equal(rf, f[order])
        np.testing.assert_array_equal(rn, n[order])
        _check_latch_invariants(lambda d, o: _run_chain(dev, d, o, 2048), data, off, f, n)
    assert f.any() and not f.all()


def test_divergent_full_batch_chain_agrees_with_tiles():
    """The high-divergence NFA at full size: 65,536 states, 1024 inputs of 1 MiB."""
    pu, (data, off) = W.workload_divergent(count=1024, length=1 << 20, seed=0)
    with DeviceNFA(pu.nfa) as dev:
        f, n = _run_chain(dev, data, off, 4096)
        bf, bn = dev.run_host(data, off, _lib.NFA_BITPACKED)[:2]
        np.testing.assert_array_equal(bf, f)
        np.testing.assert_array_equal(bn, n)
        _check_latch_invariants(lambda d, o: _run_chain(dev, d, o, 4096), data, off, f, n)


def test_general_control_full_batch_masks_agree_with_lists():
    """The general (non-chain) 65,536-state control at its bench size, 1024 inputs of
    16 KiB, with 7 accepting pattern ends so that half the inputs match thousands of bytes
    in: the masked bit-parallel step (first and diagonal masks) equals the same step with
    every edge in the gather lists, on the whole batch; latch-first-match holds; and 8
    inputs equal the oracle."""
    import dataclasses
    import os

    from oracle import native

    pu = W.pattern_union(8000, (5, 12), (16, 64), seed=0, total_states=65536)
    nfa = W.optional_positions(pu, 0.25, seed=1)
    acc = np.asarray(nfa.accept).copy()
    ends = np.flatnonzero(acc)
    acc[:] = False
    acc[ends[np.random.default_rng(7).random(ends.size) < 0.002]] = True
    nfa = dataclasses.replace(nfa, accept=acc)
    data, off = W.pattern_inputs(pu, np.full(1024, 16384, dtype=np.int64), 0.95, seed=2)
    old = os.environ.get("GFSM_GATHER_DIAG")
    try:
        os.environ["GFSM_GATHER_DIAG"] = "0"
        with DeviceNFA(nfa) as dev:
            lf, ln = dev.run_host(data, off, _lib.NFA_BITPACKED)[:2]
        os.environ["GFSM_GATHER_DIAG"] = "1"
        with DeviceNFA(nfa) as dev:
            assert dev.resolve(1024, int(off[-1])) == _lib.NFA_BITPACKED  # first masks
            f, n = dev.run_host(data, off, _lib.NFA_BITPACKED)[:2]
            np.testing.assert_array_equal(f, lf)
            np.testing.assert_array_equal(n, ln)
            _check_latch_invariants(lambda d, o: dev.run_host(d, o, _lib.NFA_BITPACKED)[:2],
                                    data, off, f, n)
    finally:
        if old is None:
            os.environ.pop("GFSM_GATHER_DIAG", None)
        else:
            os.environ["GFSM_GATHER_DIAG"] = old
    assert 0.2 < f.mean() < 0.8 and np.median(n[f > 0]) > 1000
    wf, wn, _ = native.nfa_batch(nfa, data, off[:9])
    np.testing.assert_array_equal(f[:8], wf)
    np.testing.assert_array_equal(n[:8], wn)
