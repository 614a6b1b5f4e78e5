"""The native serving loop (``gfsm_server_*``, :class:`gpufsm.serve.NativeServingLoop`).

Every served verdict is checked against the oracle and against the request's lone run;
the admission, rejection, timeout and stop paths end every request (none is left waiting).
"""

from __future__ import annotations

import ctypes
import threading

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.batch import BatchEngine
from gpufsm.core.packing import pack_bytes
from gpufsm.serve import NativeServingLoop, poisson_drive
from oracle import native

pytestmark = pytest.mark.gpu


def _oracle(a, data, off):
    fn = native.dfa_batch if hasattr(a, "trans") else native.nfa_batch
    f, n, _ = fn(a, data, off)
    return f, n


def _check_against_oracle(automata, reqs, done):
    by_name: dict = {}
    for k, (name, _) in enumerate(reqs):
        by_name.setdefault(name, []).append(k)
    for name, ks in by_name.items():
        data, off = pack_bytes([reqs[k][1] for k in ks])
        wf, wn = _oracle(automata[name], data, off)
        for j, k in enumerate(ks):
            r = done[k]
            assert r.status == "ok", (name, k, r.status)
            assert (r.accepted, r.match_len) == (bool(wf[j]), int(wn[j])), (name, k)


def test_native_configured_mix_equals_alone_and_oracle():
    """The configured serving mix (4 random DFAs of 256 to 16,384 states, 4 pattern unions
    of 1,024 to 65,536 states), Poisson arrivals, log-normal lengths plus requests of the
    4 MiB cap: every served verdict equals the oracle and the request's lone run."""
    automata, unions = W.serving_automata(seed=0)
    reqs = W.serving_requests(automata, unions, 400, seed=13)
    rng = np.random.default_rng(14)
    for name in ("dfa16384", "nfa65536", "dfa4096", "nfa4096"):
        n = 4 << 20
        if name in unions:
            d, _ = W.pattern_inputs(unions[name], np.array([n], dtype=np.int64), 0.70, seed=15)
        else:
            d = rng.integers(0, 256, size=n, dtype=np.uint8)
        reqs.insert(int(rng.integers(0, len(reqs))), (name, d.tobytes()))
    reqs.insert(7, ("dfa256", b""))  # an empty request
    loop = NativeServingLoop(automata, budget_bytes=32 << 30, slots=8)
    with loop:
        done = poisson_drive(loop, reqs, qps=8000, seed=16)
        st = loop.stats()
    assert st["ok"] == len(reqs) and st["rejected"] == 0
    assert st["launches"] < len(reqs)  # requests shared launches
    _check_against_oracle(automata, reqs, done)
    # lone runs of a sample, one request per launch
    alone = BatchEngine(tuner=None)
    try:
        for k in range(0, len(reqs), 37):
            name, data = reqs[k]
            d = np.frombuffer(data, dtype=np.uint8)
            f, n, _ = alone.run_packed(automata[name], d, np.array([0, d.size], np.int64))
            assert (done[k].accepted, done[k].match_len) == (bool(f[0]), int(n[0])), (name, k)
    finally:
        alone.close()
    assert all(r.latency >= 0 for r in done)


def test_native_budget_serialises_and_rejects_oversized():
    """A budget just above the resident automata and one request: requests are admitted
    one at a time and all complete; a request longer than a slot is rejected at once."""
    automata, unions = W.serving_automata(seed=1)
    small = {k: automata[k] for k in ("dfa256", "nfa1024")}
    probe = NativeServingLoop(small, slots=2)
    resident = probe.query()["resident_bytes"]
    probe.close()
    reqs = W.serving_requests(small, {"nfa1024": unions["nfa1024"]}, 60, seed=21,
                              max_len=64 << 10)
    loop = NativeServingLoop(small, budget_bytes=resident + (64 << 10), slots=4,
                             max_batch_bytes=1 << 20)
    with loop:
        big = loop.submit("dfa256", bytes(2 << 20))  # longer than a slot
        done = [loop.submit(name, data) for name, data in reqs]
        loop.collect(done)
        assert big.wait(30.0)
        assert big.status == "rejected"
        st = loop.query()
        assert st["rejected"] == 1 and st["completed"] == len(reqs)
        assert st["in_flight_bytes"] == 0
    _check_against_oracle(small, reqs, done)


def test_native_queue_timeout_and_stop_without_drain():
    """Requests waiting past the queue timeout end as timed out; a stop without drain ends
    every request (ok or error), none is left pending."""
    automata, unions = W.serving_automata(seed=2)
    small = {"dfa1024": automata["dfa1024"]}
    probe = NativeServingLoop(small, slots=1)
    resident = probe.query()["resident_bytes"]
    probe.close()
    # budget for one 1 MiB request: the rest wait, and time out after 1 ms unless the
    # first finishes sooner
    loop = NativeServingLoop(small, budget_bytes=resident + (1 << 20), slots=1,
                             queue_timeout_s=0.001)
    loop.start()
    reqs = [loop.submit("dfa1024", bytes(1 << 20)) for _ in range(200)]
    loop.collect(reqs, 60.0)
    statuses = {r.status for r in reqs}
    assert statuses <= {"ok", "timeout"} and "timeout" in statuses
    loop.stop()
    loop.close()

    loop = NativeServingLoop(small, slots=2)
    loop.start()
    reqs = [loop.submit("dfa1024", bytes(256 << 10)) for _ in range(400)]
    loop.stop(drain=False)
    assert all(r.status in ("ok", "error") for r in reqs)
    assert any(r.status == "ok" for r in reqs)
    with pytest.raises(_lib.HIPError):
        loop.submit("dfa1024", b"abc")  # stopped
    loop.close()


def test_native_concurrent_submitters_and_bad_arguments():
    """Several host threads submit at once; unknown names and indexes are errors."""
    automata, unions = W.serving_automata(seed=3)
    small = {k: automata[k] for k in ("dfa256", "nfa4096")}
    reqs = W.serving_requests(small, {"nfa4096": unions["nfa4096"]}, 240, seed=31,
                              max_len=256 << 10)
    loop = NativeServingLoop(small, slots=4)
    with loop:
        with pytest.raises(KeyError):
            loop.submit("nope", b"x")
        rid = ctypes.c_int64(0)
        rc = loop._so.gfsm_server_submit(loop._srv, 99, b"x", 1, ctypes.byref(rid))
        assert rc == _lib.EINVAL and b"unknown automaton" in _lib.load().gfsm_last_error()
        done: list = [None] * len(reqs)

        def worker(lo):
            for k in range(lo, len(reqs), 4):
                done[k] = loop.submit(*reqs[k])

        threads = [threading.Thread(target=worker, args=(lo,)) for lo in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        loop.collect(done)
    _check_against_oracle(small, reqs, done)


def test_native_bulk_submit_matches_oracle_and_checks_offsets():
    """submit_many: one call for a packed batch (a slice of a larger buffer, empty requests
    included) next to single submits of another automaton; every verdict equals the
    oracle. Offsets that decrease or leave the data are refused before anything is queued."""
    automata, unions = W.serving_automata(seed=4)
    small = {k: automata[k] for k in ("dfa4096", "nfa16384")}
    pu = unions["nfa16384"]
    lens = W.lognormal_lengths(700, 8192, 1.0, 0, 1 << 20, seed=41)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=42)
    loop = NativeServingLoop(small, slots=4, max_batch_bytes=32 << 20)
    with loop:
        singles = [("dfa4096", bytes(np.random.default_rng(k).integers(0, 256, 3000,
                                                                       dtype=np.uint8)))
                   for k in range(20)]
        s_reqs = [loop.submit(name, x) for name, x in singles]
        bulk = loop.submit_many("nfa16384", data, off[100:])  # offsets need not start at 0
        in_place = loop.submit_many("nfa16384", data, off[:101], copy=False)
        loop.collect(s_reqs)
        assert bulk.wait(60.0) and in_place.wait(60.0)
        with pytest.raises(ValueError):  # past the end of the data
            loop.submit_many("nfa16384", data, np.array([0, 5, data.size + 1], np.int64))
        with pytest.raises(_lib.HIPError):
            ids = np.zeros(2, np.int64)
            bad = np.array([0, 10, 4], np.int64)  # decreasing: refused by the library
            rc = loop._so.gfsm_server_submit_many(
                loop._srv, 0, ctypes.c_void_p(data.ctypes.data),
                bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 2, 0,
                ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
            _lib.check(rc)
        assert loop.query()["submitted"] == len(bulk) + len(in_place) + len(s_reqs)
    sub_off = off[100:] - off[100]
    wf, wn = _oracle(pu.nfa, data[off[100]:], sub_off)
    assert len(bulk) == sub_off.size - 1 and (bulk.status == _lib.SERVE_OK).all()
    np.testing.assert_array_equal(bulk.accepted, wf.astype(bool))
    np.testing.assert_array_equal(bulk.match_len, wn)
    assert (bulk.latency >= 0).all()
    wf, wn = _oracle(pu.nfa, data, off[:101])
    np.testing.assert_array_equal(in_place.accepted, wf.astype(bool))
    np.testing.assert_array_equal(in_place.match_len, wn)
    _check_against_oracle(small, singles, s_reqs)


def test_native_bulk_submit_from_pinned_memory():
    """submit_many in place from a caller buffer in pinned host memory: the scheduler copies
    those runs to the device directly. Launches that mix them with staged requests of the
    same automaton (single submits and a pageable bulk submit between them) keep every
    verdict equal to the oracle."""
    import torch

    automata, unions = W.serving_automata(seed=5)
    small = {"nfa16384": automata["nfa16384"]}
    pu = unions["nfa16384"]
    lens = W.lognormal_lengths(900, 6000, 1.0, 0, 1 << 20, seed=51)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=52)
    pinned_t = torch.empty(int(off[-1]), dtype=torch.uint8, pin_memory=True)
    pinned = pinned_t.numpy()
    pinned[:] = data[: int(off[-1])]
    wf, wn = _oracle(pu.nfa, data, off)
    loop = NativeServingLoop(small, slots=3, max_batch_bytes=8 << 20, max_batch_count=256)
    with loop:
        a = loop.submit_many("nfa16384", pinned, off[:401], copy=False)  # inputs 0..399
        singles = [loop.submit("nfa16384", data[off[k]:off[k + 1]].tobytes())
                   for k in range(400, 450)]
        b = loop.submit_many("nfa16384", data, off[450:701], copy=False)  # pageable
        c = loop.submit_many("nfa16384", pinned, off[700:], copy=False)  # not from 0
        assert a.wait(60.0) and b.wait(60.0) and c.wait(60.0)
        loop.collect(singles)
        assert loop.query()["launches"] >= 3  # launches of mixed pinned and staged runs
    del pinned_t
    np.testing.assert_array_equal(a.accepted, wf[:400].astype(bool))
    np.testing.assert_array_equal(a.match_len, wn[:400])
    np.testing.assert_array_equal([r.accepted for r in singles], wf[400:450].astype(bool))
    np.testing.assert_array_equal([r.match_len for r in singles], wn[400:450])
    np.testing.assert_array_equal(b.accepted, wf[450:700].astype(bool))
    np.testing.assert_array_equal(b.match_len, wn[450:700])
    np.testing.assert_array_equal(c.accepted, wf[700:].astype(bool))
    np.testing.assert_array_equal(c.match_len, wn[700:])


def test_native_chunked_staging_across_launches():
    """A pageable bulk submit of about 100 MiB of ragged requests in launches of up to
    40 MiB: each launch is staged in 16 MiB chunks (requests straddle chunk boundaries) and
    the launches' copies follow each other on the copy stream while earlier launches run;
    single submits between two bulk ones break the staged runs. Every verdict equals the
    oracle's."""
    dfa = W.dense_dfa(256, 0.0, seed=41)
    trans = np.asarray(dfa.trans).reshape(256, 256)
    trans[trans == 5] = 6
    trans[90, 2] = 5  # one (state, byte) pair enters the accepting state: matches are rare
    dfa.trans[:] = trans.reshape(-1)
    dfa.accept[5] = True
    rng = np.random.default_rng(42)
    lens = np.minimum(rng.lognormal(np.log(48 << 10), 1.2, size=1600).astype(np.int64), 6 << 20)
    off = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    data = rng.integers(0, 256, size=int(off[-1]), dtype=np.uint8)
    assert off[-1] > (80 << 20)
    wf, wn = _oracle(dfa, data, off)
    assert wf.any() and not wf.all()
    loop = NativeServingLoop({"d": dfa}, slots=4, max_batch_bytes=40 << 20,
                             max_batch_count=4096, max_inflight_per_automaton=4)
    with loop:
        half = lens.size // 2
        first = loop.submit_many("d", data, off[:half + 1], copy=False)
        singles = [loop.submit("d", data[off[i]:off[i + 1]].tobytes())
                   for i in range(half, half + 8)]
        second = loop.submit_many("d", data, off[half + 8:], copy=False)
        assert first.wait(60.0) and second.wait(60.0)
        loop.collect(singles, 60.0)
        launches = loop.stats()["launches"]
    assert launches >= 3
    np.testing.assert_array_equal(first.accepted, wf[:half].astype(bool))
    np.testing.assert_array_equal(first.match_len, wn[:half])
    np.testing.assert_array_equal(second.accepted, wf[half + 8:].astype(bool))
    np.testing.assert_array_equal(second.match_len, wn[half + 8:])
    for k, r in enumerate(singles):
        assert r.status == "ok"
        assert (r.accepted, r.match_len) == (bool(wf[half + k]), int(wn[half + k]))
