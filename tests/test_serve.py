"""Serving-loop logic on the CPU: admission by byte budget, batching, timeouts, results.

The loop's device side is replaced by a host stand-in with the same slot interface whose
verdicts come from the oracle, so these tests check the scheduling, not the kernels (the GPU
suite runs the same loop on the device, `test_hip_engine.py`).
"""

from __future__ import annotations

import collections
import threading
import time

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.serve import (Admission, Request, ServingLoop, form_batch, form_launch, pack_groups,
                          poisson_drive)
from oracle import native


class HostSlots:
    """Slot interface of gpufsm.serve.GpuSlots, computed on the host with the oracle."""

    def __init__(self, automata, n, delay=0.0, loop_ref=None):
        self.automata = automata
        self.delay = delay
        self.pending = [None] * n
        self.ready_at = [0.0] * n
        self.launch_log = []          # (slot, name, count, bytes, in_flight at launch)
        self.loop_ref = loop_ref
        self.concurrent_max = 0

    def __len__(self):
        return len(self.pending)

    def resident_bytes(self):
        return 1000

    def launch(self, i, name, data, offsets):
        a = self.automata[name]
        if hasattr(a, "trans"):
            f, n, _ = native.dfa_batch(a, data, offsets)
        else:
            f, n, _ = native.nfa_batch(a, data, offsets)
        self.pending[i] = (f, n)
        self.ready_at[i] = time.perf_counter() + self.delay
        in_flight = self.loop_ref[0].admission.in_flight if self.loop_ref else 0
        self.launch_log.append((i, name, len(offsets) - 1, int(offsets[-1]), in_flight))
        busy = sum(p is not None for p in self.pending)
        self.concurrent_max = max(self.concurrent_max, busy)

    def ready(self, i):
        return self.pending[i] is not None and time.perf_counter() >= self.ready_at[i]

    def results(self, i):
        out = self.pending[i]
        self.pending[i] = None
        return out


class GroupedHostSlots(HostSlots):
    """HostSlots with the grouped launch of gpufsm.serve.GpuSlots: each group's verdicts
    from the oracle over its slice of the packed arrays, at the group's output index."""

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.group_log = []           # automata per launch

    def launch_groups(self, i, data, offsets, table, count):
        f = np.full(count, -1, np.int32)
        n = np.full(count, -1, np.int32)
        for name, byte_at, off_at, in_at, cnt, total, max_len in table:
            assert byte_at % 256 == 0
            o = offsets[off_at:off_at + cnt + 1]
            assert o[0] == 0 and int(o[-1]) == total
            assert max_len == (int(np.diff(o).max()) if cnt else 0)
            a = self.automata[name]
            fn = native.dfa_batch if hasattr(a, "trans") else native.nfa_batch
            gf, gn, _ = fn(a, data[byte_at:byte_at + total], o)
            f[in_at:in_at + cnt], n[in_at:in_at + cnt] = gf, gn
        assert (f >= 0).all()
        self.pending[i] = (f, n)
        self.ready_at[i] = time.perf_counter() + self.delay
        self.group_log.append([t[0] for t in table])
        busy = sum(p is not None for p in self.pending)
        self.concurrent_max = max(self.concurrent_max, busy)


def test_form_launch_and_pack_groups():
    mk = lambda name, n: Request(name, bytes((7 * k + n) % 256 for k in range(n)), 0.0)  # noqa: E731
    q = collections.deque([mk("a", 10), mk("b", 5), mk("a", 300), mk("c", 0), mk("b", 7)])
    groups = form_launch(q, max_bytes=10 + 5 + 300 + 0 + 7 + 3 * 256, max_count=10)
    assert [(g, [len(r.data) for r in rs]) for g, rs in groups] == \
        [("a", [10, 300]), ("b", [5, 7]), ("c", [0])]
    assert not q
    data, offs, table, count = pack_groups(groups)
    assert count == 5 and [t[0] for t in table] == ["a", "b", "c"]
    for (name, byte_at, off_at, in_at, cnt, total, max_len), (_, rs) in zip(table, groups):
        assert byte_at % 256 == 0 and offs[off_at] == 0 and int(offs[off_at + cnt]) == total
        got = [data[byte_at + offs[off_at + k]:byte_at + offs[off_at + k + 1]].tobytes()
               for k in range(cnt)]
        assert got == [r.data for r in rs]
    # the byte cap counts the 256-byte padding of each new automaton; the head always goes
    q = collections.deque([mk("a", 100), mk("b", 100)])
    assert [g for g, _ in form_launch(q, max_bytes=300, max_count=10)] == ["a"]
    q = collections.deque([mk("a", 5000)])
    assert len(form_launch(q, max_bytes=100, max_count=4)) == 1
    q = collections.deque([mk("a", 1), mk("a", 1), mk("b", 1)])
    assert [len(rs) for _, rs in form_launch(q, max_bytes=10**6, max_count=2)] == [2]


def test_grouped_launches_mix_results_equal_alone():
    automata, pu = _automata()
    reqs = _requests(automata, pu, 150, seed=14)
    slots = GroupedHostSlots(automata, 2, delay=0.003)
    loop = ServingLoop(automata, budget_bytes=1 << 30, max_batch_bytes=64 << 10,
                       max_batch_count=64, backend=slots, group_automata=True)
    with loop:
        done = poisson_drive(loop, reqs, qps=20000, seed=15)
    assert all(r.status == "ok" for r in done)
    for (name, data), r in zip(reqs, done):
        assert (r.accepted, r.match_len) == _alone(automata, name, data)
    # at this rate launches carry several automata at once
    assert max(len(g) for g in slots.group_log) > 1
    assert loop.launches < len(reqs)


def test_inflight_cap_per_automaton_grows_launches():
    """One automaton, 4 slots, at most one launch of it in flight: requests that arrive
    meanwhile wait and leave together, results equal their lone runs."""
    automata, pu = _automata()
    reqs = [(n, d) for n, d in _requests(automata, pu, 200, seed=21) if n == "dfa"]
    slots = HostSlots(automata, 4, delay=0.004)
    loop = ServingLoop(automata, budget_bytes=1 << 30, max_batch_bytes=1 << 20,
                       max_batch_count=256, backend=slots, max_inflight_per_automaton=1)
    with loop:
        done = poisson_drive(loop, reqs, qps=20000, seed=22)
    assert all(r.status == "ok" for r in done)
    for (name, data), r in zip(reqs, done):
        assert (r.accepted, r.match_len) == _alone(automata, name, data)
    assert slots.concurrent_max == 1
    assert loop.launches < len(reqs) / 3


def _automata():
    pu = W.pattern_union(80, seed=1)
    return {"dfa": W.dense_dfa(64, 0.05, seed=2), "union": pu.nfa,
            "dfa2": W.dense_dfa(300, 0.01, seed=3)}, pu


def _requests(automata, pu, n, seed=0, max_len=5000):
    rng = np.random.default_rng(seed)
    names = sorted(automata)
    out = []
    for k in range(n):
        name = names[rng.integers(0, len(names))]
        length = int(rng.integers(0, max_len))
        if name == "union":
            d, _ = W.pattern_inputs(pu, np.array([length]), 0.7, seed=seed * 1000 + k)
        else:
            d = rng.integers(0, 256, size=length, dtype=np.uint8)
        out.append((name, d.tobytes()))
    return out


def _alone(automata, name, data):
    a = automata[name]
    d = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(0, np.uint8)
    off = np.array([0, len(data)], dtype=np.int64)
    f, n, _ = (native.dfa_batch if hasattr(a, "trans") else native.nfa_batch)(a, d, off)
    return bool(f[0]), int(n[0])


def test_admission_accounting():
    adm = Admission(100, resident_bytes=30)
    assert adm.ever_fits(70) and not adm.ever_fits(71)
    assert adm.try_admit(50) and not adm.try_admit(21) and adm.try_admit(20)
    adm.release(50)
    assert adm.in_flight == 20 and adm.try_admit(50)
    with pytest.raises(ValueError):
        Admission(10, resident_bytes=11)


def test_form_batch_groups_by_head_automaton_in_order():
    mk = lambda name, n: Request(name, b"x" * n, 0.0)  # noqa: E731
    q = collections.deque([mk("a", 10), mk("b", 5), mk("a", 20), mk("a", 40), mk("b", 1)])
    batch = form_batch(q, max_bytes=35, max_count=10)
    assert [len(r.data) for r in batch] == [10, 20]           # 40 does not fit in 35
    assert [(r.name, len(r.data)) for r in q] == [("b", 5), ("a", 40), ("b", 1)]
    batch = form_batch(q, max_bytes=100, max_count=1)
    assert [(r.name, len(r.data)) for r in batch] == [("b", 5)]
    # an oversized head still forms a batch of one
    q = collections.deque([mk("a", 500)])
    assert len(form_batch(q, max_bytes=100, max_count=4)) == 1


def test_mix_results_equal_alone_results():
    automata, pu = _automata()
    reqs = _requests(automata, pu, 120, seed=4)
    ref = [None]
    slots = HostSlots(automata, 3, delay=0.002, loop_ref=ref)
    loop = ServingLoop(automata, budget_bytes=1 << 30, max_batch_bytes=64 << 10,
                       max_batch_count=16, backend=slots)
    ref[0] = loop
    with loop:
        done = poisson_drive(loop, reqs, qps=4000, seed=5)
    assert all(r.status == "ok" for r in done)
    for (name, data), r in zip(reqs, done):
        assert (r.accepted, r.match_len) == _alone(automata, name, data)
    # continuous batching: several requests per launch and launches overlapping
    assert loop.launches < len(reqs)
    assert slots.concurrent_max >= 2
    st = loop.stats()
    assert st["ok"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


def test_budget_bounds_in_flight_bytes_and_rejects_oversized():
    automata, pu = _automata()
    reqs = _requests(automata, pu, 60, seed=6, max_len=4000)
    reqs.append(("dfa", b"\x01" * 9000))  # larger than the whole budget
    ref = [None]
    slots = HostSlots(automata, 4, delay=0.003, loop_ref=ref)
    budget = 1000 + 8000
    loop = ServingLoop(automata, budget_bytes=budget, max_batch_bytes=8000,
                       max_batch_count=64, backend=slots)
    ref[0] = loop
    with loop:
        out = [loop.submit(n, d) for n, d in reqs]
        for r in out:
            assert r.wait(30)
    assert out[-1].status == "rejected"
    assert all(r.status == "ok" for r in out[:-1])
    assert all(in_flight + 1000 <= budget for *_, in_flight in slots.launch_log)
    assert loop.admission.in_flight == 0


def test_queue_timeout():
    automata, pu = _automata()
    ref = [None]
    slots = HostSlots(automata, 1, delay=0.2, loop_ref=ref)
    loop = ServingLoop(automata, budget_bytes=1000 + 100, max_batch_bytes=100,
                       max_batch_count=1, queue_timeout_s=0.05, backend=slots)
    ref[0] = loop
    with loop:
        first = loop.submit("dfa", b"a" * 100)
        time.sleep(0.01)
        second = loop.submit("dfa", b"b" * 100)  # no budget until the first completes
        assert first.wait(5) and second.wait(5)
    assert first.status == "ok" and second.status == "timeout"


def test_unknown_automaton_and_concurrent_submitters():
    automata, pu = _automata()
    slots = HostSlots(automata, 2)
    loop = ServingLoop(automata, budget_bytes=1 << 30, backend=slots)
    with pytest.raises(KeyError):
        loop.submit("nope", b"")
    reqs = _requests(automata, pu, 80, seed=8, max_len=800)
    got = {}
    with loop:
        def client(part):
            for k in part:
                got[k] = loop.submit(*reqs[k])
        threads = [threading.Thread(target=client, args=(range(j, 80, 4),)) for j in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for r in got.values():
            assert r.wait(30)
    for k, (name, data) in enumerate(reqs):
        assert (got[k].accepted, got[k].match_len) == _alone(automata, name, data)


class FailingSlots(HostSlots):
    """Host slots whose third launch raises, as a device error would."""

    def launch(self, i, name, data, offsets):
        if len(self.launch_log) == 2:
            self.launch_log.append("boom")
            raise RuntimeError("device fault (simulated on the host)")
        super().launch(i, name, data, offsets)


def test_a_failing_launch_fails_every_pending_request_instead_of_hanging():
    dfa = W.dense_dfa(16, 0.0, seed=3)
    automata = {"d": dfa}
    slots = FailingSlots(automata, 1, delay=0.05)
    loop = ServingLoop(automata, budget_bytes=1 << 20, backend=slots, max_batch_count=1)
    with loop:
        reqs = [loop.submit("d", bytes([k]) * 100) for k in range(6)]
        assert all(r.wait(10) for r in reqs), "a request was left waiting"
        assert isinstance(loop.error, RuntimeError)
        with pytest.raises(RuntimeError, match="stopped on an error"):
            loop.submit("d", b"x")
    statuses = [r.status for r in reqs]
    assert statuses[:2] == ["ok", "ok"] and set(statuses[2:]) == {"error"}
    assert loop.stats()["error"] == 4


def test_completed_history_is_bounded():
    dfa = W.dense_dfa(16, 0.0, seed=4)
    automata = {"d": dfa}
    loop = ServingLoop(automata, budget_bytes=1 << 20, backend=HostSlots(automata, 2),
                       keep_completed=5)
    with loop:
        reqs = [loop.submit("d", b"ab" * 10) for _ in range(12)]
        assert all(r.wait(10) for r in reqs)
    assert len(loop.completed) == 5 and loop.stats()["ok"] == 5


class NotifyingSlots(HostSlots):
    """Host slots that report each completion from a timer thread, as GpuSlots' waiters do,
    and count how often the scheduler asked ready() (it sleeps between notifications)."""

    def __init__(self, automata, n, delay):
        super().__init__(automata, n, delay)
        self.notify = None
        self.done = [False] * n
        self.polls = 0

    def set_notify(self, callback):
        self.notify = callback

    def launch(self, i, name, data, offsets):
        super().launch(i, name, data, offsets)
        self.done[i] = False

        def complete():
            self.done[i] = True
            self.notify()

        threading.Timer(self.delay, complete).start()

    def ready(self, i):
        self.polls += 1
        return self.done[i]

    def results(self, i):
        self.done[i] = False
        return super().results(i)


def test_completion_notifications_wake_the_scheduler():
    """With a backend that notifies completions the scheduler sleeps until one arrives (no
    polling) and still finishes every request with its lone verdict."""
    automata, pu = _automata()
    slots = NotifyingSlots(automata, 2, delay=0.02)
    loop = ServingLoop(automata, budget_bytes=1 << 30, backend=slots, max_batch_count=4)
    reqs = _requests(automata, pu, 24, seed=11)
    with loop:
        done = [loop.submit(name, data) for name, data in reqs]
        assert all(r.wait(10) for r in done), "a notified completion was missed"
    for (name, data), r in zip(reqs, done):
        assert r.status == "ok"
        assert (r.accepted, r.match_len) == _alone(automata, name, data)
    # at most a few checks per launch: one per wake plus the pre-sleep check, never a
    # half-millisecond poll across each 20 ms launch
    assert slots.polls < 20 * len(slots.launch_log), (slots.polls, len(slots.launch_log))
