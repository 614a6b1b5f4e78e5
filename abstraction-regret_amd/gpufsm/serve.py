"""Serving loop: continuous batching of match requests against device-resident automata.

Requests ``(automaton name, bytes)`` arrive at any time from any thread. One scheduler
thread runs the loop:

1. **Completion.** Every in-flight launch slot whose completion event has fired hands its
   verdicts to the requests it carried and gives their bytes back to the budget.
2. **Admission.** Waiting requests are admitted in arrival order while the HBM byte budget
   (resident automata + the input bytes of every admitted, unfinished request) has room. A
   request that could never fit is rejected at once; one that waits longer than the queue
   timeout is rejected as timed out.
3. **Batching.** Each free slot takes the oldest admitted request and every other admitted
   request for the same automaton, up to the slot's byte and count capacity, packs them
   (uint8 bytes, int64 offsets) into pinned memory and enqueues, on the slot's own stream:
   the host-to-device copy, the launch (the persistent ragged kernels of the technique for
   that automaton) and the copy back, followed by an event. An automaton with
   ``max_inflight_per_automaton`` launches already running (2 by default) gets no new one:
   its requests wait for one of them to finish and then leave together, so under load the
   launches grow instead of multiplying (each costs the host tens of microseconds).
   With ``group_automata=True`` a slot instead takes the admitted requests in arrival
   order whatever their automata, one sub-batch per automaton, staged with one copy each
   way and launched group after group on the slot's stream. That makes fewer launches, but
   a request then completes only with the slowest group of its launch: on the configured
   mix at 16k QPS it raised the p50 from 4.6 to 27 ms, so it is off by default.

Slots have separate streams and buffers, so requests that arrive while others are on the
GPU are admitted and launched as slots and budget free up, and their work overlaps on the
device. Automata stay resident for the life of the loop (:class:`~gpufsm.batch.ResidentCache`).

The device side sits behind a small slot interface (:class:`GpuSlots`), so the scheduling
logic also runs against a host-side stand-in in the CPU test suite.

:class:`NativeServingLoop` is the same loop run by the library itself (``gfsm_server_*``
in ``csrc/server.hip``): a native scheduler thread does completion, admission, batching,
packing and launching, so no request waits for the interpreter lock. Python only submits
(one C call that copies the bytes) and collects verdicts. It has the same client surface
(``submit`` / ``Request.wait`` / ``stats`` / ``start`` / ``stop``) and is what the
benchmark serves with; :class:`ServingLoop` stays as the reference for the scheduling
rules and for host-side tests.

No upstream counterpart: upstream has no serving path.
"""

from __future__ import annotations

import collections
import itertools
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass(eq=False)
class Request:
    """One match request; :meth:`wait` blocks until it is done."""

    name: str
    data: bytes
    arrival: float
    rid: int = 0
    status: str = "queued"        # queued | admitted | running | ok | rejected | timeout | error
    accepted: bool = False
    match_len: int = 0
    done_at: float = 0.0
    _event: threading.Event = field(default_factory=threading.Event, repr=False)

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival

    def finish(self, status: str, now: float, accepted: bool = False, match_len: int = 0) -> None:
        self.status = status
        self.accepted = bool(accepted)
        self.match_len = int(match_len)
        self.done_at = now
        self._event.set()

    def wait(self, timeout: float | None = None) -> bool:
        return self._event.wait(timeout)


class Admission:
    """HBM byte budget: resident automata plus the inputs of admitted, unfinished requests."""

    def __init__(self, budget_bytes: int, resident_bytes: int = 0) -> None:
        if resident_bytes > budget_bytes:
            raise ValueError(f"resident automata need {resident_bytes} bytes, more than the "
                             f"budget of {budget_bytes}")
        self.budget = int(budget_bytes)
        self.resident = int(resident_bytes)
        self.in_flight = 0

    def ever_fits(self, nbytes: int) -> bool:
        return self.resident + nbytes <= self.budget

    def try_admit(self, nbytes: int) -> bool:
        if self.resident + self.in_flight + nbytes > self.budget:
            return False
        self.in_flight += nbytes
        return True

    def release(self, nbytes: int) -> None:
        self.in_flight -= nbytes
        assert self.in_flight >= 0


def form_batch(admitted: collections.deque, max_bytes: int, max_count: int,
               blocked=frozenset()) -> list[Request]:
    """The oldest admitted request whose automaton is not in ``blocked``, plus later ones for
    the same automaton that fit (empty when every admitted request is blocked)."""
    head = next((r for r in admitted if r.name not in blocked), None)
    if head is None:
        return []
    batch, used, keep = [], 0, collections.deque()
    while admitted:
        r = admitted.popleft()
        if (r.name == head.name and len(batch) < max_count
                and (not batch or used + len(r.data) <= max_bytes)):
            batch.append(r)
            used += len(r.data)
        else:
            keep.append(r)
    admitted.extend(keep)
    return batch


_GROUP_ALIGN = 256  # each automaton's bytes start on this boundary in a grouped launch


def form_launch(admitted: collections.deque, max_bytes: int,
                max_count: int) -> list[tuple[str, list[Request]]]:
    """The admitted requests, oldest first, that fit a launch of ``max_bytes`` /
    ``max_count`` (an oversized head alone), grouped by automaton: ``(name, requests)`` in
    order of each automaton's first request, requests in arrival order within a group."""
    if not admitted:
        return []
    groups: dict[str, list[Request]] = {}
    used = n = 0
    while admitted and n < max_count:
        r = admitted[0]
        pad = 0 if r.name in groups else _GROUP_ALIGN
        if n and used + len(r.data) + pad > max_bytes:
            break
        admitted.popleft()
        groups.setdefault(r.name, []).append(r)
        used += len(r.data) + pad
        n += 1
    return list(groups.items())


def pack_groups(groups: list[tuple[str, list[Request]]]):
    """Host arrays of a grouped launch: the bytes of every group back to back, each group at
    a multiple of 256; the groups' offset arrays back to back, each starting at 0; and per
    group ``(name, byte_start, off_start, in_start, count, total, max_len)``."""
    table, parts, offs = [], [], []
    byte_at = off_at = in_at = 0
    for name, reqs in groups:
        lens = np.fromiter((len(r.data) for r in reqs), dtype=np.int64, count=len(reqs))
        o = np.zeros(len(reqs) + 1, dtype=np.int64)
        np.cumsum(lens, out=o[1:])
        total = int(o[-1])
        table.append((name, byte_at, off_at, in_at, len(reqs), total,
                      int(lens.max()) if len(reqs) else 0))
        parts.append(b"".join(r.data for r in reqs))
        pad = (-total) % _GROUP_ALIGN
        if pad:
            parts.append(bytes(pad))
        offs.append(o)
        byte_at += total + pad
        off_at += len(reqs) + 1
        in_at += len(reqs)
    data = np.frombuffer(b"".join(parts), dtype=np.uint8) if byte_at else np.zeros(0, np.uint8)
    return data, np.concatenate(offs), table, in_at


class GpuSlots:
    """Launch slots on the GPU: each has a stream, a resident batch buffer and an event.

    A waiter thread per slot blocks on the slot's completion event (without the GIL) and
    calls the notify callback the loop registers, so the scheduler sleeps until a request
    arrives or a launch completes instead of polling the events.
    """

    def __init__(self, engine, automata: dict, slots: int, capacity_bytes: int,
                 capacity_count: int, technique: str = "auto") -> None:
        import torch

        from .batch import DeviceBatch

        self.torch = torch
        self.engine = engine
        self.automata = automata
        self.technique = technique
        empty_off = np.zeros(1, dtype=np.int64)
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        # a grouped launch pads each automaton's bytes to 256 and adds an offset entry per
        # automaton
        groups = len(automata)
        self.batches = [DeviceBatch(np.zeros(0, np.uint8), empty_off,
                                    capacity_bytes=capacity_bytes + _GROUP_ALIGN * groups,
                                    capacity_count=capacity_count + groups)
                        for _ in range(slots)]
        self.events: list = [None] * slots
        for a in automata.values():  # resident before the first request
            engine.cache.get(a)
        torch.cuda.synchronize()
        self._done = [False] * slots
        self._error: list[BaseException | None] = [None] * slots
        self._armed = [threading.Event() for _ in range(slots)]
        self._notify = None
        self._closing = False
        self._waiters = [threading.Thread(target=self._wait_slot, args=(i,), daemon=True,
                                          name=f"gpufsm-slot{i}") for i in range(slots)]
        for t in self._waiters:
            t.start()

    def __len__(self) -> int:
        return len(self.streams)

    def resident_bytes(self) -> int:
        return self.engine.cache.bytes

    def set_notify(self, callback) -> None:
        """``callback()`` is called from a waiter thread after a slot's launch completes."""
        self._notify = callback

    def _wait_slot(self, i: int) -> None:
        while True:
            self._armed[i].wait()
            self._armed[i].clear()
            if self._closing:
                return
            try:
                self.events[i].synchronize()  # releases the GIL while it blocks
            except BaseException as exc:  # surfaced by results() on the scheduler thread
                self._error[i] = exc
            self._done[i] = True
            if self._notify is not None:
                self._notify()

    def launch(self, i: int, name: str, data: np.ndarray, offsets: np.ndarray) -> None:
        if self._closing:  # no waiter would ever report this launch
            raise RuntimeError("launch on closed GPU slots")
        s, b = self.streams[i], self.batches[i]
        b.load(data, offsets, stream=s)
        self.engine.enqueue(self.automata[name], b, self.technique, stream=s)
        self.events[i] = b.download(s)
        self._done[i] = False
        self._armed[i].set()

    def launch_groups(self, i: int, data: np.ndarray, offsets: np.ndarray, table, count: int
                      ) -> None:
        """One staged copy each way for several automata's requests (see pack_groups)."""
        if self._closing:
            raise RuntimeError("launch on closed GPU slots")
        s, b = self.streams[i], self.batches[i]
        b.load_groups(data, offsets, count, stream=s)
        for name, byte_at, off_at, in_at, n, total, max_len in table:
            self.engine.enqueue_range(self.automata[name], b, self.technique, s, byte_at,
                                      off_at, in_at, n, total, max_len)
        self.events[i] = b.download(s)
        self._done[i] = False
        self._armed[i].set()

    def ready(self, i: int) -> bool:
        return self._done[i]

    def results(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        self.events[i] = None
        self._done[i] = False
        err, self._error[i] = self._error[i], None
        if err is not None:
            raise err
        return self.batches[i].results()

    def close(self) -> None:
        """Stop the waiter threads (a waiter inside a wait first sees its launch finish)."""
        self._closing = True
        for e in self._armed:
            e.set()
        for t in self._waiters:
            t.join(10.0)


class ServingLoop:
    """Continuous batching over device-resident automata; see the module docstring."""

    def __init__(self, automata: dict, *, budget_bytes: int = 32 << 30, slots: int = 8,
                 max_batch_bytes: int = 16 << 20, max_batch_count: int = 4096,
                 queue_timeout_s: float | None = None, technique: str = "auto",
                 engine=None, backend=None, keep_completed: int = 1 << 20,
                 group_automata: bool = False, max_inflight_per_automaton: int = 2) -> None:
        self.automata = dict(automata)
        self.max_batch_bytes = int(max_batch_bytes)
        self.max_batch_count = int(max_batch_count)
        self.queue_timeout_s = queue_timeout_s
        if backend is None:
            from .batch import BatchEngine

            self.engine = engine or BatchEngine(tuner=None)
            backend = GpuSlots(self.engine, self.automata, slots, self.max_batch_bytes,
                               self.max_batch_count, technique)
        self.slots = backend
        self.admission = Admission(budget_bytes, backend.resident_bytes())
        self._waiting: collections.deque[Request] = collections.deque()
        self._admitted: collections.deque[Request] = collections.deque()
        self._running: dict[int, list[Request]] = {}
        self._free = list(range(len(backend)))
        self._grouped = group_automata and hasattr(backend, "launch_groups")
        # launches in flight per automaton: at the cap, an automaton's new requests wait for
        # one of its launches to finish and then go together in a larger one
        self.max_inflight = max(1, int(max_inflight_per_automaton))
        self._inflight: collections.Counter = collections.Counter()
        self._cv = threading.Condition()
        self._ids = itertools.count()
        self._stop = False
        self._thread: threading.Thread | None = None
        # finished requests, newest last, for stats(); bounded so a long-lived loop does
        # not grow without limit
        self.completed: collections.deque[Request] = collections.deque(maxlen=keep_completed)
        self.launches = 0
        self.error: BaseException | None = None  # what stopped the scheduler, if anything
        # a backend that reports completions lets the scheduler sleep between events; one
        # that does not is polled every half millisecond while launches are in flight
        self._notified = hasattr(backend, "set_notify")
        if self._notified:
            backend.set_notify(self._wake)

    def _wake(self) -> None:
        with self._cv:
            self._cv.notify()

    # -- client side --------------------------------------------------------------------
    def submit(self, name: str, data: bytes) -> Request:
        if name not in self.automata:
            raise KeyError(f"unknown automaton {name!r}; have {sorted(self.automata)}")
        r = Request(name, bytes(data), time.perf_counter(), next(self._ids))
        with self._cv:
            if self.error is not None:
                raise RuntimeError("the serving loop stopped on an error") from self.error
            self._waiting.append(r)
            self._cv.notify()
        return r

    def start(self) -> ServingLoop:
        self._thread = threading.Thread(target=self._run, name="gpufsm-serve", daemon=True)
        self._thread.start()
        return self

    def stop(self, drain: bool = True, timeout: float = 60.0) -> None:
        if drain:
            deadline = time.perf_counter() + timeout
            while time.perf_counter() < deadline:
                with self._cv:
                    idle = not (self._waiting or self._admitted or self._running)
                if idle:
                    break
                time.sleep(0.001)
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread is not None:
            self._thread.join(timeout)
            self._thread = None
        if hasattr(self.slots, "close"):
            self.slots.close()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc) -> None:
        self.stop()

    # -- scheduler ----------------------------------------------------------------------
    def step(self) -> bool:
        """One pass of completion, admission and batching; True if anything happened."""
        now = time.perf_counter()
        busy = False
        for i, reqs in list(self._running.items()):
            if not self.slots.ready(i):
                continue
            flags, lens = self.slots.results(i)
            if not self._grouped:
                self._inflight[reqs[0].name] -= 1
            t = time.perf_counter()
            with self._cv:  # completed is read by stats() from other threads
                for r, f, n in zip(reqs, flags, lens):
                    self.admission.release(len(r.data))
                    r.finish("ok", t, bool(f), int(n))
                    self.completed.append(r)
            del self._running[i]
            self._free.append(i)
            busy = True
        with self._cv:
            while self._waiting:
                r = self._waiting[0]
                if not self.admission.ever_fits(len(r.data)) or len(r.data) > self.max_batch_bytes:
                    self._waiting.popleft()
                    r.finish("rejected", now)
                    self.completed.append(r)
                    busy = True
                    continue
                if self.admission.try_admit(len(r.data)):
                    self._waiting.popleft()
                    r.status = "admitted"
                    self._admitted.append(r)
                    busy = True
                    continue
                if self.queue_timeout_s is not None and now - r.arrival > self.queue_timeout_s:
                    self._waiting.popleft()
                    r.finish("timeout", now)
                    self.completed.append(r)
                    busy = True
                    continue
                break
        while self._free and self._admitted and self._grouped:
            groups = form_launch(self._admitted, self.max_batch_bytes, self.max_batch_count)
            batch = [r for _, rs in groups for r in rs]
            data, offsets, table, count = pack_groups(groups)
            i = self._free.pop()
            for r in batch:
                r.status = "running"
            self._running[i] = batch  # before the launch, so a failing launch fails these too
            self.slots.launch_groups(i, data, offsets, table, count)
            self.launches += 1
            busy = True
        while self._free and self._admitted:
            blocked = {k for k, v in self._inflight.items() if v >= self.max_inflight}
            batch = form_batch(self._admitted, self.max_batch_bytes, self.max_batch_count,
                               blocked)
            if not batch:
                break  # every admitted request's automaton is at its in-flight cap
            self._inflight[batch[0].name] += 1
            i = self._free.pop()
            lens = np.fromiter((len(r.data) for r in batch), dtype=np.int64, count=len(batch))
            offsets = np.zeros(len(batch) + 1, dtype=np.int64)
            np.cumsum(lens, out=offsets[1:])
            data = np.frombuffer(b"".join(r.data for r in batch), dtype=np.uint8)
            for r in batch:
                r.status = "running"
            self._running[i] = batch  # before the launch, so a failing launch fails these too
            self.slots.launch(i, batch[0].name, data, offsets)
            self.launches += 1
            busy = True
        return busy

    def _run(self) -> None:
        try:
            while True:
                with self._cv:
                    if self._stop:
                        return
                if not self.step():
                    with self._cv:
                        # a completion reported before this check is seen here; one after it
                        # notifies the wait below (the waiter needs the lock to notify)
                        done = any(self.slots.ready(i) for i in self._running)
                        if not self._waiting and not self._stop and not done:
                            if self._notified:
                                self._cv.wait(0.05)
                            else:
                                self._cv.wait(0.0005 if self._running else 0.01)
        except BaseException as exc:  # a launch or device error: fail everything pending
            self._fail_all(exc)

    def _fail_all(self, exc: BaseException) -> None:
        """Finish every unfinished request with status "error", so no caller waits forever,
        and refuse new ones. The exception is kept in :attr:`error`."""
        now = time.perf_counter()
        with self._cv:
            self.error = exc
            pending = list(self._waiting) + list(self._admitted)
            pending += [r for reqs in self._running.values() for r in reqs]
            self._waiting.clear()
            self._admitted.clear()
            self._running.clear()
            self._inflight.clear()
            for r in pending:
                if not r._event.is_set():
                    r.finish("error", now)
                    self.completed.append(r)

    # -- reporting ----------------------------------------------------------------------
    def stats(self, since: float = 0.0) -> dict:
        with self._cv:  # the scheduler appends while another thread may be asking
            snapshot = list(self.completed)
        done = [r for r in snapshot if r.arrival >= since]
        ok = [r for r in done if r.status == "ok"]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
        span = (max(r.done_at for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
        credited = sum(len(r.data) for r in ok)
        examined = sum((r.match_len if r.accepted else len(r.data)) for r in ok)
        return {
            "requests": len(done),
            "ok": len(ok),
            "rejected": sum(r.status == "rejected" for r in done),
            "timeout": sum(r.status == "timeout" for r in done),
            "error": sum(r.status == "error" for r in done),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if ok else math.nan,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if ok else math.nan,
            "credited_GBps": credited / span / 1e9 if span > 0 else math.nan,
            "examined_GBps_upper": examined / span / 1e9 if span > 0 else math.nan,
            "launches": self.launches,
        }


class NativeRequest:
    """One request of a :class:`NativeServingLoop`; :meth:`wait` collects its verdict."""

    __slots__ = ("loop", "rid", "name", "nbytes", "arrival", "status", "accepted",
                 "match_len", "done_at")

    def __init__(self, loop: "NativeServingLoop", rid: int, name: str, nbytes: int) -> None:
        self.loop = loop
        self.rid = rid
        self.name = name
        self.nbytes = nbytes
        self.arrival = 0.0
        self.status = "queued"
        self.accepted = False
        self.match_len = 0
        self.done_at = 0.0

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival

    def wait(self, timeout: float | None = None) -> bool:
        if self.status != "queued":
            return True
        return self.loop.collect([self], timeout)


class NativeBatch:
    """Requests submitted together by :meth:`NativeServingLoop.submit_many`.

    After :meth:`wait`, ``status`` (the ``GFSM_SERVE_*`` codes), ``accepted``,
    ``match_len`` and ``latency`` (seconds) are arrays in submit order.
    """

    def __init__(self, loop: "NativeServingLoop", name: str, ids: np.ndarray) -> None:
        self.loop = loop
        self.name = name
        self.ids = ids                      # ids not collected yet
        self._pos = np.arange(ids.size)     # their positions in submit order
        self._keep = None                   # the caller's bytes, read in place
        self.done = ids.size == 0
        self.status = np.zeros(ids.size, dtype=np.int32)
        self.accepted = np.zeros(ids.size, dtype=bool)
        self.match_len = np.zeros(ids.size, dtype=np.int32)
        self.latency = np.zeros(ids.size, dtype=np.float64)

    def __len__(self) -> int:
        return int(self.status.size)

    def wait(self, timeout: float | None = None) -> bool:
        """Wait up to ``timeout`` seconds (None: no limit) for every request; True when all
        have finished (their verdicts are then in the arrays)."""
        if self.done:
            return True
        lp = self.loop
        out = (lp._lib.ServerResult * self.ids.size)()
        lp._lib.check(lp._so.gfsm_server_wait(
            lp._srv, self.ids.ctypes.data_as(lp._ct.POINTER(lp._ct.c_int64)), self.ids.size,
            -1.0 if timeout is None else float(timeout), out))
        res = np.ctypeslib.as_array(out)
        fin = res["status"] != lp._lib.SERVE_PENDING  # these are collected now
        at = self._pos[fin]
        self.status[at] = res["status"][fin]
        self.accepted[at] = res["accepted"][fin] != 0
        self.match_len[at] = res["match_len"][fin]
        self.latency[at] = res["done_s"][fin] - res["arrival_s"][fin]
        self.ids = self.ids[~fin]
        self._pos = self._pos[~fin]
        self.done = self.ids.size == 0
        if self.done:
            self._keep = None
        return self.done


class NativeServingLoop:
    """The serving loop run by the library's native scheduler (``gfsm_server_*``).

    Same rules as :class:`ServingLoop` (HBM byte budget, per-automaton batching with an
    in-flight cap, one launch per free slot, rejection and queue timeout), the same client
    surface, and the same statistics; one difference: a launch takes its automaton's
    admitted requests in arrival order and stops at the first that does not fit (each
    automaton's requests wait in a queue of their own), where :class:`ServingLoop` also takes
    later ones that fit. Automata are uploaded once (their handles live as
    long as the loop). ``technique`` applies to every automaton: "auto" (resolved by the
    library per launch for the launch's shape), a technique name, or a label with tuning
    such as ``"chain:chunk=512:triple"`` (:meth:`gpufsm.tune.Choice.parse`). ``slots`` defaults to 4, the hardware queues
    a process gets: more slots share queues, and a long launch then delays the launches
    queued behind it (on the configured mix at 16k QPS, p50 0.85 ms with 4 slots, 0.91 with
    8, 1.8 with 16).
    """

    def __init__(self, automata: dict, *, budget_bytes: int = 32 << 30, slots: int = 4,
                 max_batch_bytes: int = 16 << 20, max_batch_count: int = 4096,
                 queue_timeout_s: float | None = None, technique: str = "auto",
                 max_inflight_per_automaton: int = 2, poll_us: int = 20,
                 keep_completed: int = 1 << 20) -> None:
        import ctypes

        from .backends.hip import _lib
        from .backends.hip.handles import DeviceDFA, DeviceNFA
        from .core.dfa import DFA
        from .tune import Choice

        self._ct = ctypes
        self._lib = _lib
        self._so = _lib.load()
        self.automata = dict(automata)
        self._handles: dict = {}
        self._index: dict[str, int] = {}
        self._srv = None
        cfg = _lib.ServerConfig(int(slots), int(max_batch_count), int(max_batch_bytes),
                                int(budget_bytes), int(max_inflight_per_automaton),
                                int(poll_us), float(queue_timeout_s or 0.0))
        srv = ctypes.c_void_p()
        _lib.check(self._so.gfsm_server_create(ctypes.byref(cfg), ctypes.byref(srv)))
        self._srv = srv
        choice = Choice.parse(technique)
        tuning = choice.tuning()  # the longest input is set per launch by the scheduler
        tref = ctypes.byref(tuning) if tuning is not None else None
        try:
            for name, a in self.automata.items():
                idx = ctypes.c_int32(0)
                if isinstance(a, DFA):
                    h = DeviceDFA(a)
                    tech = _lib.DFA_TECHNIQUES[choice.technique]
                    fn = self._so.gfsm_server_add_dfa
                else:
                    h = DeviceNFA(a)
                    tech = _lib.NFA_TECHNIQUES[choice.technique]
                    fn = self._so.gfsm_server_add_nfa
                self._handles[name] = h
                _lib.check(fn(self._srv, h.handle, tech, tref, ctypes.byref(idx)))
                self._index[name] = int(idx.value)
        except BaseException:
            self.close()
            raise
        self._pending: dict[int, NativeRequest] = {}
        self.completed: collections.deque = collections.deque(maxlen=keep_completed)
        self._lock = threading.Lock()
        self._started = False

    # -- lifecycle ------------------------------------------------------------------------
    def start(self) -> "NativeServingLoop":
        self._lib.check(self._so.gfsm_server_start(self._srv))
        self._started = True
        return self

    def stop(self, drain: bool = True) -> None:
        """Stop the scheduler: ``drain`` finishes every queued request first; without it the
        requests not yet launched end with status "error"."""
        if self._srv and self._started:
            self._lib.check(self._so.gfsm_server_stop(self._srv, 1 if drain else 0))
            self.collect(list(self._pending.values()), 0.0)

    def close(self) -> None:
        """Stop without draining and free the server, then the automata's handles."""
        if self._srv:
            self._so.gfsm_server_destroy(self._srv)
            self._srv = None
        for h in self._handles.values():
            h.close()
        self._handles = {}

    def __enter__(self) -> "NativeServingLoop":
        return self.start()

    def __exit__(self, *exc) -> None:
        self.stop()
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    # -- client side ----------------------------------------------------------------------
    def submit(self, name: str, data: bytes) -> NativeRequest:
        if name not in self._index:
            raise KeyError(f"unknown automaton {name!r}; have {sorted(self.automata)}")
        data = bytes(data)
        rid = self._ct.c_int64(0)
        self._lib.check(self._so.gfsm_server_submit(self._srv, self._index[name], data,
                                                    len(data), self._ct.byref(rid)))
        r = NativeRequest(self, int(rid.value), name, len(data))
        with self._lock:
            self._pending[r.rid] = r
        return r

    def submit_many(self, name: str, data, offsets, copy: bool = True) -> "NativeBatch":
        """Requests ``data[offsets[i]:offsets[i + 1]]`` for ``name`` in one call (``data``
        uint8 and ``offsets`` int64 numpy arrays). Their verdicts come back as arrays through
        the returned :class:`NativeBatch`, with no per-request Python object (they are not
        part of :meth:`stats`). ``copy=False`` lets the library read ``data`` in place
        instead of copying it first; the batch keeps a reference to it, and the caller must
        not change it until the batch has finished."""
        if name not in self._index:
            raise KeyError(f"unknown automaton {name!r}; have {sorted(self.automata)}")
        data = np.ascontiguousarray(data, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = int(offsets.size) - 1
        ids = np.empty(max(n, 0), dtype=np.int64)
        if n > 0:
            if int(offsets[0]) < 0 or int(offsets[-1]) > data.size:
                raise ValueError("offsets must lie inside the data")
            i64p = self._ct.POINTER(self._ct.c_int64)
            self._lib.check(self._so.gfsm_server_submit_many(
                self._srv, self._index[name],
                self._ct.c_void_p(data.ctypes.data if data.size else 0),
                offsets.ctypes.data_as(i64p), n, 0 if copy else self._lib.SERVE_NOCOPY,
                ids.ctypes.data_as(i64p)))
        batch = NativeBatch(self, name, ids)
        batch._keep = None if copy else data
        return batch

    def collect(self, requests: list, timeout: float | None = None) -> bool:
        """Wait up to ``timeout`` seconds (None: no limit, 0: poll) for ``requests`` and
        record the verdicts of those that finished; True when all of them have."""
        todo = [r for r in requests if r.status == "queued"]
        if not todo:
            return True
        ids = np.fromiter((r.rid for r in todo), dtype=np.int64, count=len(todo))
        out = (self._lib.ServerResult * len(todo))()
        self._lib.check(self._so.gfsm_server_wait(
            self._srv, ids.ctypes.data_as(self._ct.POINTER(self._ct.c_int64)), len(todo),
            -1.0 if timeout is None else float(timeout), out))
        done = True
        with self._lock:
            for r, res in zip(todo, out):
                if res.status == self._lib.SERVE_PENDING:
                    done = False
                    continue
                r.status = self._lib.SERVE_STATUS[res.status]
                r.accepted = bool(res.accepted)
                r.match_len = int(res.match_len)
                r.arrival = float(res.arrival_s)
                r.done_at = float(res.done_s)
                self._pending.pop(r.rid, None)
                self.completed.append(r)
        return done

    def error(self) -> str:
        """The message that stopped the scheduler ("" while it runs)."""
        buf = self._ct.create_string_buffer(1024)
        self._lib.check(self._so.gfsm_server_error(self._srv, buf, len(buf)))
        return buf.value.decode(errors="replace")

    # -- reporting ------------------------------------------------------------------------
    def query(self) -> dict:
        """The scheduler's counters (``gfsm_server_stats``)."""
        st = self._lib.ServerStats()
        self._lib.check(self._so.gfsm_server_query(self._srv, self._ct.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_ if name != "pad"}

    def stats(self, since: float = 0.0) -> dict:
        """As :meth:`ServingLoop.stats`, over the requests collected so far (requests still
        pending are polled first). Times are CLOCK_MONOTONIC, as ``time.perf_counter``."""
        self.collect(list(self._pending.values()), 0.0)
        with self._lock:
            snapshot = list(self.completed)
        done = [r for r in snapshot if r.arrival >= since]
        ok = [r for r in done if r.status == "ok"]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
        span = (max(r.done_at for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
        credited = sum(r.nbytes for r in ok)
        examined = sum((r.match_len if r.accepted else r.nbytes) for r in ok)
        return {
            "requests": len(done),
            "ok": len(ok),
            "rejected": sum(r.status == "rejected" for r in done),
            "timeout": sum(r.status == "timeout" for r in done),
            "error": sum(r.status == "error" for r in done),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if ok else math.nan,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if ok else math.nan,
            "credited_GBps": credited / span / 1e9 if span > 0 else math.nan,
            "examined_GBps_upper": examined / span / 1e9 if span > 0 else math.nan,
            "launches": int(self.query()["launches"]),
        }


def poisson_drive(loop: ServingLoop, requests: list[tuple[str, bytes]], qps: float,
                  seed: int = 0) -> list[Request]:
    """Submit ``requests`` open-loop at Poisson arrivals of rate ``qps``; wait for all."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=len(requests))
    t0 = time.perf_counter()
    at = t0 + np.cumsum(gaps)
    out = []
    for (name, data), when in zip(requests, at):
        delay = when - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        out.append(loop.submit(name, data))
    if isinstance(loop, NativeServingLoop):
        loop.collect(out)  # one wait for all of them
    for r in out:
        r.wait()
    return out
