"""Batched, device-resident ``run_batch``: the resident form of the hot path.

The plain ``gpufsm.run_batch`` builds a fresh executor per call (as upstream does,
`src/gpufsm/api.py:48-49`): automaton preprocessing and upload, packing, host-to-device
copy, one launch, copy back. :class:`BatchEngine` keeps what repeats resident instead:

* automata stay in HBM in a :class:`ResidentCache` keyed by automaton identity (``NFA`` and
  ``DFA`` hash by identity, `src/gpufsm/core/nfa.py:43-54`), least recently used first out
  under a byte budget;
* a :class:`DeviceBatch` holds packed inputs (uint8 bytes, int64 offsets), outputs and the
  launch workspace in HBM; inputs are staged through pinned host memory with asynchronous
  copies;
* the technique and its tuning come from the :class:`~gpufsm.tune.Autotuner`;
* a launch is the persistent ragged kernel sequence of the chosen technique enqueued with
  caller workspace (no allocation, no memset), so it can be captured once into a HIP graph
  and replayed (``graph=True``).

PyTorch provides device memory, pinned host memory, streams, events and graph capture; the
work itself is the C ABI of ``libgpufsm_hip.so``.
"""

from __future__ import annotations

import collections
import concurrent.futures
import itertools
import threading
import time
import weakref

import numpy as np

from .backends.hip import _lib
from .backends.hip.handles import DeviceDFA, DeviceNFA
from .core.packing import pack_bytes
from .core.registry import Kind
from .core.result import Result, batch_results
from .tune import Autotuner, Choice


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.HIPUnavailable("BatchEngine needs a visible GPU (torch.cuda.is_available())")
    return torch


class ResidentCache:
    """Device-resident automaton handles keyed by identity, LRU under ``budget_bytes``.

    ``on_evict(handle)`` runs before an evicted handle's device memory is freed (the engine
    drops the graphs captured on it).
    """

    def __init__(self, budget_bytes: int | None = None, on_evict=None) -> None:
        self.budget_bytes = budget_bytes
        self.on_evict = on_evict
        self._items: collections.OrderedDict[int, tuple[object, object]] = collections.OrderedDict()
        self._lock = threading.Lock()
        self.uploads = 0

    @property
    def bytes(self) -> int:
        return sum(h.props.device_bytes for _, h in self._items.values())

    def get(self, automaton):
        key = id(automaton)
        # one lock over lookup and upload: two threads missing on the same automaton must
        # not both upload it
        with self._lock:
            hit = self._items.get(key)
            if hit is not None and hit[0] is automaton:
                self._items.move_to_end(key)
                return hit[1]
            kind = Kind.of(automaton)
            handle = DeviceNFA(automaton) if kind is Kind.NFA else DeviceDFA(automaton)
            self.uploads += 1
            self._items[key] = (automaton, handle)  # holds the automaton: id stays unique
            if self.budget_bytes is not None:
                while len(self._items) > 1 and self.bytes > self.budget_bytes:
                    _, (_, old) = self._items.popitem(last=False)
                    if self.on_evict is not None:
                        self.on_evict(old)
                    old.close()
        return handle

    def close(self) -> None:
        with self._lock:
            for _, h in self._items.values():
                h.close()
            self._items.clear()


_batch_serials = itertools.count(1)

_stream_pool: list = []
_stream_pool_lock = threading.Lock()


def slot_streams(n: int) -> list:
    """``n`` streams from a process-wide pool, created once (default priority) and reused.

    The HIP runtime gives each new stream the next of the process's hardware queues in turn
    (4 by default, one of them the default stream's), so streams created afresh call after
    call wrap around onto queues that other streams hold, and launches on two streams that
    share a queue run one after the other. Slots that take their streams from this pool
    keep the same queues across calls: up to three slots on three distinct queues, as long
    as the pool's streams are the process's first.
    """
    torch = _torch()
    with _stream_pool_lock:
        while len(_stream_pool) < n:
            _stream_pool.append(torch.cuda.Stream())
        return _stream_pool[:n]


_capture_stream = None


def capture_stream():
    """The one stream graph captures run on (created after three pooled slot streams, so
    that it takes no queue before them; captured work never runs on it)."""
    global _capture_stream
    if _capture_stream is None:
        slot_streams(3)
        with _stream_pool_lock:
            if _capture_stream is None:
                _capture_stream = _torch().cuda.Stream()
    return _capture_stream

# Host copies into pinned staging are split over a few threads above this size: one thread
# copies at a fraction of what the host's memory and the PCIe link carry, and numpy drops
# the GIL inside a contiguous copy.
_COPY_SPLIT_BYTES = 16 << 20
_COPY_THREADS = 8
_copy_pool: concurrent.futures.ThreadPoolExecutor | None = None
_copy_pool_lock = threading.Lock()


def _copy_into(dst: np.ndarray, src: np.ndarray) -> None:
    """``dst[:] = src`` for 1-D uint8 arrays, split over threads when large."""
    global _copy_pool
    n = int(src.size)
    if n < _COPY_SPLIT_BYTES:
        dst[:] = src
        return
    with _copy_pool_lock:
        if _copy_pool is None:
            _copy_pool = concurrent.futures.ThreadPoolExecutor(_COPY_THREADS,
                                                               thread_name_prefix="gpufsm-copy")
    step = -(-n // _COPY_THREADS)
    futs = [_copy_pool.submit(np.copyto, dst[a:a + step], src[a:a + step])
            for a in range(0, n, step)]
    for f in futs:
        f.result()


class DeviceBatch:
    """Packed inputs, outputs and launch workspace resident in HBM.

    ``count`` inputs at ``data[offsets[i]:offsets[i+1]]``. The buffers keep their addresses
    for the life of the batch (a graph captured on them stays valid); :meth:`load` replaces
    the contents with another batch that fits the capacity. ``serial`` is unique for the life
    of the process (``id()`` of a collected batch is reused by the next one), so graphs
    captured on a batch are keyed by it and never replayed on another batch's buffers.
    """

    def __init__(self, data: np.ndarray, offsets: np.ndarray, *, capacity_bytes: int = 0,
                 capacity_count: int = 0, stream=None) -> None:
        torch = _torch()
        self._torch = torch
        self.serial = next(_batch_serials)
        count = int(offsets.size) - 1
        total = int(offsets[-1]) if count >= 0 else 0
        self.cap_bytes = max(int(capacity_bytes), total, 1)
        self.cap_count = max(int(capacity_count), count, 1)
        dev = torch.device("cuda")
        self.data = torch.empty(self.cap_bytes, dtype=torch.uint8, device=dev)
        self.offsets = torch.zeros(self.cap_count + 1, dtype=torch.int64, device=dev)
        self.flags = torch.zeros(self.cap_count, dtype=torch.int32, device=dev)
        self.lens = torch.zeros(self.cap_count, dtype=torch.int32, device=dev)
        self._h_data = torch.empty(self.cap_bytes, dtype=torch.uint8, pin_memory=True)
        self._h_off = torch.empty(self.cap_count + 1, dtype=torch.int64, pin_memory=True)
        self._h_out = torch.empty(2 * self.cap_count, dtype=torch.int32, pin_memory=True)
        self.workspace = None
        self._retired = []   # superseded workspaces: graphs captured on them still point there
        self.count = 0
        self.total = 0
        self.max_len = 0     # longest input of the loaded batch
        # event after the last copy from or into the pinned buffers: the copies are raw
        # hipMemcpyAsync calls, which PyTorch's allocators do not track, so it is waited for
        # before a pinned buffer is rewritten and before the buffers are released
        self._staged = None
        self._event = None   # one event, re-recorded after each staging copy (see _mark)
        # the buffers were allocated, and zero-filled, on the current stream; a batch staged
        # on another stream (non-blocking, as PyTorch's streams are) must not copy into them
        # before that stream has got there, or a late fill overwrites the staged bytes
        _order_after_current(torch, stream)
        self.load(data, offsets, stream=stream)

    def __del__(self) -> None:
        ev = getattr(self, "_staged", None)
        if ev is not None:
            try:
                ev.synchronize()
            except Exception:  # interpreter shutdown or a failed device: nothing to wait on
                pass

    def load(self, data, offsets: np.ndarray, stream=None) -> None:
        """Stage a packed batch (asynchronously on ``stream``) into the resident buffers.

        ``data`` may also be a list of byte arrays that lie back to back in the batch (the
        offsets index their concatenation); each is copied straight into the staging buffer.
        """
        torch = self._torch
        parts = data if isinstance(data, (list, tuple)) else [data]
        offsets = np.asarray(offsets, dtype=np.int64)
        count = int(offsets.size) - 1
        if count < 0:
            raise ValueError("offsets must have count + 1 entries")
        total = int(offsets[-1])
        # the kernels trust the offsets: check them here, where it costs O(count) on the host
        lengths = np.diff(offsets)
        if offsets[0] != 0 or (lengths < 0).any():
            raise ValueError("offsets must start at 0 and never decrease")
        if count and int(lengths.max()) > np.iinfo(np.int32).max:
            raise ValueError("an input is longer than 2^31 - 1 bytes (match_len is int32)")
        size = sum(len(p) for p in parts)
        if size < total:
            raise ValueError(f"offsets end at {total}, past the {size}-byte data")
        if count > self.cap_count or total > self.cap_bytes:
            raise ValueError(f"batch of {count} inputs / {total} bytes exceeds the capacity "
                             f"{self.cap_count} / {self.cap_bytes}")
        stream = stream or torch.cuda.current_stream()
        # the previous copy from or into the pinned buffers may still be in flight
        if self._staged is not None:
            self._staged.synchronize()
        staging, at = self._h_data.numpy(), 0
        for p in parts:
            n = min(len(p), total - at)
            if n > 0:
                _copy_into(staging[at:at + n], np.asarray(p[:n], dtype=np.uint8))
                at += n
        self._h_off[: count + 1].numpy()[:] = offsets
        # asynchronous copies on the stream's raw handle (no stream context switch)
        s = stream.cuda_stream
        _lib.memcpy_async(self.data.data_ptr(), self._h_data.data_ptr(), total, _lib.MEMCPY_H2D, s)
        _lib.memcpy_async(self.offsets.data_ptr(), self._h_off.data_ptr(), 8 * (count + 1),
                          _lib.MEMCPY_H2D, s)
        self._mark(stream)
        self.count, self.total = count, total
        self.max_len = int(lengths.max()) if count else 0

    def load_groups(self, data: np.ndarray, offsets: np.ndarray, count: int,
                    stream=None) -> None:
        """Stage several sub-batches at once: ``data`` holds them back to back (each at a
        256-byte aligned start), ``offsets`` their offset arrays back to back (each starting
        at 0), ``count`` the inputs of all of them. The caller launches each sub-batch over
        its own slice (:meth:`BatchEngine.enqueue_range`); the verdicts of input k of the
        whole group land at index k. The caller has checked every sub-batch's offsets."""
        torch = self._torch
        total, n_off = int(data.size), int(offsets.size)
        if total > self.cap_bytes or n_off > self.cap_count + 1 or count > self.cap_count:
            raise ValueError(f"grouped batch of {count} inputs / {total} bytes exceeds the "
                             f"capacity {self.cap_count} / {self.cap_bytes}")
        stream = stream or torch.cuda.current_stream()
        if self._staged is not None:
            self._staged.synchronize()
        _copy_into(self._h_data[:total].numpy(), np.asarray(data, dtype=np.uint8))
        self._h_off[:n_off].numpy()[:] = offsets
        s = stream.cuda_stream
        _lib.memcpy_async(self.data.data_ptr(), self._h_data.data_ptr(), total, _lib.MEMCPY_H2D, s)
        _lib.memcpy_async(self.offsets.data_ptr(), self._h_off.data_ptr(), 8 * n_off,
                          _lib.MEMCPY_H2D, s)
        self._mark(stream)
        self.count, self.total, self.max_len = count, total, 0

    def download(self, stream):
        """Enqueue the copy of the verdicts into the pinned output buffer on ``stream``;
        returns the event recorded after it (the verdicts are in place once it completes)."""
        c, s = self.count, stream.cuda_stream
        h = self._h_out.data_ptr()
        _lib.memcpy_async(h, self.flags.data_ptr(), 4 * c, _lib.MEMCPY_D2H, s)
        _lib.memcpy_async(h + 4 * self.cap_count, self.lens.data_ptr(), 4 * c, _lib.MEMCPY_D2H, s)
        return self._mark(stream)

    def _mark(self, stream):
        """Record the batch's event after the copies just enqueued on ``stream``. One event
        serves every copy: a pinned buffer is rewritten only after the previous copy's event
        has been waited for, so no one still waits on an older recording."""
        if self._event is None:
            self._event = self._torch.cuda.Event()
        self._event.record(stream)
        self._staged = self._event
        return self._event

    def results(self) -> tuple[np.ndarray, np.ndarray]:
        """The verdicts of the last :meth:`download` (once its stream has passed it)."""
        out = self._h_out.numpy()
        c = self.count
        return out[:c].copy(), out[self.cap_count:self.cap_count + c].copy()

    def ensure_workspace(self, nbytes: int, stream=None) -> int:
        """Device pointer of a workspace of at least ``nbytes`` (256-byte aligned), to be used
        on ``stream`` (the current stream if None)."""
        torch = self._torch
        if self.workspace is None or self.workspace.numel() < nbytes:
            if self.workspace is not None:
                self._retired.append(self.workspace)
            self.workspace = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")
            # allocated on the current stream: its memory may have served work still queued
            # there, which another stream must not overtake
            _order_after_current(torch, stream)
        return self.workspace.data_ptr()

    def fetch(self, stream=None) -> tuple[np.ndarray, np.ndarray]:
        """Copy the verdicts back (waits for ``stream``)."""
        stream = stream or self._torch.cuda.current_stream()
        self.download(stream)
        stream.synchronize()
        return self.results()


def _order_after_current(torch, stream) -> None:
    """Make ``stream`` wait for the work queued so far on the current stream (a no-op when
    they are the same stream): device memory allocated, or filled, on the current stream is
    then safe to use on ``stream``."""
    if stream is None:
        return
    cur = torch.cuda.current_stream()
    if stream != cur:
        stream.wait_stream(cur)


def _techniques(kind: Kind) -> dict[str, int]:
    return _lib.NFA_TECHNIQUES if kind is Kind.NFA else _lib.DFA_TECHNIQUES


class BatchEngine:
    """Resident automata + resident batches + autotuned, optionally graph-captured launches."""

    def __init__(self, *, budget_bytes: int | None = None, tuner: Autotuner | None | bool = True,
                 arch: str | None = None, graph_min_kernels: int = 3) -> None:
        self._torch = _torch()
        self.cache = ResidentCache(budget_bytes, on_evict=self._drop_graphs_of)
        if tuner is True:
            tuner = Autotuner()
        self.tuner = tuner or None
        self.arch = arch or _lib.device_arch(0)
        # (id(handle), batch serial, choice, count, total, longest input) -> (graph, kernels);
        # the cache holds every handle it names, and a batch's graphs are dropped when the
        # batch is collected
        self._graphs: dict[tuple, tuple] = {}
        self._graphs_lock = threading.Lock()
        self.captures = 0
        # hipGraphLaunch costs about 17 us of host time on ROCm, three times a kernel launch:
        # a launch sequence of fewer kernels than this is launched directly, not replayed
        self.graph_min_kernels = int(graph_min_kernels)
        self.replays = 0

    def _drop_graphs_of(self, handle) -> None:
        """Forget the graphs captured on ``handle``: its device memory is about to be freed."""
        with self._graphs_lock:
            for key in [k for k in self._graphs if k[0] == id(handle)]:
                del self._graphs[key]

    def _drop_graphs_of_batch(self, serial: int) -> None:
        """Forget the graphs captured on the batch with ``serial`` (it has been collected)."""
        with self._graphs_lock:
            for key in [k for k in self._graphs if k[1] == serial]:
                del self._graphs[key]

    # -- technique selection ------------------------------------------------------------
    def choose(self, automaton, batch: DeviceBatch, technique: str = "auto",
               stream=None) -> Choice:
        """The technique + tuning a launch of ``batch`` uses (``technique`` unless "auto";
        a label such as ``"chain:chunk=2048"`` fixes the tuning too)."""
        if technique != "auto" or self.tuner is None or batch.count == 0:
            return Choice.parse(technique)
        kind = Kind.of(automaton)
        dev = self.cache.get(automaton)
        key = self.tuner.signature(kind.value, self.arch, dev.props, batch.count, batch.total)
        offsets = None

        def timer(choice: Choice, give_up: float | None = None):
            nonlocal offsets
            torch = self._torch
            if offsets is None:
                offsets = batch.offsets[: batch.count + 1].cpu().numpy()
            k = self.tuner.sample_count(kind.value, offsets, batch.count)
            s = stream or torch.cuda.current_stream()
            w0 = torch.cuda.Event(enable_timing=True)
            w1 = torch.cuda.Event(enable_timing=True)
            w0.record(s)
            self._launch(dev, kind, batch, choice, s, count=k)  # warm-up (and attributes)
            w1.record(s)
            w1.synchronize()
            times = [w0.elapsed_time(w1) / 1e3]
            # at least one timed run after the warm-up, which may include one-time work
            # (the chain lane tier builds its byte-triple lists on first use)
            repeats = self.tuner.repeats if give_up is None or times[0] < give_up else 1
            for _ in range(repeats):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(s)
                self._launch(dev, kind, batch, choice, s, count=k)
                e1.record(s)
                e1.synchronize()
                times.append(e0.elapsed_time(e1) / 1e3)
            c = k
            flags = batch.flags[:c].cpu().numpy()
            lens = batch.lens[:c].cpu().numpy().astype(np.int64)
            span = np.diff(offsets[: c + 1])
            examined = int(np.where(flags != 0, lens, span).sum())
            return min(times), max(examined, 1)

        return self.tuner.decide(key, kind.value, dev.props, batch.count, batch.total,
                                 timer).choice

    # -- launches -----------------------------------------------------------------------
    def _launch(self, dev, kind: Kind, batch: DeviceBatch, choice: Choice, stream,
                count: int | None = None) -> None:
        tech = _techniques(kind)[choice.technique]
        c = batch.count if count is None else count
        total = int(batch.total) if count is None else None
        if total is None:
            total = int(batch.offsets[c].item())
        # the longest input of a prefix is at most the batch's
        tuning = choice.tuning(max(batch.max_len, 1))
        need = dev.workspace_size(c, total, tech, tuning)
        ws = batch.ensure_workspace(need, stream)
        dev.run_device_ws(batch.data.data_ptr(), batch.offsets.data_ptr(), c, total,
                          batch.flags.data_ptr(), batch.lens.data_ptr(), ws,
                          batch.workspace.numel(), stream.cuda_stream, tech, tuning)

    def enqueue_range(self, automaton, batch: DeviceBatch, technique: str, stream,
                      byte_start: int, off_start: int, in_start: int, count: int, total: int,
                      max_len: int) -> None:
        """Enqueue one run over a slice of ``batch`` (one sub-batch of :meth:`DeviceBatch.
        load_groups`): inputs from data byte ``byte_start`` (256-byte aligned) with offsets
        from entry ``off_start``, verdicts into index ``in_start`` onward. Runs of one batch
        on one stream run in order, so they share the batch's workspace. ``technique`` is
        fixed or "auto" (resolved by the library for the slice's shape; not autotuned)."""
        if count == 0:
            return
        kind = Kind.of(automaton)
        dev = self.cache.get(automaton)
        choice = Choice.parse(technique)
        tech = _techniques(kind)[choice.technique]
        tuning = choice.tuning(max(int(max_len), 1))
        need = dev.workspace_size(count, total, tech, tuning)
        ws = batch.ensure_workspace(need, stream)
        dev.run_device_ws(batch.data.data_ptr() + byte_start, batch.offsets.data_ptr() + 8 * off_start,
                          count, total, batch.flags.data_ptr() + 4 * in_start,
                          batch.lens.data_ptr() + 4 * in_start, ws, batch.workspace.numel(),
                          stream.cuda_stream, tech, tuning)

    def enqueue(self, automaton, batch: DeviceBatch, technique: str = "auto", stream=None,
                graph: bool = False) -> Choice:
        """Enqueue one run of ``automaton`` over ``batch`` on ``stream``; returns the choice.

        With ``graph=True`` the launch sequence is captured into a HIP graph on first use
        (per automaton, batch, choice and shape) and replayed afterwards with
        ``hipGraphLaunch``; a sequence of fewer than ``graph_min_kernels`` kernels (one-kernel
        techniques) is launched directly instead, since a graph launch costs more host time
        than it saves there.
        """
        torch = self._torch
        stream = stream or torch.cuda.current_stream()
        kind = Kind.of(automaton)
        dev = self.cache.get(automaton)
        choice = self.choose(automaton, batch, technique, stream)
        if batch.count == 0:
            return choice
        if not graph:
            self._launch(dev, kind, batch, choice, stream)
            return choice
        key = (id(dev), batch.serial, choice, batch.count, batch.total, batch.max_len)
        with self._graphs_lock:
            entry = self._graphs.get(key)
        if entry is None:
            # one eager run first: it sizes the workspace and sets kernel attributes, so the
            # capture records kernel launches only
            self._launch(dev, kind, batch, choice, stream)
            stream.synchronize()
            g = torch.cuda.CUDAGraph(keep_graph=True)
            side = capture_stream()
            side.wait_stream(stream)
            with torch.cuda.graph(g, stream=side):
                self._launch(dev, kind, batch, choice, torch.cuda.current_stream())
            kernels = _lib.graph_node_count(g.raw_cuda_graph())  # kernel nodes only
            if kernels >= self.graph_min_kernels:
                g.instantiate()
            entry = (g, kernels)
            with self._graphs_lock:
                first = not any(k[1] == batch.serial for k in self._graphs)
                self._graphs[key] = entry
            if first:
                weakref.finalize(batch, _drop_batch_graphs, weakref.ref(self), batch.serial)
            self.captures += 1
            stream.wait_stream(side)
        g, kernels = entry
        if kernels < self.graph_min_kernels:
            self._launch(dev, kind, batch, choice, stream)
        else:
            _lib.graph_launch(g.raw_cuda_graph_exec(), stream.cuda_stream)
            self.replays += 1
        return choice

    def run_packed(self, automaton, data: np.ndarray, offsets: np.ndarray,
                   technique: str = "auto") -> tuple[np.ndarray, np.ndarray, float]:
        """(flags, lens, kernel_ms) of one batch given packed host arrays."""
        torch = self._torch
        stream = torch.cuda.current_stream()
        batch = DeviceBatch(data, offsets, stream=stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        self.enqueue(automaton, batch, technique, stream)
        e1.record(stream)
        flags, lens = batch.fetch(stream)
        return flags, lens, e0.elapsed_time(e1)

    def auto_coalesce(self, offsets, max_bytes: int = 4 << 30) -> int:
        """Batches per launch for a stream of batches shaped like ``offsets``: enough for 256
        inputs per CU, at most ``max_bytes`` of input per launch (at least one batch)."""
        offsets = np.asarray(offsets, dtype=np.int64)
        count = max(int(offsets.size) - 1, 1)
        total = max(int(offsets[-1]) if offsets.size else 0, 1)
        torch = self._torch
        cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        want = -(-256 * cus // count)
        return max(1, min(want, int(max_bytes) // total))

    def run_stream(self, automaton, batches, technique: str = "auto", depth: int = 2,
                   graph: bool = False, capacity_bytes: int = 0, capacity_count: int = 0,
                   coalesce: int | str = 1, coalesce_bytes: int = 4 << 30):
        """Verdicts ``(flags, lens)`` of a stream of host batches ``(data, offsets)``, in order.

        ``depth`` slots run at once, each with its own stream (from the process-wide pool)
        and resident batch buffers, launched directly: graph replays on several streams at
        once overlapped less on ROCm (2.9 ms per batch of the ragged mix against 2.25 ms,
        three slots), and a graph saves microseconds of launch work against milliseconds of
        kernels; ``graph=True`` replays a captured graph per slot instead. While the GPU
        walks batch k, the host copies batch k + 1 into the
        next slot's pinned staging and its host-to-device copy is enqueued, so the staging
        cost of a host-resident stream overlaps the kernels instead of adding to them. A
        slot is reused once its previous verdicts have been taken. With ``coalesce`` above
        1, up to that many consecutive batches share one slot and one launch (their bytes
        back to back): a launch whose passes are latency-bound (a DFA table read from L2)
        then walks several batches' chains at once. ``coalesce="auto"`` sizes the groups
        from the first batch: enough batches for 256 inputs per CU (where the DFA's `auto`
        walks one lane per input, with no repair passes), at most ``coalesce_bytes`` of
        input per launch; throughput for latency, since a batch's verdicts come with its
        group's launch. Slots must hold the first launch's bytes and inputs unless
        ``capacity_bytes`` / ``capacity_count`` say more.
        """
        torch = self._torch
        if depth < 1:
            raise ValueError("depth must be >= 1")
        batches = iter(batches)
        if coalesce == "auto":
            first = next(batches, None)
            if first is None:
                return
            coalesce = self.auto_coalesce(first[1], coalesce_bytes)
            batches = itertools.chain([first], batches)
        if not isinstance(coalesce, int) or coalesce < 1:
            raise ValueError("coalesce must be an int >= 1 or 'auto'")
        slots: list[tuple[DeviceBatch, object]] = []
        pending: collections.deque = collections.deque()  # (slot, event, counts) in order

        def groups():
            group: list = []
            for b in batches:
                group.append(b)
                if len(group) == coalesce:
                    yield group
                    group = []
            if group:
                yield group

        def results(slot: int, counts: list[int]):
            f, n = slots[slot][0].results()
            at = 0
            for c in counts:
                yield f[at:at + c], n[at:at + c]
                at += c

        for k, group in enumerate(groups()):
            counts = [int(np.asarray(o).size) - 1 for _, o in group]
            if len(group) == 1:
                data, offsets = group[0]
            else:  # one launch over the group: bytes back to back, offsets shifted
                data, offs, base = [], [np.zeros(1, np.int64)], 0
                for d, o in group:
                    o = np.asarray(o, dtype=np.int64)
                    if o.size < 1 or o[0] != 0:
                        raise ValueError("offsets must start at 0 and never decrease")
                    data.append(d[: int(o[-1])])
                    offs.append(o[1:] + base)
                    base += int(o[-1])
                offsets = np.concatenate(offs)
            j = k % depth
            while pending and any(p[0] == j for p in pending):
                slot, ev, cts = pending.popleft()
                ev.synchronize()
                yield from results(slot, cts)
            if j == len(slots):
                st = slot_streams(depth)[j]  # pooled: the same hardware queues every call
                slots.append((DeviceBatch(data, offsets, capacity_bytes=capacity_bytes,
                                          capacity_count=capacity_count, stream=st), st))
            else:
                slots[j][0].load(data, offsets, stream=slots[j][1])
            batch, st = slots[j]
            self.enqueue(automaton, batch, technique, stream=st, graph=graph)
            pending.append((j, batch.download(st), counts))
        while pending:
            slot, ev, cts = pending.popleft()
            ev.synchronize()
            yield from results(slot, cts)

    def run_batch(self, automaton, inputs: list[bytes], technique: str = "auto") -> list[Result]:
        """Drop-in ``run_batch`` over the resident path (same results, same Result layout)."""
        if not inputs:
            return []
        t0 = time.perf_counter()
        data, offsets = pack_bytes(inputs)
        transfer_ms = (time.perf_counter() - t0) * 1000.0
        flags, lens, kernel_ms = self.run_packed(automaton, data, offsets, technique)
        return batch_results(flags, lens, kernel_ms, transfer_ms)

    def close(self) -> None:
        with self._graphs_lock:
            self._graphs.clear()
        self.cache.close()


def _drop_batch_graphs(engine_ref, serial: int) -> None:
    engine = engine_ref()
    if engine is not None:
        engine._drop_graphs_of_batch(serial)
