"""Device-resident automata: Python owners of the C-ABI handles.

A :class:`DeviceNFA` / :class:`DeviceDFA` uploads the automaton once (``gfsm_*_create``)
and frees it on :meth:`close` (or garbage collection). Runs take either host numpy
buffers (synchronous, ``gfsm_*_run_host``) or raw device pointers on a stream
(asynchronous, ``gfsm_*_run_device``).
"""

from __future__ import annotations

import ctypes

import numpy as np

from ...core.dfa import DFA
from ...core.nfa import NFA
from . import _lib
from ._lib import DfaProps, NfaProps, check, ptr


def _i32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a if a.size else np.zeros(1, dtype=np.int32)


def _as_batch(data: np.ndarray, offsets: np.ndarray) -> tuple[np.ndarray, np.ndarray, int]:
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    count = int(offsets.size) - 1
    if count < 0:
        raise ValueError("offsets must have num_strings + 1 entries")
    # the library copies data[0:offsets[-1]]: offsets past the buffer would over-read it
    if count > 0 and not (0 <= int(offsets[-1]) <= data.size):
        raise ValueError(f"offsets end at {int(offsets[-1])}, past the {data.size}-byte data")
    if data.size == 0:
        data = np.zeros(1, dtype=np.uint8)
    return data, offsets, count


def analyze_nfa(nfa: NFA) -> NfaProps:
    """Structural properties of ``nfa`` computed on the host (no device, no upload)."""
    acc = np.ascontiguousarray(nfa.accept, dtype=np.uint8)
    arrays = [_i32(a) for a in (nfa.sym_row_ptr, nfa.sym_targets, nfa.sym_symbols,
                                nfa.eps_row_ptr, nfa.eps_targets)]
    props = NfaProps()
    check(_lib.load().gfsm_nfa_analyze(
        int(nfa.num_states), int(nfa.start_state), ptr(acc, ctypes.c_uint8),
        *[ptr(a, ctypes.c_int32) for a in arrays], ctypes.byref(props)))
    return props


def make_tuning(chunk_bytes: int = 0, segment_bytes: int = 0, flags: int = 0,
                max_input_bytes: int = 0):
    """A ``gfsm_tuning`` (None when every field is the library's choice).

    ``max_input_bytes`` is the batch's longest input (0 = not known); the DFA ``auto`` rule
    walks one lane per input only when it is known and short against the batch.
    """
    if not (chunk_bytes or segment_bytes or flags or max_input_bytes):
        return None
    return _lib.Tuning(int(chunk_bytes), int(segment_bytes), int(flags),
                       min(int(max_input_bytes), 2**31 - 1))


class _Handle:
    _destroy = ""
    _prefix = ""

    def __init__(self) -> None:
        self._h = ctypes.c_void_p()

    def _fn(self, name: str):
        return getattr(_lib.load(), f"{self._prefix}_{name}")

    def resolve(self, count: int, total_bytes: int, technique: int = 0, tuning=None) -> int:
        """The concrete technique id ``technique`` (auto = 0) resolves to for this shape."""
        out = ctypes.c_int32(0)
        check(self._fn("resolve")(self.handle, technique,
                                  ctypes.byref(tuning) if tuning is not None else None,
                                  int(count), int(total_bytes), ctypes.byref(out)))
        return int(out.value)

    def workspace_size(self, count: int, total_bytes: int, technique: int = 0,
                       tuning=None) -> int:
        """Device scratch bytes :meth:`run_device_ws` needs for this batch shape."""
        out = ctypes.c_int64(0)
        check(self._fn("workspace_size")(
            self.handle, technique, int(count), int(total_bytes),
            ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(out)))
        return int(out.value)

    def run_device_ws(self, d_data: int, d_offsets: int, count: int, total_bytes: int,
                      d_flags: int, d_lens: int, d_ws: int, ws_bytes: int, stream: int = 0,
                      technique: int = 0, tuning=None) -> None:
        """Enqueue on ``stream`` with caller scratch: no allocation, graph-capturable."""
        check(self._fn("run_device_ws")(
            self.handle, technique, ctypes.byref(tuning) if tuning is not None else None,
            ctypes.c_void_p(d_data), ctypes.c_void_p(d_offsets), int(count),
            int(total_bytes), ctypes.c_void_p(d_flags), ctypes.c_void_p(d_lens),
            ctypes.c_void_p(d_ws), int(ws_bytes), ctypes.c_void_p(stream)))

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._h:
            raise RuntimeError("handle is closed")
        return self._h

    def close(self) -> None:
        if self._h:
            getattr(_lib.load(), self._destroy)(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()


class DeviceNFA(_Handle):
    """An :class:`~gpufsm.core.nfa.NFA` resident in HBM in both device layouts."""

    _destroy = "gfsm_nfa_destroy"
    _prefix = "gfsm_nfa"

    def __init__(self, nfa: NFA) -> None:
        super().__init__()
        lib = _lib.load()
        self.nfa = nfa
        acc = np.ascontiguousarray(nfa.accept, dtype=np.uint8)
        arrays = [_i32(a) for a in (nfa.sym_row_ptr, nfa.sym_targets, nfa.sym_symbols,
                                    nfa.eps_row_ptr, nfa.eps_targets)]
        check(lib.gfsm_nfa_create(
            int(nfa.num_states), int(nfa.start_state), ptr(acc, ctypes.c_uint8),
            *[ptr(a, ctypes.c_int32) for a in arrays], ctypes.byref(self._h)))
        self.props = NfaProps()
        check(lib.gfsm_nfa_query(self._h, ctypes.byref(self.props)))

    def run_host(self, data: np.ndarray, offsets: np.ndarray, technique: int = _lib.NFA_AUTO
                 ) -> tuple[np.ndarray, np.ndarray, float]:
        data, offsets, count = _as_batch(data, offsets)
        flags = np.zeros(max(count, 1), dtype=np.int32)
        lens = np.zeros(max(count, 1), dtype=np.int32)
        ms = ctypes.c_float(0.0)
        check(_lib.load().gfsm_nfa_run_host(
            self.handle, technique, ptr(data, ctypes.c_uint8), ptr(offsets, ctypes.c_int64),
            count, ptr(flags, ctypes.c_int32), ptr(lens, ctypes.c_int32), ctypes.byref(ms)))
        return flags[:count], lens[:count], float(ms.value)

    def run_device(self, d_data: int, d_offsets: int, count: int, total_bytes: int,
                   d_flags: int, d_lens: int, stream: int = 0,
                   technique: int = _lib.NFA_AUTO) -> None:
        """Enqueue on ``stream``; ``total_bytes`` must equal ``offsets[count]``."""
        check(_lib.load().gfsm_nfa_run_device(
            self.handle, technique, ctypes.c_void_p(d_data), ctypes.c_void_p(d_offsets), count,
            int(total_bytes), ctypes.c_void_p(d_flags), ctypes.c_void_p(d_lens),
            ctypes.c_void_p(stream)))


class DeviceDFA(_Handle):
    """A :class:`~gpufsm.core.dfa.DFA` resident in HBM (uint16 or uint32 entries)."""

    _destroy = "gfsm_dfa_destroy"
    _prefix = "gfsm_dfa"

    def __init__(self, dfa: DFA) -> None:
        super().__init__()
        lib = _lib.load()
        self.dfa = dfa
        acc = np.ascontiguousarray(dfa.accept, dtype=np.uint8)
        trans = np.ascontiguousarray(dfa.trans, dtype=np.int32)
        if trans.size != int(dfa.num_states) * 256 or acc.size != int(dfa.num_states):
            raise ValueError("DFA arrays do not match num_states")
        check(lib.gfsm_dfa_create(int(dfa.num_states), int(dfa.start_state),
                                  ptr(acc, ctypes.c_uint8), ptr(trans, ctypes.c_int32),
                                  ctypes.byref(self._h)))
        self.props = DfaProps()
        check(lib.gfsm_dfa_query(self._h, ctypes.byref(self.props)))

    def run_host(self, data: np.ndarray, offsets: np.ndarray, technique: int = _lib.DFA_AUTO
                 ) -> tuple[np.ndarray, np.ndarray, float]:
        data, offsets, count = _as_batch(data, offsets)
        flags = np.zeros(max(count, 1), dtype=np.int32)
        lens = np.zeros(max(count, 1), dtype=np.int32)
        ms = ctypes.c_float(0.0)
        check(_lib.load().gfsm_dfa_run_host(
            self.handle, technique, ptr(data, ctypes.c_uint8), ptr(offsets, ctypes.c_int64),
            count, ptr(flags, ctypes.c_int32), ptr(lens, ctypes.c_int32), ctypes.byref(ms)))
        return flags[:count], lens[:count], float(ms.value)

    def run_device(self, d_data: int, d_offsets: int, count: int, total_bytes: int,
                   d_flags: int, d_lens: int, stream: int = 0,
                   technique: int = _lib.DFA_AUTO) -> None:
        """Enqueue on ``stream``; ``total_bytes`` must equal ``offsets[count]``."""
        check(_lib.load().gfsm_dfa_run_device(
            self.handle, technique, ctypes.c_void_p(d_data), ctypes.c_void_p(d_offsets), count,
            int(total_bytes), ctypes.c_void_p(d_flags), ctypes.c_void_p(d_lens),
            ctypes.c_void_p(stream)))
