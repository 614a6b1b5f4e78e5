"""ctypes binding of the C oracle (``liboracle.so``). TEST INFRASTRUCTURE ONLY.

Batch verdicts of `src/gpufsm/reference.py` semantics on all host cores, with examined
byte counts. See ``fsm_oracle.c`` for the restatement and its two NFA modes.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from oracle.reference import DFA, NFA

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build() -> str:
    """Compile ``liboracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_nfa_batch.restype = ctypes.c_int
        lib.oracle_nfa_batch.argtypes = [
            ctypes.c_int32, ctypes.c_int32, _u8p,
            _i32p, _i32p, _i32p, _i32p, _i32p,
            _u8p, _i64p, ctypes.c_int32, _i32p, _i32p, _i64p,
            ctypes.c_int32, ctypes.c_int32,
        ]
        lib.oracle_dfa_batch.restype = ctypes.c_int
        lib.oracle_dfa_batch.argtypes = [
            ctypes.c_int32, ctypes.c_int32, _u8p, _i32p,
            _u8p, _i64p, ctypes.c_int32, _i32p, _i32p, _i64p, ctypes.c_int32,
        ]
        _lib = lib
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _batch_arrays(data: np.ndarray, offsets: np.ndarray):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    if data.size == 0:
        data = np.zeros(1, dtype=np.uint8)
    return data, np.ascontiguousarray(offsets, dtype=np.int64)


def nfa_batch(nfa: NFA, data: np.ndarray, offsets: np.ndarray, *, indexed: bool = True,
              threads: int | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """``(flags int32, lens int32, examined int64)`` for every input of the batch."""
    lib = _load()
    data, offsets = _batch_arrays(data, offsets)
    count = offsets.size - 1
    flags = np.zeros(max(count, 1), dtype=np.int32)
    lens = np.zeros(max(count, 1), dtype=np.int32)
    examined = np.zeros(max(count, 1), dtype=np.int64)
    acc = np.ascontiguousarray(nfa.accept, dtype=np.uint8)
    arrs = [np.ascontiguousarray(a, dtype=np.int32) if a.size else np.zeros(1, np.int32)
            for a in (nfa.sym_row_ptr, nfa.sym_targets, nfa.sym_symbols, nfa.eps_row_ptr,
                      nfa.eps_targets)]
    rc = lib.oracle_nfa_batch(
        nfa.num_states, nfa.start_state, _p(acc, _u8p),
        *[_p(a, _i32p) for a in arrs],
        _p(data, _u8p), _p(offsets, _i64p), count,
        _p(flags, _i32p), _p(lens, _i32p), _p(examined, _i64p),
        1 if indexed else 0, threads or default_threads(),
    )
    if rc != 0:
        raise RuntimeError(f"oracle_nfa_batch failed with {rc}")
    return flags[:count], lens[:count], examined[:count]


def dfa_batch(dfa: DFA, data: np.ndarray, offsets: np.ndarray, *,
              threads: int | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    lib = _load()
    data, offsets = _batch_arrays(data, offsets)
    count = offsets.size - 1
    flags = np.zeros(max(count, 1), dtype=np.int32)
    lens = np.zeros(max(count, 1), dtype=np.int32)
    examined = np.zeros(max(count, 1), dtype=np.int64)
    acc = np.ascontiguousarray(dfa.accept, dtype=np.uint8)
    trans = np.ascontiguousarray(dfa.trans, dtype=np.int32)
    rc = lib.oracle_dfa_batch(
        dfa.num_states, dfa.start_state, _p(acc, _u8p), _p(trans, _i32p),
        _p(data, _u8p), _p(offsets, _i64p), count,
        _p(flags, _i32p), _p(lens, _i32p), _p(examined, _i64p), threads or default_threads(),
    )
    if rc != 0:
        raise RuntimeError(f"oracle_dfa_batch failed with {rc}")
    return flags[:count], lens[:count], examined[:count]
