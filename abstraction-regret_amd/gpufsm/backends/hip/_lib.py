"""ctypes binding of ``libgpufsm_hip.so`` (C ABI declared in ``include/gpufsm_hip.h``).

This module is the only place that knows the library may be absent: :func:`load` raises
:class:`HIPUnavailable` with the reason, and every caller fails loudly with it. There is no
fallback to any other implementation.
"""

from __future__ import annotations

import ctypes
import importlib.util
import os

LIB_NAME = "libgpufsm_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# status codes and technique ids (gpufsm_hip.h)
OK, EINVAL, ENOMEM, EUNSUPPORTED, EDEVICE = 0, 1, 2, 3, 4
NFA_AUTO, NFA_WORKLIST, NFA_BITPACKED, NFA_CHAIN, NFA_REGISTER = 0, 1, 2, 3, 4
DFA_AUTO, DFA_GATHER, DFA_LDS, DFA_SPLIT = 0, 1, 2, 3

NFA_TECHNIQUES = {
    "auto": NFA_AUTO,
    "worklist": NFA_WORKLIST,
    "bitpacked": NFA_BITPACKED,
    "chain": NFA_CHAIN,
    "register": NFA_REGISTER,
}
DFA_TECHNIQUES = {"auto": DFA_AUTO, "gather": DFA_GATHER, "lds": DFA_LDS, "split": DFA_SPLIT}

# Every symbol the header declares; the CPU test suite checks the library exports them all.
EXPORTS = (
    "gfsm_abi_version",
    "gfsm_last_error",
    "gfsm_device_arch",
    "gfsm_nfa_create",
    "gfsm_nfa_destroy",
    "gfsm_nfa_query",
    "gfsm_nfa_analyze",
    "gfsm_nfa_run_host",
    "gfsm_nfa_run_device",
    "gfsm_dfa_create",
    "gfsm_dfa_destroy",
    "gfsm_dfa_query",
    "gfsm_dfa_run_host",
    "gfsm_dfa_run_device",
    "gfsm_nfa_workspace_size",
    "gfsm_nfa_run_device_ws",
    "gfsm_nfa_resolve",
    "gfsm_dfa_workspace_size",
    "gfsm_dfa_run_device_ws",
    "gfsm_dfa_resolve",
    "gfsm_dfa_run_device_static",
    "gfsm_server_create",
    "gfsm_server_add_nfa",
    "gfsm_server_add_dfa",
    "gfsm_server_start",
    "gfsm_server_submit",
    "gfsm_server_submit_many",
    "gfsm_server_wait",
    "gfsm_server_query",
    "gfsm_server_error",
    "gfsm_server_stop",
    "gfsm_server_destroy",
)

TUNE_CHAIN_FIRST = 1
TUNE_CHAIN_WAVE64 = 2
TUNE_CHAIN_TRIPLE = 4


class HIPUnavailable(RuntimeError):
    """The HIP library could not be loaded; the message says why."""


class HIPError(RuntimeError):
    """A C-ABI call returned a non-zero status; ``code`` is the status."""

    def __init__(self, code: int, message: str) -> None:
        super().__init__(message)
        self.code = code


class NfaProps(ctypes.Structure):
    _fields_ = [
        ("num_states", ctypes.c_int32),
        ("start_state", ctypes.c_int32),
        ("num_sym_edges", ctypes.c_int64),
        ("num_eps_edges", ctypes.c_int64),
        ("start_accepts", ctypes.c_int32),
        ("chain_ok", ctypes.c_int32),
        ("num_tiles", ctypes.c_int32),
        ("start_always", ctypes.c_int32),
        ("max_chain", ctypes.c_int32),
        ("general_ok", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
    ]


class DfaProps(ctypes.Structure):
    _fields_ = [
        ("num_states", ctypes.c_int32),
        ("start_state", ctypes.c_int32),
        ("start_accepts", ctypes.c_int32),
        ("entry_bytes", ctypes.c_int32),
        ("lds_ok", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
    ]


class ServerConfig(ctypes.Structure):
    """``gfsm_server_config``: zero fields mean the defaults."""

    _fields_ = [
        ("slots", ctypes.c_int32),
        ("max_batch_count", ctypes.c_int32),
        ("max_batch_bytes", ctypes.c_int64),
        ("budget_bytes", ctypes.c_int64),
        ("max_inflight_per_automaton", ctypes.c_int32),
        ("poll_us", ctypes.c_int32),
        ("queue_timeout_s", ctypes.c_double),
    ]


class ServerResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("accepted", ctypes.c_int32),
        ("match_len", ctypes.c_int32),
        ("automaton", ctypes.c_int32),
        ("bytes", ctypes.c_int64),
        ("arrival_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "submitted", "completed", "rejected", "timed_out", "errors", "launches",
        "credited_bytes", "waiting", "admitted", "in_flight_bytes", "resident_bytes")] + [
        ("failed", ctypes.c_int32), ("pad", ctypes.c_int32)]


# request status (GFSM_SERVE_*)
SERVE_PENDING, SERVE_OK, SERVE_REJECTED, SERVE_TIMEOUT, SERVE_ERROR = 0, 1, 2, 3, 4
SERVE_NOCOPY = 1  # submit_many flag: the library reads the caller's buffer in place
SERVE_STATUS = {SERVE_PENDING: "queued", SERVE_OK: "ok", SERVE_REJECTED: "rejected",
                SERVE_TIMEOUT: "timeout", SERVE_ERROR: "error"}


class Tuning(ctypes.Structure):
    """``gfsm_tuning``: zero fields mean the library's choice."""

    _fields_ = [
        ("chunk_bytes", ctypes.c_int32),
        ("segment_bytes", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("max_input_bytes", ctypes.c_int32),
    ]


_c = ctypes
_vp = _c.c_void_p
_i32 = _c.c_int32
_i32p = _c.POINTER(_c.c_int32)
_i64p = _c.POINTER(_c.c_int64)
_u8p = _c.POINTER(_c.c_uint8)
_fp = _c.POINTER(_c.c_float)

_SIGNATURES = {
    "gfsm_abi_version": (_c.c_int, []),
    "gfsm_last_error": (_c.c_char_p, []),
    "gfsm_device_arch": (_c.c_int, [_i32, _c.c_char_p, _i32]),
    "gfsm_nfa_create": (_c.c_int, [_i32, _i32, _u8p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                   _c.POINTER(_vp)]),
    "gfsm_nfa_destroy": (_c.c_int, [_vp]),
    "gfsm_nfa_query": (_c.c_int, [_vp, _c.POINTER(NfaProps)]),
    "gfsm_nfa_analyze": (_c.c_int, [_i32, _i32, _u8p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                    _c.POINTER(NfaProps)]),
    "gfsm_nfa_run_host": (_c.c_int, [_vp, _i32, _u8p, _i64p, _i32, _i32p, _i32p, _fp]),
    "gfsm_nfa_run_device": (_c.c_int, [_vp, _i32, _vp, _vp, _i32, _c.c_int64, _vp, _vp, _vp]),
    "gfsm_dfa_create": (_c.c_int, [_i32, _i32, _u8p, _i32p, _c.POINTER(_vp)]),
    "gfsm_dfa_destroy": (_c.c_int, [_vp]),
    "gfsm_dfa_query": (_c.c_int, [_vp, _c.POINTER(DfaProps)]),
    "gfsm_dfa_run_host": (_c.c_int, [_vp, _i32, _u8p, _i64p, _i32, _i32p, _i32p, _fp]),
    "gfsm_dfa_run_device": (_c.c_int, [_vp, _i32, _vp, _vp, _i32, _c.c_int64, _vp, _vp, _vp]),
    "gfsm_nfa_workspace_size": (_c.c_int, [_vp, _i32, _i32, _c.c_int64, _c.POINTER(Tuning),
                                           _i64p]),
    "gfsm_nfa_run_device_ws": (_c.c_int, [_vp, _i32, _c.POINTER(Tuning), _vp, _vp, _i32,
                                          _c.c_int64, _vp, _vp, _vp, _c.c_int64, _vp]),
    "gfsm_nfa_resolve": (_c.c_int, [_vp, _i32, _c.POINTER(Tuning), _i32, _c.c_int64, _i32p]),
    "gfsm_dfa_workspace_size": (_c.c_int, [_vp, _i32, _i32, _c.c_int64, _c.POINTER(Tuning),
                                           _i64p]),
    "gfsm_dfa_run_device_ws": (_c.c_int, [_vp, _i32, _c.POINTER(Tuning), _vp, _vp, _i32,
                                          _c.c_int64, _vp, _vp, _vp, _c.c_int64, _vp]),
    "gfsm_dfa_resolve": (_c.c_int, [_vp, _i32, _c.POINTER(Tuning), _i32, _c.c_int64, _i32p]),
    "gfsm_dfa_run_device_static": (_c.c_int, [_vp, _i32, _c.POINTER(Tuning), _vp, _vp, _i32,
                                              _c.c_int64, _vp, _vp, _vp]),
    "gfsm_server_create": (_c.c_int, [_c.POINTER(ServerConfig), _c.POINTER(_vp)]),
    "gfsm_server_add_nfa": (_c.c_int, [_vp, _vp, _i32, _c.POINTER(Tuning), _i32p]),
    "gfsm_server_add_dfa": (_c.c_int, [_vp, _vp, _i32, _c.POINTER(Tuning), _i32p]),
    "gfsm_server_start": (_c.c_int, [_vp]),
    "gfsm_server_submit": (_c.c_int, [_vp, _i32, _vp, _c.c_int64, _i64p]),
    "gfsm_server_submit_many": (_c.c_int, [_vp, _i32, _vp, _i64p, _i32, _i32, _i64p]),
    "gfsm_server_wait": (_c.c_int, [_vp, _i64p, _c.c_int64, _c.c_double,
                                    _c.POINTER(ServerResult)]),
    "gfsm_server_query": (_c.c_int, [_vp, _c.POINTER(ServerStats)]),
    "gfsm_server_error": (_c.c_int, [_vp, _c.c_char_p, _i32]),
    "gfsm_server_stop": (_c.c_int, [_vp, _i32]),
    "gfsm_server_destroy": (_c.c_int, [_vp]),
}

STATIC_SCRATCH_BYTES = 32 << 20  # GFSM_STATIC_SCRATCH_BYTES

_lib: ctypes.CDLL | None = None
_load_error: str | None = None


def _share_torch_hip_runtime() -> None:
    """Make this library use the same HIP runtime as PyTorch, if PyTorch-ROCm is installed.

    PyTorch wheels carry their own ``libamdhip64.so`` (same soname as the system one) and
    reach it by file name through their RPATH. Loaded in the other order, a process would
    hold two HIP runtimes, and the second to initialise finds no device. Loading PyTorch's
    copy first (without importing torch) lets the dynamic loader bind this library to it by
    soname, and lets torch reuse the same file later. ``GFSM_HIP_RUNTIME=system`` skips this.
    """
    if os.environ.get("GFSM_HIP_RUNTIME") == "system":
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.origin:
        return
    candidate = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(candidate):
        try:
            ctypes.CDLL(candidate, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass


def load() -> ctypes.CDLL:
    """Load the library once; raise :class:`HIPUnavailable` naming the cause otherwise."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise HIPUnavailable(_load_error)
    _share_torch_hip_runtime()
    if not os.path.exists(LIB_PATH):
        _load_error = (
            f"{LIB_NAME} is not built ({LIB_PATH} missing); run "
            f"`python abstraction-regret_amd/csrc/build.py` or __graft_entry__.build()"
        )
        raise HIPUnavailable(_load_error)
    try:
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    except (OSError, AttributeError) as exc:
        _load_error = f"{LIB_NAME} failed to load: {type(exc).__name__}: {exc}"
        raise HIPUnavailable(_load_error) from exc
    _lib = lib
    return lib


def check(rc: int) -> None:
    """Raise :class:`HIPError` with the library's message for a non-zero status."""
    if rc != OK:
        msg = load().gfsm_last_error()
        raise HIPError(rc, msg.decode() if msg else f"status {rc}")


MEMCPY_H2D = 1  # hipMemcpyHostToDevice
MEMCPY_D2H = 2  # hipMemcpyDeviceToHost
_memcpy_async = None


def memcpy_async(dst: int, src: int, nbytes: int, kind: int, stream: int) -> None:
    """``hipMemcpyAsync`` on a raw stream handle (pinned host memory for a true async copy).

    The symbol is looked up through this library's handle, so it is the HIP runtime the
    library itself is bound to. Staging copies of the serving loop go through here instead
    of PyTorch's stream context managers, which cost tens of microseconds per launch.
    """
    global _memcpy_async
    if nbytes <= 0:
        return
    if _memcpy_async is None:
        fn = load().hipMemcpyAsync
        fn.restype = ctypes.c_int
        fn.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        _memcpy_async = fn
    rc = _memcpy_async(dst, src, nbytes, kind, stream)
    if rc != 0:
        raise HIPError(rc, f"hipMemcpyAsync of {nbytes} bytes failed with HIP error {rc}")


_graph_fns = None


def _graph_api():
    global _graph_fns
    if _graph_fns is None:
        lib = load()  # the HIP runtime this library is bound to (see memcpy_async)
        launch = lib.hipGraphLaunch
        launch.restype = ctypes.c_int
        launch.argtypes = [_vp, _vp]
        nodes = lib.hipGraphGetNodes
        nodes.restype = ctypes.c_int
        nodes.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_size_t)]
        _graph_fns = (launch, nodes)
    return _graph_fns


def graph_launch(graph_exec: int, stream: int) -> None:
    """``hipGraphLaunch`` of an instantiated graph on a raw stream handle (no stream context
    switch, which costs more host time than the launch)."""
    rc = _graph_api()[0](graph_exec, stream)
    if rc != 0:
        raise HIPError(rc, f"hipGraphLaunch failed with HIP error {rc}")


def graph_node_count(graph: int) -> int:
    """Number of nodes of a captured (kept) graph: ``hipGraphGetNodes`` with no node array."""
    n = ctypes.c_size_t(0)
    rc = _graph_api()[1](graph, None, ctypes.byref(n))
    if rc != 0:
        raise HIPError(rc, f"hipGraphGetNodes failed with HIP error {rc}")
    return int(n.value)


def device_arch(ordinal: int = 0) -> str:
    buf = ctypes.create_string_buffer(64)
    check(load().gfsm_device_arch(ordinal, buf, 64))
    return buf.value.decode()


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array's data."""
    return a.ctypes.data_as(ctypes.POINTER(ctype))
