"""C-ABI boundary checks that need no GPU (cf. upstream `tests/test_cuda_api_surface.py:36-105`,
which pins its pybind11 surface three ways).

* the header, the ctypes table and the shared library agree on the exported symbols;
* host-only entry points work: version, structural analysis, argument validation;
* every failure comes back as a status code plus a message, never a crash.
"""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.backends.hip.handles import analyze_nfa
from gpufsm.core.nfa import NFA, NFABuilder

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "gpufsm_hip.h"


def _header_functions() -> set[str]:
    text = HEADER.read_text()
    return set(re.findall(r"^\s*(?:int|const char \*)\s*(gfsm_\w+)\s*\(", text, re.M))


def test_header_and_ctypes_table_agree():
    assert _header_functions() == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _lib.EXPORTS:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gfsm_\w+)", out))
    assert exported == set(_lib.EXPORTS)


def test_header_constants_match_python():
    text = HEADER.read_text()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define (GFSM_\w+) (\d+)", text))
    assert consts["GFSM_ABI_VERSION"] == _lib.load().gfsm_abi_version()
    assert consts["GFSM_NFA_WORKLIST"] == _lib.NFA_WORKLIST
    assert consts["GFSM_NFA_BITPACKED"] == _lib.NFA_BITPACKED
    assert consts["GFSM_NFA_CHAIN"] == _lib.NFA_CHAIN
    assert consts["GFSM_DFA_GATHER"] == _lib.DFA_GATHER
    assert consts["GFSM_DFA_LDS"] == _lib.DFA_LDS
    assert consts["GFSM_DFA_SPLIT"] == _lib.DFA_SPLIT
    assert consts["GFSM_EINVAL"] == _lib.EINVAL
    assert consts["GFSM_EUNSUPPORTED"] == _lib.EUNSUPPORTED
    assert consts["GFSM_TUNE_CHAIN_FIRST"] == _lib.TUNE_CHAIN_FIRST
    assert consts["GFSM_TUNE_CHAIN_WAVE64"] == _lib.TUNE_CHAIN_WAVE64
    assert [consts[f"GFSM_SERVE_{k}"] for k in ("PENDING", "OK", "REJECTED", "TIMEOUT",
                                                 "ERROR")] == [
        _lib.SERVE_PENDING, _lib.SERVE_OK, _lib.SERVE_REJECTED, _lib.SERVE_TIMEOUT,
        _lib.SERVE_ERROR]
    assert consts["GFSM_SERVE_NOCOPY"] == _lib.SERVE_NOCOPY


def test_props_struct_layout(tmp_path):
    """The ctypes mirrors have the C structs' size and field offsets (compiled with gcc)."""
    fields = {"gfsm_nfa_props": _lib.NfaProps, "gfsm_dfa_props": _lib.DfaProps,
              "gfsm_tuning": _lib.Tuning, "gfsm_server_config": _lib.ServerConfig,
              "gfsm_server_result": _lib.ServerResult, "gfsm_server_stats": _lib.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gpufsm_hip.h"',
             "int main(void) {"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.splitlines())
    for cname, py in fields.items():
        assert int(got[cname]) == ctypes.sizeof(py)
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, fname


def test_analyze_pattern_union():
    pu = W.pattern_union(300, seed=1)
    p = analyze_nfa(pu.nfa)
    assert p.chain_ok == 1 and p.start_always == 1 and p.start_accepts == 0
    assert p.num_states == pu.nfa.num_states
    assert p.max_chain == int(pu.lengths.max())
    assert p.num_tiles == -(-(pu.nfa.num_states - 1) // 4096)


def test_analyze_rejects_non_chain_shapes():
    nfa, _ = W.workload_plain()
    assert analyze_nfa(nfa).chain_ok == 0
    b = NFABuilder()
    s = [b.add_state() for _ in range(3)]
    b.set_start(s[0])
    b.add_transition(s[0], 256, s[0])
    b.add_transition(s[0], "a", s[1])
    b.add_transition(s[1], "b", s[2])
    b.set_accept(s[2])
    assert analyze_nfa(b.build()).chain_ok == 1
    b.add_transition(s[2], "c", s[2])  # self-loop on a chain state: a cycle
    assert analyze_nfa(b.build()).chain_ok == 0
    b2 = NFABuilder()
    t = [b2.add_state() for _ in range(3)]
    b2.add_transition(t[0], "a", t[1])
    b2.add_transition(t[0], "x", t[1])
    b2.add_transition(t[1], "b", t[2])
    b2.add_transition(t[0], "b", t[2])  # same labels from start and chain: homogeneous
    assert analyze_nfa(b2.build()).chain_ok == 1
    b2.add_transition(t[0], "q", t[2])  # now inhomogeneous
    assert analyze_nfa(b2.build()).chain_ok == 0
    b3 = NFABuilder()
    u = [b3.add_state() for _ in range(2)]
    b3.add_epsilon(u[0], u[1])
    p = analyze_nfa(b3.build())
    assert p.chain_ok == 0 and p.num_eps_edges == 1


def test_analyze_start_closure_accepts():
    b = NFABuilder()
    s0, s1 = b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_epsilon(s0, s1)
    assert analyze_nfa(b.build()).start_accepts == 1


def _raw_nfa(n, start, row, tgt, sym, erow=None, etgt=None) -> NFA:
    return NFA(num_states=n, start_state=start, accept=np.zeros(n, dtype=bool),
               sym_row_ptr=np.array(row, np.int32), sym_targets=np.array(tgt, np.int32),
               sym_symbols=np.array(sym, np.int32),
               eps_row_ptr=np.array(erow if erow else [0] * (n + 1), np.int32),
               eps_targets=np.array(etgt if etgt else [], np.int32))


@pytest.mark.parametrize("bad,msg", [
    (lambda: _raw_nfa(2, 5, [0, 1, 1], [1], [97]), "start_state"),
    (lambda: _raw_nfa(2, 0, [0, 1, 1], [7], [97]), "sym_targets"),
    (lambda: _raw_nfa(2, 0, [0, 1, 1], [1], [300]), "sym_symbols"),
    (lambda: _raw_nfa(2, 0, [0, 2, 1], [1, 1], [97, 98]), "non-decreasing"),
    (lambda: _raw_nfa(2, 0, [0, 1, 1], [1], [97], [0, 1, 1], [9]), "eps_targets"),
])
def test_malformed_automata_are_rejected_with_a_message(bad, msg):
    with pytest.raises(_lib.HIPError, match=msg) as exc:
        analyze_nfa(bad())
    assert exc.value.code == _lib.EINVAL


def test_null_handles_are_errors_not_crashes():
    lib = _lib.load()
    flags = np.zeros(1, np.int32)
    rc = lib.gfsm_nfa_run_host(None, 0, None, None, 0, _lib.ptr(flags, ctypes.c_int32),
                               _lib.ptr(flags, ctypes.c_int32), None)
    assert rc == _lib.EINVAL and b"null handle" in lib.gfsm_last_error()
    assert lib.gfsm_nfa_destroy(None) == _lib.OK
    assert lib.gfsm_dfa_query(None, None) == _lib.EINVAL
    out = ctypes.c_int64(0)
    assert lib.gfsm_nfa_workspace_size(None, 0, 1, 1, None, ctypes.byref(out)) == _lib.EINVAL
    assert lib.gfsm_dfa_workspace_size(None, 0, 1, 1, None, ctypes.byref(out)) == _lib.EINVAL
    tech = ctypes.c_int32(0)
    assert lib.gfsm_nfa_resolve(None, 0, None, 1, 1, ctypes.byref(tech)) == _lib.EINVAL
    assert lib.gfsm_dfa_run_device_ws(None, 0, None, None, None, 0, 0, None, None, None, 0,
                                      None) == _lib.EINVAL


def test_dfa_create_validates_before_touching_the_device():
    lib = _lib.load()
    trans = np.full(2 * 256, 5, dtype=np.int32)  # target 5 out of range
    acc = np.zeros(2, np.uint8)
    h = ctypes.c_void_p()
    rc = lib.gfsm_dfa_create(2, 0, _lib.ptr(acc, ctypes.c_uint8), _lib.ptr(trans, ctypes.c_int32),
                             ctypes.byref(h))
    assert rc == _lib.EINVAL and b"out of range" in lib.gfsm_last_error() and not h


def test_batch_offsets_must_stay_inside_the_data():
    """The host form copies data[0:offsets[-1]]; offsets past the buffer are refused before
    anything reaches the library."""
    from gpufsm.backends.hip.handles import _as_batch

    data = np.zeros(10, dtype=np.uint8)
    _as_batch(data, np.array([0, 4, 10], dtype=np.int64))
    with pytest.raises(ValueError, match="past"):
        _as_batch(data, np.array([0, 4, 11], dtype=np.int64))
    with pytest.raises(ValueError, match="num_strings"):
        _as_batch(data, np.array([], dtype=np.int64))
    d, o, n = _as_batch(np.zeros(0, dtype=np.uint8), np.array([0], dtype=np.int64))
    assert n == 0 and d.size == 1
