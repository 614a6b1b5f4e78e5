"""Registry and public API (CPU). Mirrors `tests/test_registry.py:47-172`,
`tests/test_api.py:11-41` and the registry/gate parts of `tests/test_validation.py:105-132`.

The CPU oracle is registered as ``Backend.CPU`` by ``tests/conftest.py``.
"""

from __future__ import annotations

import re
from pathlib import Path

import pytest

from fixtures import EXAMPLES
from gpufsm import (
    Backend,
    DFABuilder,
    Kind,
    NFABuilder,
    benchmark,
    list_kinds,
    list_techniques,
    random_dfa,
    run,
    run_batch,
)
from gpufsm.core.registry import _DEFAULTS, _REGISTRY, get_factory, register
from oracle.gate import compare, matches, require
from oracle.reference import simulate, simulate_dfa

ROOT = Path(__file__).resolve().parent.parent


def _nfa():
    b = NFABuilder()
    s0, s1 = b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, "a", s1)
    return b.build()


def _dfa():
    b = DFABuilder()
    s0, s1 = b.add_state(), b.add_state(accept=True)
    b.set_start(s0)
    b.add_transition(s0, ord("a"), s1)
    return b.build()


def test_kind_dispatch():
    assert Kind.of(_nfa()) is Kind.NFA and Kind.of(_dfa()) is Kind.DFA
    with pytest.raises(TypeError, match="expected an NFA or a DFA"):
        Kind.of("x")  # type: ignore[arg-type]


def test_cpu_oracle_backend_registered():
    assert list_kinds(Backend.CPU) == [Kind.NFA, Kind.DFA]
    assert list_techniques(Backend.CPU, Kind.NFA) == ["reference", "bitmap"]
    assert list_techniques(Backend.CPU, Kind.DFA) == ["reference"]


def test_hip_techniques_registered():
    assert list_techniques(Backend.HIP, Kind.NFA) == ["auto", "worklist", "bitpacked", "chain",
                                                      "register"]
    assert list_techniques(Backend.HIP, Kind.DFA) == ["auto", "gather", "lds", "split"]
    assert _DEFAULTS[(Kind.NFA, Backend.HIP)] == "auto"
    assert _DEFAULTS[(Kind.DFA, Backend.HIP)] == "auto"


def test_run_dispatches_to_cpu_oracles():
    nfa, dfa = _nfa(), _dfa()
    assert (run(nfa, b"a", backend=Backend.CPU).accepted, 1) == (True, 1)
    for data in (b"a", b"b", b""):
        r = run(dfa, data, backend="cpu")
        assert (r.accepted, r.match_len) == simulate_dfa(dfa, data)


def test_run_batch_loop_fallback_puts_batch_time_on_first_result():
    dfa = random_dfa(32, seed=7)
    batch = [b"abc", b"", b"zzzz"]
    res = run_batch(dfa, batch, backend=Backend.CPU)
    assert [(r.accepted, r.match_len) for r in res] == [simulate_dfa(dfa, b) for b in batch]
    assert all(r.kernel_ms == 0.0 for r in res[1:])
    assert run_batch(dfa, [], backend=Backend.CPU) == []


def test_unknown_technique_and_backend_messages():
    with pytest.raises(KeyError, match="available: "):
        run(_dfa(), b"a", backend=Backend.CPU, technique="bitmap")
    with pytest.raises(KeyError) as exc:
        get_factory(Kind.NFA, Backend.CUDA, None)
    assert "cuda" in str(exc.value) and "gpufsm env" in str(exc.value)
    with pytest.raises(KeyError, match="not registered"):
        get_factory(Kind.NFA, Backend.HIP, "dense")


def test_second_default_raises():
    register(Kind.DFA, Backend.WARP, "one", default=True)(lambda a, t: None)
    try:
        with pytest.raises(RuntimeError, match="two defaults"):
            register(Kind.DFA, Backend.WARP, "two", default=True)(lambda a, t: None)
    finally:
        for key in [k for k in _REGISTRY if k[1] == Backend.WARP]:
            del _REGISTRY[key]
        _DEFAULTS.pop((Kind.DFA, Backend.WARP), None)


def test_every_registered_pair_declares_a_default():
    pairs = {(k, b) for k, b, _ in _REGISTRY}
    assert not [p for p in pairs if p not in _DEFAULTS]
    for (kind, backend), tech in _DEFAULTS.items():
        assert (kind, backend, tech) in _REGISTRY


def test_hip_sources_declare_one_default_per_kind():
    text = (ROOT / "abstraction-regret_amd/gpufsm/backends/hip/__init__.py").read_text()
    found = re.findall(r'register\(Kind\.(\w+), Backend\.HIP, _tech, default=\(_tech == "(\w+)"\)', text)
    assert sorted(found) == [("DFA", "auto"), ("NFA", "auto")]


def test_examples_on_cpu_backend():
    for name, make in EXAMPLES.items():
        nfa, labelled = make()
        for data, want in labelled:
            r = run(nfa, data, backend=Backend.CPU)
            assert r.accepted == want and (r.accepted, r.match_len) == simulate(nfa, data)
            r2 = run(nfa, data, backend=Backend.CPU, technique="bitmap")
            assert (r2.accepted, r2.match_len) == (r.accepted, r.match_len)


def test_benchmark_stats():
    nfa, _ = EXAMPLES["ab_star_c_plus_d"]()
    stats = benchmark(nfa, b"abcd" * 64, backend=Backend.CPU, repeats=5, warmup=2)
    assert stats.n == 5 and len(stats.raw_ms) == 5
    assert stats.mean_ms >= 0 and stats.ci95_ms >= 0
    row = stats.as_row()
    assert row["backend"] == "cpu" and row["n"] == 5 and row["technique"] == "reference"
    with pytest.raises(ValueError):
        benchmark(nfa, b"abcd", backend=Backend.CPU, repeats=0)


def test_gate_is_never_vacuous():
    with pytest.raises(ValueError, match="vacuously"):
        compare(_nfa(), [], Backend.CPU)
    with pytest.raises(ValueError, match="limit must be"):
        require(_nfa(), [b"a"], Backend.CPU, limit=0)
    with pytest.raises(ValueError, match="limit must be"):
        matches(_nfa(), [b"a"], Backend.CPU, limit=-1)
    require(_nfa(), [b"a", b"b"], Backend.CPU)
    require(_nfa(), [b"a", b"b"], Backend.CPU, technique="bitmap")


def test_gate_catches_a_lying_backend(monkeypatch):
    """A backend that corrupts one verdict must make the gate fail (cf. upstream
    `tests/test_bench_harness.py:109-128`)."""
    import oracle.gate as g

    real = g.run_batch

    def lying(automaton, inputs, backend, technique=None):
        res = real(automaton, inputs, backend=backend, technique=technique)
        res[1].accepted = not res[1].accepted
        return res

    monkeypatch.setattr(g, "run_batch", lying)
    with pytest.raises(g.OracleMismatch, match="first at index 1"):
        require(_nfa(), [b"a", b"b", b"aa"], Backend.CPU)
