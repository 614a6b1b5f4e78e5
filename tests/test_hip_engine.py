"""Resident batch engine, autotuner, graph replay and serving loop on the GPU.

Every verdict is checked against the oracle; graph replays against the eager run of the
same launch (several replays, with other eager work between them).
"""

from __future__ import annotations

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.batch import BatchEngine, DeviceBatch
from gpufsm.core.packing import pack_bytes
from gpufsm.serve import ServingLoop, poisson_drive
from gpufsm.tune import Autotuner
from fixtures import random_nfa
from oracle import native

pytestmark = pytest.mark.gpu


def _oracle(a, data, off):
    fn = native.dfa_batch if hasattr(a, "trans") else native.nfa_batch
    f, n, _ = fn(a, data, off)
    return f, n


@pytest.fixture(scope="module")
def engine(tmp_path_factory):
    path = tmp_path_factory.mktemp("tune") / "table.json"
    eng = BatchEngine(tuner=Autotuner(str(path), sample_bytes=4 << 20, repeats=2))
    yield eng
    eng.close()


def test_engine_matches_oracle_on_every_kind(engine):
    import random

    pu = W.pattern_union(300, seed=21)
    lens = np.random.default_rng(21).integers(0, 20000, size=200)
    ud, uo = W.pattern_inputs(pu, lens, 0.7, seed=22)
    gen = random.Random(23)
    gnfa = random_nfa(gen, 300, b"abcdefgh", max_sym=3, max_eps=1, p_accept=0.01, p_any=0.02)
    gd = np.frombuffer(bytes(gen.choice(b"abcdefghi") for _ in range(40000)), dtype=np.uint8)
    go = np.linspace(0, gd.size, 101).astype(np.int64)
    dfa = W.dense_dfa(4096, 0.0, seed=24)
    dd, do = W.uniform_batch(64, 1 << 16, seed=25)
    small = W.dense_dfa(256, 0.001, seed=26)
    for a, d, o in ((pu.nfa, ud, uo), (gnfa, gd, go), (dfa, dd, do), (small, dd, do)):
        f, n, _ = engine.run_packed(a, d, o)
        wf, wn = _oracle(a, d, o)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
    table = engine.tuner.table
    assert len(table) >= 3
    for entry in table.values():
        assert entry["rates"] and entry["choice"]["technique"]


def test_run_batch_is_a_drop_in(engine):
    pu = W.pattern_union(100, seed=31)
    lens = np.random.default_rng(31).integers(0, 3000, size=50)
    d, o = W.pattern_inputs(pu, lens, 0.7, seed=32)
    inputs = [d[o[i]:o[i + 1]].tobytes() for i in range(50)]
    res = engine.run_batch(pu.nfa, inputs)
    wf, wn = _oracle(pu.nfa, *pack_bytes(inputs))
    assert [(r.accepted, r.match_len) for r in res] == [(bool(a), int(b)) for a, b in zip(wf, wn)]
    assert res[0].kernel_ms > 0 and all(r.kernel_ms == 0 for r in res[1:])
    assert engine.run_batch(pu.nfa, []) == []


@pytest.mark.parametrize("technique", ["chain", "split"])
def test_graph_replays_match_eager(engine, technique):
    import torch

    if technique == "chain":
        pu = W.pattern_union(200, seed=41)
        a = pu.nfa
        d, o = W.pattern_inputs(pu, np.full(64, 9000, dtype=np.int64), 0.7, seed=42)
    else:
        a = W.dense_dfa(256, 0.0005, seed=43)
        d, o = W.uniform_batch(32, 1 << 15, seed=44)
    other = W.dense_dfa(64, 0.01, seed=45)
    od, oo = W.uniform_batch(16, 4096, seed=46)
    wf, wn = _oracle(a, d, o)
    s = torch.cuda.current_stream()
    batch = DeviceBatch(d, o, stream=s)
    obatch = DeviceBatch(od, oo, stream=s)
    engine.enqueue(a, batch, technique, stream=s)
    f, n = batch.fetch(s)
    np.testing.assert_array_equal(f, wf)
    np.testing.assert_array_equal(n, wn)
    for rep in range(4):
        batch.flags.fill_(-1)
        batch.lens.fill_(-1)
        engine.enqueue(a, batch, technique, stream=s, graph=True)
        f, n = batch.fetch(s)
        np.testing.assert_array_equal(f, wf, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(n, wn, err_msg=f"replay {rep}")
        engine.enqueue(other, obatch, "gather", stream=s)  # eager work between replays
        obatch.fetch(s)


def test_run_stream_auto_coalesce_groups_small_batches():
    """coalesce="auto": batches of 2000 inputs are grouped until a launch gives every CU 256
    inputs (capped by bytes per launch); each batch's verdicts still equal the oracle."""
    import torch

    eng = BatchEngine(tuner=None)
    try:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        dfa = W.dense_dfa(256, 0.0005, seed=151)
        rng = np.random.default_rng(152)
        batches = []
        for k in range(5):
            lens = rng.integers(0, 600, size=2000)
            o = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            batches.append((rng.integers(0, 256, size=int(o[-1]), dtype=np.uint8), o))
        k = eng.auto_coalesce(batches[0][1])
        assert k == -(-256 * cus // 2000)
        assert eng.auto_coalesce(batches[0][1], max_bytes=int(batches[0][1][-1]) * 3) == 3
        cap_b = sum(int(o[-1]) for _, o in batches)
        got = list(eng.run_stream(dfa, batches, "auto", depth=2, coalesce="auto",
                                  capacity_bytes=cap_b, capacity_count=2000 * 5))
        assert len(got) == len(batches)
        for j, ((d, o), (f, n)) in enumerate(zip(batches, got)):
            wf, wn = _oracle(dfa, d, o)
            np.testing.assert_array_equal(f, wf, err_msg=f"batch {j}")
            np.testing.assert_array_equal(n, wn, err_msg=f"batch {j}")
        assert list(eng.run_stream(dfa, [], coalesce="auto")) == []
        with pytest.raises(ValueError):
            list(eng.run_stream(dfa, batches, coalesce="many"))
    finally:
        eng.close()


def test_graph_replays_and_direct_launches_by_kernel_count():
    """graph=True replays a captured multi-kernel sequence (split: six kernels) with
    hipGraphLaunch and launches a one-kernel technique (gather) directly; with
    graph_min_kernels=1 the one-kernel graph is replayed too. Every run equals the oracle,
    with eager work on another batch between runs."""
    import torch

    a = W.dense_dfa(256, 0.0005, seed=143)
    d, o = W.uniform_batch(32, 1 << 14, seed=144)
    wf, wn = _oracle(a, d, o)
    s = torch.cuda.current_stream()
    for min_kernels in (3, 1):
        eng = BatchEngine(tuner=None, graph_min_kernels=min_kernels)
        try:
            batch = DeviceBatch(d, o, stream=s)
            other = DeviceBatch(*W.uniform_batch(8, 2048, seed=145), stream=s)
            for technique, replayed in (("split:segment=1024", True),
                                        ("gather", min_kernels == 1)):
                before = eng.replays
                for rep in range(3):
                    batch.flags.fill_(-1)
                    batch.lens.fill_(-1)
                    eng.enqueue(a, batch, technique, stream=s, graph=True)
                    f, n = batch.fetch(s)
                    np.testing.assert_array_equal(f, wf, err_msg=f"{technique} run {rep}")
                    np.testing.assert_array_equal(n, wn, err_msg=f"{technique} run {rep}")
                    eng.enqueue(a, other, "lds", stream=s)
                    other.fetch(s)
                assert (eng.replays - before == 3) == replayed, (technique, min_kernels)
        finally:
            eng.close()


@pytest.mark.parametrize("segment", [256, 512, 1024, 8192])
def test_split_segment_sizes(engine, segment):
    """Split at fixed segment lengths: up to 64 checkpoint intervals per segment the fix
    pass finds flagged intervals from one mask word per head, above that by a scan. Rare
    accepting states put the first matches deep inside heads, after their merge points."""
    import torch

    from gpufsm.core.dfa import DFA

    s = torch.cuda.current_stream()
    for n, seed in ((256, 61), (48, 62)):
        rng = np.random.default_rng(seed)
        base = W.dense_dfa(n, 0.0, seed=seed)
        acc_state = (base.start_state + 1) % n
        trans = np.asarray(base.trans, dtype=np.int32).reshape(n, 256).copy()
        trans[trans == acc_state] = (acc_state + 1) % n
        # about one step in 2 segments enters the accepting state
        k = max(2, (n * 256) // (2 * segment))
        trans.reshape(-1)[rng.choice(n * 256, size=k, replace=False)] = acc_state
        accept = np.zeros(n, dtype=bool)
        accept[acc_state] = True
        a = DFA(num_states=n, start_state=base.start_state, accept=accept,
                trans=trans.reshape(-1))
        d, o = W.uniform_batch(40, 4 * segment + 5000, seed=seed + 1)
        wf, wn = _oracle(a, d, o)
        assert wf.any() and wn[wf.astype(bool)].max() > segment
        batch = DeviceBatch(d, o, stream=s)
        engine.enqueue(a, batch, f"split:segment={segment}", stream=s)
        f, m = batch.fetch(s)
        np.testing.assert_array_equal(f, wf, err_msg=f"n={n}")
        np.testing.assert_array_equal(m, wn, err_msg=f"n={n}")


def test_graphs_survive_workspace_growth_and_eviction():
    """A graph keeps pointing at the workspace it was captured with, so a later, larger
    workspace on the same batch must not free it; an evicted automaton's graphs must be
    dropped, and the next graph run recaptures on the re-uploaded automaton."""
    import torch

    # room for the 2 MiB table (4096 states) or the 64 KiB one (256 states), not both
    eng = BatchEngine(tuner=None, budget_bytes=(2 << 20) + (32 << 10))
    try:
        s = torch.cuda.current_stream()
        small = W.dense_dfa(256, 0.0005, seed=81)
        d, o = W.uniform_batch(32, 1 << 15, seed=82)
        wf, wn = _oracle(small, d, o)
        batch = DeviceBatch(d, o, stream=s)
        # 256-byte segments: no transition maps, a small workspace
        eng.enqueue(small, batch, "split:segment=256", stream=s, graph=True)
        f, n = batch.fetch(s)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
        before = batch.workspace.data_ptr()
        eng.enqueue(small, batch, "split:segment=4096", stream=s)  # map scratch: more bytes
        f, n = batch.fetch(s)
        np.testing.assert_array_equal(f, wf)
        assert batch.workspace.data_ptr() != before, "the workspace did not grow"
        for rep in range(3):  # the first graph replays on its own, retained workspace
            batch.flags.fill_(-1)
            eng.enqueue(small, batch, "split:segment=256", stream=s, graph=True)
            f, n = batch.fetch(s)
            np.testing.assert_array_equal(f, wf, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(n, wn, err_msg=f"replay {rep}")
        big = W.dense_dfa(4096, 0.1, seed=83)   # 2 MiB table: evicts the other automaton
        eng.run_packed(big, d, o)
        assert eng.cache.uploads == 2 and not eng._graphs  # its graph went with it
        batch.flags.fill_(-1)
        eng.enqueue(small, batch, "split:segment=256", stream=s, graph=True)  # re-uploaded
        f, n = batch.fetch(s)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
        assert eng.cache.uploads == 3
    finally:
        eng.close()


def test_resident_cache_reuses_and_evicts():
    eng = BatchEngine(tuner=None, budget_bytes=3 << 20)
    try:
        a = W.dense_dfa(4096, 0.1, seed=51)   # 2 MiB table
        b = W.dense_dfa(4096, 0.1, seed=52)
        d, o = W.uniform_batch(8, 512, seed=53)
        eng.run_packed(a, d, o)
        eng.run_packed(a, d, o)
        assert eng.cache.uploads == 1
        eng.run_packed(b, d, o)               # over budget: a is evicted
        assert eng.cache.uploads == 2 and eng.cache.bytes <= 3 << 20
        f, n, _ = eng.run_packed(a, d, o)
        assert eng.cache.uploads == 3
        wf, wn = _oracle(a, d, o)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
    finally:
        eng.close()


def test_workspace_too_small_is_an_error(engine):
    import torch

    pu = W.pattern_union(50, seed=61)
    d, o = W.pattern_inputs(pu, np.full(8, 2000, dtype=np.int64), 0.7, seed=62)
    batch = DeviceBatch(d, o)
    dev = engine.cache.get(pu.nfa)
    need = dev.workspace_size(batch.count, batch.total, _lib.NFA_CHAIN)
    assert need > 0
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.HIPError, match="workspace"):
        dev.run_device_ws(batch.data.data_ptr(), batch.offsets.data_ptr(), batch.count,
                          batch.total, batch.flags.data_ptr(), batch.lens.data_ptr(),
                          ws.data_ptr(), need - 256, 0, _lib.NFA_CHAIN)


def test_serving_mix_equals_alone():
    """Mixed automata, Poisson arrivals, 3 slots: each result equals its lone run."""
    automata = {
        "dfa256": W.dense_dfa(256, 0.0005, seed=71),
        "dfa4096": W.dense_dfa(4096, 0.0002, seed=72),
        "pu1k": W.pattern_union(40, seed=73).nfa,
        "pu16k": W.pattern_union(700, seed=74).nfa,
    }
    rng = np.random.default_rng(75)
    names = sorted(automata)
    reqs = []
    for k in range(150):
        name = names[rng.integers(0, 4)]
        length = int(min(np.exp(rng.normal(np.log(16384), 1.0)), 1 << 20))
        reqs.append((name, rng.integers(0, 256, size=length, dtype=np.uint8).tobytes()))
    loop = ServingLoop(automata, budget_bytes=1 << 30, slots=3, max_batch_bytes=4 << 20,
                       max_batch_count=64)
    with loop:
        done = poisson_drive(loop, reqs, qps=3000, seed=76)
    assert all(r.status == "ok" for r in done)
    alone = BatchEngine(tuner=None)
    try:
        for (name, data), r in zip(reqs, done):
            d = np.frombuffer(data, dtype=np.uint8)
            o = np.array([0, len(data)], dtype=np.int64)
            f, n, _ = alone.run_packed(automata[name], d, o)
            wf, wn = _oracle(automata[name], d, o)
            assert (r.accepted, r.match_len) == (bool(f[0]), int(n[0])) == (bool(wf[0]), int(wn[0]))
    finally:
        alone.close()
    assert loop.launches < len(reqs)


def test_graph_cache_is_keyed_per_batch_not_by_id():
    """Batches of one shape built and freed in a loop (CPython may give a new batch the id of
    a collected one): each must capture its own graph (replaying another batch's graph would
    read freed buffers) and match the oracle, and a collected batch's graphs must leave the
    engine."""
    import gc

    import torch

    eng = BatchEngine(tuner=None)
    try:
        s = torch.cuda.current_stream()
        dfa = W.dense_dfa(256, 0.01, seed=91)
        for it in range(6):
            d, o = W.uniform_batch(48, 3000, seed=100 + it)   # same shape, new bytes
            wf, wn = _oracle(dfa, d, o)
            batch = DeviceBatch(d, o, stream=s)
            before = eng.captures
            for rep in range(2):
                batch.flags.fill_(-1)
                eng.enqueue(dfa, batch, "split:segment=256", stream=s, graph=True)
                f, n = batch.fetch(s)
                np.testing.assert_array_equal(f, wf, err_msg=f"iteration {it} replay {rep}")
                np.testing.assert_array_equal(n, wn, err_msg=f"iteration {it} replay {rep}")
            assert eng.captures == before + 1, "a new batch replayed an older batch's graph"
            del batch
            gc.collect()
            assert not eng._graphs, "a collected batch left its graph in the engine"
    finally:
        eng.close()


def test_tuner_picks_the_per_input_walk_for_many_short_inputs():
    """262,144 x 4 KiB on the 256-state table: the per-input walk runs 1024 lanes per CU
    on the whole batch and beats the segmented walk, so the tuner, timing candidates on the
    whole batch, must choose it; the engine's verdicts match the oracle."""
    import torch

    eng = BatchEngine(tuner=Autotuner(persist=False))
    try:
        s = torch.cuda.current_stream()
        dfa = W.dense_dfa(256, 0.0, seed=92)
        # one accepting state entered from 4 of the 65,536 (state, byte) pairs: most inputs
        # walk every byte, a few hundred latch
        trans = np.asarray(dfa.trans).reshape(256, 256)
        trans[trans == 11] = 12
        trans[[3, 70, 140, 210], [1, 2, 3, 4]] = 11
        dfa.trans[:] = trans.reshape(-1)
        dfa.accept[11] = True
        d, o = W.uniform_batch(262144, 4096, seed=93)
        batch = DeviceBatch(d, o, stream=s)
        choice = eng.enqueue(dfa, batch, "auto", stream=s)
        f, n = batch.fetch(s)
        wf, wn = _oracle(dfa, d, o)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
        assert choice.technique == "lds", choice
    finally:
        eng.close()


def test_serving_configured_mix_equals_alone_and_oracle():
    """The configured serving mix: 4 random DFAs of 256 to 16,384 states and 4 pattern
    unions of 1,024 to 65,536 states resident together, Poisson arrivals, log-normal request
    lengths (median 16 KiB) plus requests of the 4 MiB cap on the largest automata. Every
    served verdict equals the request's lone run and the oracle."""
    import torch  # noqa: F401  (the loop's device buffers)

    automata, unions = W.serving_automata(seed=0)
    reqs = W.serving_requests(automata, unions, 240, seed=3)
    rng = np.random.default_rng(4)
    for name in ("dfa16384", "nfa65536", "dfa256", "nfa1024"):
        n = 4 << 20
        if name in unions:
            d, _ = W.pattern_inputs(unions[name], np.array([n], dtype=np.int64), 0.70, seed=5)
        else:
            d = rng.integers(0, 256, size=n, dtype=np.uint8)
        reqs.insert(int(rng.integers(0, len(reqs))), (name, d.tobytes()))
    assert max(len(x) for _, x in reqs) == 4 << 20
    loop = ServingLoop(automata, budget_bytes=32 << 30, slots=8)
    with loop:
        done = poisson_drive(loop, reqs, qps=4000, seed=6)
    assert all(r.status == "ok" for r in done)
    assert loop.launches < len(reqs)  # requests shared launches
    # the oracle per automaton over all its requests at once; lone runs one by one
    by_name: dict = {}
    for k, (name, data) in enumerate(reqs):
        by_name.setdefault(name, []).append(k)
    alone = BatchEngine(tuner=None)
    try:
        for name, ks in by_name.items():
            data, off = pack_bytes([reqs[k][1] for k in ks])
            wf, wn = _oracle(automata[name], data, off)
            for j, k in enumerate(ks):
                r = done[k]
                d = np.frombuffer(reqs[k][1], dtype=np.uint8)
                f, n, _ = alone.run_packed(automata[name], d, np.array([0, d.size], np.int64))
                assert (r.accepted, r.match_len) == (bool(f[0]), int(n[0])), (name, k)
                assert (r.accepted, r.match_len) == (bool(wf[j]), int(wn[j])), (name, k)
    finally:
        alone.close()


@pytest.mark.parametrize("depth,graph", [(1, False), (2, False), (3, False), (3, True)])
def test_run_stream_matches_oracle_in_order(depth, graph):
    """A stream of host batches through several slots (their own streams, buffers and
    graphs): every batch's verdicts come back in order and match the oracle, also when a
    later batch is smaller than the first."""
    eng = BatchEngine(tuner=None)
    try:
        pu = W.pattern_union(300, (4, 20), (1, 24), seed=61)
        batches = []
        for k in range(7):
            count = 400 if k != 5 else 90
            lens = W.lognormal_lengths(count, 2000, 1.0, 0, 40000, seed=62 + k)
            batches.append(W.pattern_inputs(pu, lens, 0.7, seed=70 + k))
        cap_b = max(int(o[-1]) for _, o in batches)
        got = list(eng.run_stream(pu.nfa, batches, "auto", depth=depth, graph=graph,
                                  capacity_bytes=cap_b, capacity_count=400))
        assert len(got) == len(batches)
        for k, ((d, o), (f, n)) in enumerate(zip(batches, got)):
            wf, wn = _oracle(pu.nfa, d, o)
            np.testing.assert_array_equal(f, wf, err_msg=f"batch {k}")
            np.testing.assert_array_equal(n, wn, err_msg=f"batch {k}")
    finally:
        eng.close()


@pytest.mark.parametrize("coalesce,depth", [(2, 3), (3, 2)])
def test_run_stream_coalesced_matches_oracle(coalesce, depth):
    """Several host batches per launch (bytes back to back, offsets shifted): each batch's
    verdicts come back as its own slice, in order, equal to the oracle, for a DFA and an
    NFA, with a short last group, an empty batch and a batch of empty inputs."""
    eng = BatchEngine(tuner=None)
    try:
        rng = np.random.default_rng(80 + coalesce)
        dfa = W.dense_dfa(4096, 0.3, seed=81)
        pu = W.pattern_union(200, (4, 16), (1, 20), seed=82)
        for a in (dfa, pu.nfa):
            batches = []
            for k in range(8):
                count = (0, 1, 250, 37, 250, 250, 5, 120)[k]
                lens = W.lognormal_lengths(count, 3000, 1.0, 0, 30000, seed=83 + k)
                if k == 6:
                    lens = np.zeros(count, np.int64)
                if a is pu.nfa:
                    d, o = W.pattern_inputs(pu, lens, 0.7, seed=90 + k)
                else:
                    o = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
                    d = rng.integers(0, 256, size=int(o[-1]) + 7, dtype=np.uint8)  # slack
                batches.append((d, o))
            cap_b = sum(sorted(int(o[-1]) for _, o in batches)[-coalesce:])
            got = list(eng.run_stream(a, batches, "auto", depth=depth, coalesce=coalesce,
                                      capacity_bytes=cap_b, capacity_count=250 * coalesce))
            assert len(got) == len(batches)
            for k, ((d, o), (f, n)) in enumerate(zip(batches, got)):
                wf, wn = _oracle(a, d, o)
                assert f.shape == wf.shape, k
                np.testing.assert_array_equal(f, wf, err_msg=f"batch {k}")
                np.testing.assert_array_equal(n, wn, err_msg=f"batch {k}")
        with pytest.raises(ValueError):
            list(eng.run_stream(dfa, [], coalesce=0))
    finally:
        eng.close()


def test_batch_staged_on_another_stream_waits_for_its_buffers():
    """A DeviceBatch staged on a slot stream allocates and zero-fills its buffers on the
    current stream; the slot stream must wait for that stream before copying into them (a
    late fill would zero the staged offsets). With the current stream held by a bounded GPU
    sleep, the slot stream's first copy is still waiting a few milliseconds later, and the
    verdicts equal the oracle's."""
    import time

    import torch

    from gpufsm.batch import slot_streams

    pu = W.pattern_union(200, (4, 16), (1, 24), seed=81)
    lens = W.lognormal_lengths(300, 1500, 1.0, 0, 20000, seed=82)
    d, o = W.pattern_inputs(pu, lens, 0.7, seed=83)
    eng = BatchEngine(tuner=None)
    try:
        st = slot_streams(1)[0]
        torch.cuda.synchronize()
        torch.cuda._sleep(20_000_000)  # the current stream busy for milliseconds (bounded)
        batch = DeviceBatch(d, o, stream=st)
        staged = torch.cuda.Event()
        staged.record(st)
        time.sleep(0.002)
        waiting = not staged.query()
        eng.enqueue(pu.nfa, batch, "auto", stream=st)
        f, n = batch.fetch(st)
        torch.cuda.synchronize()
        assert waiting, "the slot stream did not wait for the current stream"
        wf, wn = _oracle(pu.nfa, d, o)
        np.testing.assert_array_equal(f, wf)
        np.testing.assert_array_equal(n, wn)
    finally:
        eng.close()
