#!/usr/bin/env python3
"""Benchmark of the gpufsm HIP hot path on one MI355X; prints ONE JSON line.

Headline: the DFA memory-bound face — a seeded random complete DFA (``--states``, default
256: the uint16-table variant that fits LDS; 4096 is reported beside it), a batch of 4096
inputs of 64 KiB uniform random bytes already resident in HBM. The ``value`` is measured on
the non-latching control (``accept_prob = 0``, as the upstream latch control does), where
every input byte is examined, so input GB/s counts bytes actually read. The latching
variant as configured (10% accepting states) is reported beside it with both examined and
credited bytes, and so is the plain per-call drop-in (``gpufsm.run_batch``: upload, copy,
launch, copy back).

Extras (each gated against the C oracle before it is timed): the ragged NFA mix, plain
per-call ``run_batch`` versus the resident, autotuned, graph-captured engine; the plain NFA
drop-in in microseconds per call; the serving loop's p50/p99 latency and throughput over a
QPS sweep.

Order of work:
  1. build the workload, upload it, run once, and gate every verdict against the C oracle
     (this run also times the CPU baseline on the host cores: the full batch);
  2. ``--warmup`` untimed steps, then ``--steps`` timed steps bracketed by synchronisation,
     each step one launch over the whole batch; per-launch kernel time from HIP events
     recorded on the launch stream;
  3. report throughput, the roofline of the dominant kernel, and the CPU baseline.

``--gpus`` above 1 exits with a message: this benchmark runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "abstraction-regret_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _traffic(technique: str, workload: str):
    """HBM bytes per step from the committed rocprofv3 counter summary, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table.get(f"{technique}:{workload}")
    except (OSError, ValueError):
        return None


class DeviceBatch:
    """A batch resident in HBM plus output buffers (torch tensors as plumbing)."""

    def __init__(self, torch, data: np.ndarray, offsets: np.ndarray) -> None:
        self.count = int(offsets.size) - 1
        self.total = int(offsets[-1])
        self.data = torch.from_numpy(data if data.size else np.zeros(1, np.uint8)).to("cuda")
        self.offsets = torch.from_numpy(offsets).to("cuda")
        self.flags = torch.empty(max(self.count, 1), dtype=torch.int32, device="cuda")
        self.lens = torch.empty(max(self.count, 1), dtype=torch.int32, device="cuda")

    def run(self, dev, technique: int, stream: int) -> None:
        dev.run_device(self.data.data_ptr(), self.offsets.data_ptr(), self.count, self.total,
                       self.flags.data_ptr(), self.lens.data_ptr(), stream, technique)

    def results(self):
        return (self.flags[: self.count].cpu().numpy(), self.lens[: self.count].cpu().numpy())


def _time_steps(torch, fn, steps: int, warmup: int):
    """Warmup, then ``steps`` timed calls: (wall seconds, per-launch kernel ms list)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        fn()
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [s.elapsed_time(e) for s, e in zip(starts, ends)]


# The kernels one step of each technique launches (the roofline covers the whole step).
KERNELS = {
    "gather": "dfa_walk_kernel<gather>",
    "lds": "dfa_walk_kernel<lds>",
    "split": "dfa_split (init, walk, fix, repair below 2 KiB segments, event, map, finish)",
}


def _measure_dfa(torch, batch, dfa, technique: str, data, offsets, steps: int, warmup: int,
                 what: str, depth: int = 1, coalesce: int = 1) -> dict:
    """Gate one DFA/technique on the full batch against the C oracle, then time it.

    Two timings: one batch at a time on one stream (``ms_per_step``), and, with ``depth``
    above 1, a stream of the same batch run back to back on ``depth`` streams, each with
    its own outputs and workspace, so that a launch's tail overlaps the next launch
    (``stream``). With ``coalesce`` above 1 each launch of the stream covers that many
    consecutive batches (copies of the batch back to back in HBM), as a stream engine
    coalesces batches that are queued together. A step of the stream is one launch:
    ``warmup`` untimed launches, then exactly ``steps`` timed ones; ``ms_per_step`` is per
    launch and ``ms_per_batch`` per batch. Every slot's verdicts are gated against the
    oracle after the timed stream.
    """
    from gpufsm.backends.hip import _lib
    from gpufsm.backends.hip.handles import DeviceDFA, make_tuning

    from oracle import gate, native

    tech = _lib.DFA_TECHNIQUES[technique]
    stream = torch.cuda.current_stream().cuda_stream
    count, total = batch.count, batch.total
    length_max = int(np.diff(offsets).max()) if count else 0
    with DeviceDFA(dfa) as dev:
        batch.run(dev, tech, stream)
        torch.cuda.synchronize()
        f, m = batch.results()
        wf, wm, wex = native.dfa_batch(dfa, data, offsets)
        gate.require_arrays(f, m, wf, wm, f"hip dfa {technique} ({what})")
        wall, ks = _time_steps(torch, lambda: batch.run(dev, tech, stream), steps, warmup)
        entry_bytes = dev.props.entry_bytes
        names = {v: k for k, v in _lib.DFA_TECHNIQUES.items()}
        resolved = names[dev.resolve(batch.count, batch.total, tech)]
        piped = None
        if depth > 1 and count:
            tu = make_tuning(max_input_bytes=length_max)
            ccount, ctotal = count * coalesce, total * coalesce
            if coalesce > 1:
                cdata = batch.data[:total].repeat(coalesce)
                coff = torch.from_numpy(np.concatenate(
                    [offsets[:-1] + k * total for k in range(coalesce)] + [[ctotal]])).to("cuda")
            else:
                cdata, coff = batch.data, batch.offsets
            need = dev.workspace_size(ccount, ctotal, tech, tu)
            from gpufsm.batch import slot_streams

            # streams from the process-wide pool: each measurement's slots keep distinct
            # hardware queues (fresh streams would wrap onto queues other slots hold)
            slots = [(st,
                      torch.empty(ccount, dtype=torch.int32, device="cuda"),
                      torch.empty(ccount, dtype=torch.int32, device="cuda"),
                      torch.empty(max(need, 256), dtype=torch.uint8, device="cuda"))
                     for st in slot_streams(depth)]
            cur = torch.cuda.current_stream()

            def run(k: int) -> None:
                s, fl, ln, ws = slots[k % depth]
                dev.run_device_ws(cdata.data_ptr(), coff.data_ptr(), ccount, ctotal,
                                  fl.data_ptr(), ln.data_ptr(), ws.data_ptr(), ws.numel(),
                                  s.cuda_stream, tech, tu)

            def stream_of(k0: int, n: int) -> None:
                for s, *_ in slots:
                    s.wait_stream(cur)
                for k in range(k0, k0 + n):
                    run(k)
                for s, *_ in slots:
                    cur.wait_stream(s)

            # A step of the stream is one launch (``coalesce`` batches): exactly ``warmup``
            # untimed launches, then exactly ``steps`` timed ones, round-robin over the slots.
            launches = steps
            if warmup > 0:
                stream_of(0, warmup)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(cur)
            stream_of(0, launches)
            e1.record(cur)
            torch.cuda.synchronize()
            pwall = time.perf_counter() - t0
            for _, fl, ln, _ in slots[:min(depth, launches)]:  # every slot that ran
                gate.require_arrays(fl.cpu().numpy(), ln.cpu().numpy(), np.tile(wf, coalesce),
                                    np.tile(wm, coalesce),
                                    f"hip dfa {technique} ({what}, stream of {depth})")
            nb = launches * coalesce
            piped = {"depth": depth, "batches_per_launch": coalesce,
                     "technique": names[dev.resolve(ccount, ctotal, tech, tu)],
                     "launch_inputs": ccount, "launch_bytes": ctotal,
                     "steps": launches, "warmup": warmup,
                     "ms_per_step": 1e3 * pwall / launches,  # a step: one launch
                     "event_ms_per_step": e0.elapsed_time(e1) / launches,
                     "ms_per_batch": 1e3 * pwall / nb,
                     "event_ms_per_batch": e0.elapsed_time(e1) / nb,
                     # one launch's duration, launches overlapping
                     "ms_per_launch": e0.elapsed_time(e1) / launches * min(depth, launches)}
            del cdata, coff, slots
    examined = int(wex.sum())
    credited = int(offsets[-1])
    step_s = wall / steps
    out = {
        "workload": what,
        "technique": technique,
        "resolved_technique": resolved,
        "accepted": int(f.sum()),
        "examined_bytes": examined,
        "credited_bytes": credited,
        "ms_per_step": 1e3 * step_s,
        "kernel_ms": float(np.mean(ks)),
        "examined_GBps": examined / step_s / 1e9,
        "credited_GBps": credited / step_s / 1e9,
        "table_bytes": dfa.num_states * 256 * entry_bytes,
    }
    msg = f"{what} [{technique}]: {out['examined_GBps']:.1f} GB/s examined, " \
          f"{out['ms_per_step']:.3f} ms/step"
    if piped:
        piped["examined_GBps"] = examined / (piped["ms_per_batch"] * 1e-3) / 1e9
        piped["credited_GBps"] = credited / (piped["ms_per_batch"] * 1e-3) / 1e9
        out["stream"] = piped
        per = f", {coalesce} batches per launch" if coalesce > 1 else ""
        msg += f"; stream of {depth}{per}: {piped['examined_GBps']:.1f} GB/s, " \
               f"{piped['ms_per_batch']:.3f} ms/batch"
    _log(msg)
    return out


def bench_dfa(args, torch) -> dict:
    from gpufsm import run_batch, workloads as W

    from oracle import native

    n, count, length = args.states, args.count, args.length
    _log(f"workload dfa n={n} count={count} length={length} technique={args.technique}")
    data, offsets = W.uniform_batch(count, length, seed=0)
    batch = DeviceBatch(torch, data, offsets)
    extras: dict = {"variants": []}

    # headline: non-latching control of the chosen table size
    dfa = W.dense_dfa(n, 0.0, seed=0)
    head = _measure_dfa(torch, batch, dfa, args.technique, data, offsets, args.steps,
                        args.warmup, f"dfa{n} non-latching", depth=args.pipeline,
                        coalesce=args.coalesce)
    if args.profile:
        # a plain device copy of the same input bytes: the reference point that scales the
        # FETCH_SIZE / WRITE_SIZE counters of a profiled run into HBM bytes
        copy = batch.data.clone()
        torch.cuda.synchronize()
        del copy
    # the other variants of the same workload (not the headline)
    for other_n in (() if args.profile else (256, 4096)):
        for latching in (False, True):
            if other_n == n and not latching:
                continue
            techs = ["auto"] if latching else ["auto", "gather", "split"]
            for t in techs:
                d = W.dense_dfa(other_n, 0.10 if latching else 0.0, seed=0)
                extras["variants"].append(_measure_dfa(
                    torch, batch, d, t, data, offsets, max(2, args.steps // 2), 1,
                    f"dfa{other_n} {'latching 10%' if latching else 'non-latching'}",
                    # the same stream as the headline: 2 launches of 64 batches in flight
                    # (for the 4096-state table auto then takes the per-input gather walk)
                    depth=2 if (t == "auto" and not latching) else 1,
                    coalesce=64 if (t == "auto" and not latching) else 1))
    _log("oracle gate passed on every variant (full batch)")

    # CPU baseline: the C oracle over all host cores, full batch, headline automaton
    threads = _threads()
    t0 = time.perf_counter()
    reps = 0
    while True:
        native.dfa_batch(dfa, data, offsets, threads=threads)
        reps += 1
        if time.perf_counter() - t0 > 2.0 or reps >= 8:
            break
    cpu_s = (time.perf_counter() - t0) / reps
    # the reference CPU backend's own algorithm (Python, one core), on whole inputs
    from oracle.reference import simulate_dfa

    extras["cpu_python_reference"] = _python_cpu_baseline(
        "reference", lambda x: simulate_dfa(dfa, x),
        (data[offsets[i]:offsets[i + 1]].tobytes() for i in range(count)), 2.0,
        f"whole {length}-byte inputs of the headline batch, in order, about 2 s")
    _log(f"cpu python reference: {extras['cpu_python_reference']['value'] * 1e3:.2f} MB/s "
         f"on {extras['cpu_python_reference']['inputs']} inputs")

    # plain per-call drop-in (gpufsm.run_batch builds an executor, copies, launches, copies back)
    if not args.profile:
        inputs = [data[offsets[i]:offsets[i + 1]].tobytes() for i in range(count)]
        run_batch(dfa, inputs[:4], backend="hip", technique=args.technique)
        t0 = time.perf_counter()
        res = run_batch(dfa, inputs, backend="hip", technique=args.technique)
        plain_s = time.perf_counter() - t0
        assert not any(r.accepted for r in res)
        extras["plain_per_call"] = {"seconds": plain_s,
                                    "examined_GBps": head["examined_bytes"] / plain_s / 1e9,
                                    "kernel_ms": res[0].kernel_ms}

    examined = head["examined_bytes"]
    compulsory = head["table_bytes"] + 8 * (count + 1) + 8 * count
    algorithmic = examined + compulsory
    tech = head["resolved_technique"]  # what `auto` picked for this shape
    kernel_name = KERNELS.get(tech, tech)
    workload_name = f"dfa{n}_{count}x{length}_nonlatching"
    piped = head.get("stream")
    single = {"examined_GBps": head["examined_GBps"], "ms_per_step": head["ms_per_step"],
              "kernel_ms": head["kernel_ms"],
              "roofline_frac": algorithmic / (head["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS}
    per_step = 1  # configured batches per step
    if piped:
        # a stream of batches: a step is one launch of the stream (its batches coalesced),
        # timed in steady state with the next launches overlapping
        value, step_ms, kernel_ms = piped["examined_GBps"], piped["ms_per_step"], \
            piped["event_ms_per_step"]
        per_step = piped["batches_per_launch"]
        tech = piped["technique"]  # what `auto` picked for the launch shape
        kernel_name = KERNELS.get(tech, tech)
        per = (f", {piped['batches_per_launch']} batches per launch"
               if piped["batches_per_launch"] > 1 else "")
        kernel_name = (f"{kernel_name}, {piped['depth']} launches in flight on "
                       f"{piped['depth']} streams{per} (time per step, one launch, in steady "
                       f"state)")
        how = ("one lane walks each input of the launch's batches (the per-input walk: "
               f"{piped['launch_inputs']} inputs per launch)" if tech in ("lds", "gather")
               else "a batch's repair passes overlap the next batch's walk")
        mode = (f"batches run back to back, {piped['depth']} launches in flight on their own "
                f"streams and workspaces{per} ({how}; {piped['ms_per_launch']:.2f} ms per "
                f"launch); a step is one launch ({per_step} batch"
                f"{'es' if per_step > 1 else ''}, {piped['ms_per_batch']:.4f} ms per batch); "
                f"every slot's verdicts gated; one batch alone is under single_batch")
    else:
        value, step_ms, kernel_ms = head["examined_GBps"], head["ms_per_step"], \
            head["kernel_ms"]
        mode = "one batch at a time"
    achieved = algorithmic * per_step / (kernel_ms * 1e-3) / 1e9
    return {
        "metric": "input GB/s batched vs plain run_batch on MI355X; p50/p99 latency at fixed QPS",
        "value": value,
        "unit": "GB/s",
        "ms_per_step": step_ms,
        "dtype": "u8",
        "config": {
            "workload": f"DFA memory-bound face (batched: inputs resident in HBM; value = "
                        f"examined input GB/s), non-latching control: random complete DFA "
                        f"{n} states x 256 (byte table, accept_prob 0), {count} inputs x "
                        f"{length} uniform bytes (seed 0); {mode}; plain per-call, the "
                        f"latching variant, the 4096-state table, the ragged NFA mix, the "
                        f"plain NFA drop-in and the serving sweep are under extras",
            "num_states": n, "num_inputs": count, "input_bytes": length,
            "technique": args.technique, "resolved_technique": tech,
            "batches_in_flight": piped["depth"] if piped else 1,
            "batches_per_launch": piped["batches_per_launch"] if piped else 1,
            "batches_per_step": per_step,
            "global_batch": count * per_step,  # inputs per step
            "ms_per_batch": piped["ms_per_batch"] if piped else head["ms_per_step"],
        },
        "single_batch": single,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            # HBM bytes per step: the profile's per-batch figure times the step's batches
            "traffic": (lambda t: t * per_step if t is not None else None)(
                _traffic(tech, workload_name)),
            "kernel": kernel_name,
            "kernel_ms": kernel_ms,  # per step, launches overlapping
            "algorithmic_bytes_per_step": algorithmic * per_step,
            "algorithmic_bytes_per_batch": algorithmic,
        },
        "cpu_baseline": {
            "value": examined / cpu_s / 1e9,
            "unit": "GB/s",
            "cores": threads,
            "kind": "port",
            "sample": f"full batch ({count} x {length} B, non-latching, {n} states), C oracle "
                      f"over {threads} threads, mean of {reps} runs",
        },
        "extras": extras,
    }


def _python_cpu_baseline(technique: str, fn, inputs, budget_s: float, what: str) -> dict:
    """The reference CPU backend's algorithm, restated in Python, on one core.

    `src/gpufsm/backends/cpu.py:28-67` times each input with ``time.perf_counter``; so does
    this, over a bounded sample (inputs in order until ``budget_s`` has passed, at least 4),
    and reports examined bytes per second with a bootstrap CI95 over the sampled inputs
    (as `bench/timing.py:21-76` reports its CIs). ``fn(bytes) -> (accepted, match_len)``.
    Examined bytes are match_len for an accepted input, else the whole input (none of the
    timed automata can die: the DFAs are complete, the pattern unions keep their start).
    """
    times, nbytes = [], []
    t_end = time.perf_counter() + budget_s
    for x in inputs:
        t0 = time.perf_counter()
        accepted, match_len = fn(x)
        times.append(time.perf_counter() - t0)
        nbytes.append(match_len if accepted else len(x))
        if len(times) >= 4 and time.perf_counter() > t_end:
            break
    t, b = np.asarray(times), np.asarray(nbytes, dtype=np.float64)
    rng = np.random.default_rng(0)
    boot = []
    for _ in range(1000):
        idx = rng.integers(0, t.size, t.size)
        boot.append(b[idx].sum() / t[idx].sum() / 1e9)
    lo, hi = np.percentile(boot, [2.5, 97.5])
    return {"value": float(b.sum() / t.sum() / 1e9), "unit": "GB/s", "cores": 1,
            "kind": "reference algorithm restated in Python", "technique": technique,
            "inputs": int(t.size), "examined_bytes": int(b.sum()),
            "median_ms_per_input": float(np.median(t) * 1e3),
            "ci95": [float(lo), float(hi)], "sample": what}


def _examined_upper(flags: np.ndarray, lens: np.ndarray, offsets: np.ndarray) -> int:
    """Bytes up to each first match (whole input when nothing matched)."""
    return int(np.where(flags != 0, lens, np.diff(offsets)).sum())


def _serve_bulk(torch, pu, data, off, inputs, wf, wn, technique: str) -> dict:
    """The ragged mix served: every input a request to the native serving loop (the
    library's scheduler batches an automaton's admitted requests into launches of up to
    64 MiB / 16384 requests, stages them through pinned memory in chunks whose copies
    overlap the staging, and launches on its slot streams), all submitted at once; from the
    first submit to the last verdict, every verdict gated.
    Twice: from the caller's ordinary (pageable) buffer, and from a caller buffer in pinned
    memory, which the scheduler copies to the device without staging. Then one submit call
    per request. Per buffer one untimed bulk submit (it touches the staging pages), then 5
    timed ones: the median is reported (as for the plain per-call runs), the best beside it."""
    from gpufsm.serve import NativeServingLoop

    from oracle import gate

    count = len(inputs)
    served, served_pinned = [], []
    pinned_t = torch.empty(int(off[-1]), dtype=torch.uint8, pin_memory=True)
    pinned_buf = pinned_t.numpy()
    pinned_buf[:] = data[: int(off[-1])]
    loop = NativeServingLoop({"ragged": pu.nfa}, slots=4, max_batch_bytes=64 << 20,
                             max_batch_count=16384, max_inflight_per_automaton=4,
                             technique=technique)
    with loop:
        loop.collect([loop.submit("ragged", x) for x in inputs[:256]])  # builds the lists
        for buf, times in ((data, served), (pinned_buf, served_pinned)):
            for rep in range(6):
                t0 = time.perf_counter()
                bulk = loop.submit_many("ragged", buf, off, copy=False)  # one call, no copy
                bulk.wait()
                if rep > 0:
                    times.append(time.perf_counter() - t0)
                gate.require_arrays(bulk.accepted.astype(np.int32), bulk.match_len, wf, wn,
                                    "hip nfa served in bulk (ragged)")
        one_by_one = time.perf_counter()
        single = [loop.submit("ragged", x) for x in inputs]
        loop.collect(single)
        one_by_one = time.perf_counter() - one_by_one
        launches = loop.query()["launches"]
    gate.require_arrays(np.fromiter((r.accepted for r in single), np.int32, count),
                        np.fromiter((r.match_len for r in single), np.int32, count), wf, wn,
                        "hip nfa served one request per call (ragged)")
    del pinned_t
    return {"served_s": float(np.median(served)),
            "served_pinned_s": float(np.median(served_pinned)),
            "served_best_s": min(served), "served_pinned_best_s": min(served_pinned),
            "one_by_one_s": one_by_one, "launches": int(launches)}


def bench_ragged(args, torch) -> dict:
    """Ragged NFA mix: plain per-call ``gpufsm.run_batch`` vs the resident engine.

    Plain: a fresh executor per call (automaton preprocessing and upload, packing of the
    16384 ``bytes`` objects, host-to-device copy, launch, copy back, Result objects), as
    upstream's per-call path. Batched: the automaton resident, the packed batch resident in
    HBM, technique and tuning from the autotuner, the launch sequence captured into a HIP
    graph and replayed. Both are gated against the C oracle (which also gives the examined
    byte count); throughput is examined bytes per second of wall time per batch.
    """
    from gpufsm import run_batch, workloads as W
    from gpufsm.batch import BatchEngine, DeviceBatch
    from gpufsm.tune import Autotuner

    from oracle import gate, native

    count = args.ragged_count
    t0 = time.perf_counter()
    pu, (data, off) = W.workload_ragged(count=count)
    _log(f"ragged nfa: {pu.nfa.num_states} states, {count} inputs, {off[-1] / 1e6:.1f} MB "
         f"({time.perf_counter() - t0:.1f} s to generate)")
    threads = _threads()
    t0 = time.perf_counter()
    wf, wn, wex = native.nfa_batch(pu.nfa, data, off, threads=threads)
    cpu_s = time.perf_counter() - t0
    examined, credited = int(wex.sum()), int(off[-1])
    _log(f"ragged nfa: oracle {cpu_s:.1f} s on {threads} threads, {int(wf.sum())} accepted, "
         f"{examined / 1e6:.1f} MB examined")

    # the reference CPU backend's faster technique (bitmap, Python, one core) on prefixes:
    # the union's start state has thousands of edges, about 1 ms per byte in Python
    py_cpu = None
    if not args.profile:
        from oracle.bitmap import BitmapProgram

        t0 = time.perf_counter()
        prog = BitmapProgram(pu.nfa)
        compile_s = time.perf_counter() - t0
        py_cpu = _python_cpu_baseline(
            "bitmap", prog.run,
            (data[off[i]:min(off[i + 1], off[i] + 512)].tobytes() for i in range(count)), 3.0,
            "first 512 bytes of the batch's inputs, in order, about 3 s; the per-automaton "
            f"compile ({compile_s:.1f} s) is not counted")
        _log(f"ragged nfa: cpu python bitmap {py_cpu['value'] * 1e6:.1f} KB/s on "
             f"{py_cpu['inputs']} prefixes")

    # batched: resident automaton + batch, autotuned, graph-captured
    s = torch.cuda.current_stream()
    engine = BatchEngine(tuner=Autotuner(persist=False))
    t0 = time.perf_counter()
    batch = DeviceBatch(data, off, stream=s)
    tech = args.ragged_technique
    choice = engine.enqueue(pu.nfa, batch, tech, stream=s)
    f, n = batch.fetch(s)
    tune_s = time.perf_counter() - t0
    gate.require_arrays(f, n, wf, wn, f"hip nfa engine {choice.label()} (ragged)")
    steps = args.steps if args.workload == "ragged" else max(3, args.steps // 2)
    wall, ks = _time_steps(torch, lambda: engine.enqueue(pu.nfa, batch, tech, stream=s,
                                                         graph=True), steps,
                           args.warmup if args.workload == "ragged" else 1)
    f, n = batch.fetch(s)
    gate.require_arrays(f, n, wf, wn, "hip nfa engine graph replay (ragged)")
    batched_s = wall / steps
    dev = engine.cache.get(pu.nfa)
    resident_bytes = int(dev.props.device_bytes)
    if args.profile:
        # one device copy of the input bytes: the calibration point of FETCH/WRITE_SIZE
        copy = batch.data.clone()
        torch.cuda.synchronize()
        del copy
        engine.close()
        return {"examined_bytes": examined, "credited_bytes": credited, "count": count,
                "batched_s": batched_s, "kernel_ms": float(np.mean(ks)),
                "technique": choice.label(), "resident_bytes": resident_bytes,
                "cpu_s": cpu_s, "threads": threads, "states": pu.nfa.num_states}
    # the same with the host batch staged every time (pinned copy + H2D + launch + D2H)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        batch.load(data, off, stream=s)
        engine.enqueue(pu.nfa, batch, tech, stream=s, graph=True)
        batch.fetch(s)
    staged_s = (time.perf_counter() - t0) / reps
    # the same host-resident batches as a pipelined stream: while the GPU walks batch k, the
    # host stages batch k + 1 into another slot (pinned copy on 8 threads, H2D on that slot's
    # stream), so staging overlaps the kernels. Timed in steady state: from the result of
    # batch `depth` to the last one; every result is gated.
    depth, n_stream = 3, max(8, steps) + 3
    marks = []
    for f_, n_ in engine.run_stream(pu.nfa, [(data, off)] * n_stream, tech, depth=depth):
        gate.require_arrays(f_, n_, wf, wn, "hip nfa host-staged stream (ragged)")
        marks.append(time.perf_counter())
    pipelined_s = (marks[-1] - marks[depth - 1]) / (n_stream - depth)
    key = next(iter(engine.tuner.table), None)
    decision = engine.tuner.table.get(key, {"rates": {}}) if key else {"rates": {}}

    # plain per-call drop-in
    inputs = [data[off[i]:off[i + 1]].tobytes() for i in range(count)]
    res = run_batch(pu.nfa, inputs, backend="hip")
    got_f = np.fromiter((r.accepted for r in res), dtype=np.int32, count=count)
    got_n = np.fromiter((r.match_len for r in res), dtype=np.int32, count=count)
    gate.require_arrays(got_f, got_n, wf, wn, "hip nfa plain run_batch (ragged)")
    plain = []
    for _ in range(3):
        t0 = time.perf_counter()
        res = run_batch(pu.nfa, inputs, backend="hip")
        plain.append(time.perf_counter() - t0)
    plain_s = float(np.median(plain))
    engine.close()
    sv = _serve_bulk(torch, pu, data, off, inputs, wf, wn, choice.label())
    served_s, served_pinned_s = sv["served_s"], sv["served_pinned_s"]
    one_by_one, launches = sv["one_by_one_s"], sv["launches"]
    out = {
        "workload": f"ragged NFA mix: {pu.nfa.num_states}-state union of 2000 patterns "
                    f"(lengths 5-40, class widths 1-32, wildcard start), {count} inputs, "
                    f"log-normal lengths (median 4 KiB, sigma 1, 64 B-1 MiB), 70% pattern bytes",
        "examined_bytes": examined,
        "credited_bytes": credited,
        "accepted": int(wf.sum()),
        "batched": {
            "technique": choice.label(),
            "ms_per_batch": 1e3 * batched_s,
            "kernel_ms": float(np.mean(ks)),
            "examined_GBps": examined / batched_s / 1e9,
            "credited_GBps": credited / batched_s / 1e9,
            "tuning_seconds": tune_s,
            "autotuner": {"key": key, "rates_GBps_sample": decision["rates"]},
        },
        "batched_host_staged": {"ms_per_batch": 1e3 * staged_s,
                                "examined_GBps": examined / staged_s / 1e9,
                                "note": "one batch at a time: pinned copy, H2D, graph "
                                        "replay, D2H, each waited for"},
        "batched_host_pipelined": {"ms_per_batch": 1e3 * pipelined_s,
                                   "examined_GBps": examined / pipelined_s / 1e9,
                                   "credited_GBps": credited / pipelined_s / 1e9,
                                   "slots": depth, "batches": n_stream,
                                   "note": "host-resident batches streamed through "
                                           "BatchEngine.run_stream: staging of the next "
                                           "batch overlaps the current batch's kernels"},
        "plain_per_call": {"ms_per_batch": 1e3 * plain_s,
                           "examined_GBps": examined / plain_s / 1e9,
                           "credited_GBps": credited / plain_s / 1e9,
                           "kernel_ms": res[0].kernel_ms},
        "served": {"ms_per_batch": 1e3 * served_s,
                   "best_ms_per_batch": 1e3 * sv["served_best_s"],
                   "examined_GBps": examined / served_s / 1e9,
                   "credited_GBps": credited / served_s / 1e9,
                   "one_request_per_call_ms": 1e3 * one_by_one,
                   "one_request_per_call_GBps": examined / one_by_one / 1e9,
                   "launches": int(launches),
                   "note": f"the {count} inputs submitted as requests to the native serving "
                           f"loop in one bulk call, read in place (4 slots, launches of up to "
                           f"64 MiB / 16384 requests, technique {choice.label()}); first "
                           f"submit to last verdict, host copies included, median of 5 after "
                           f"one untimed submit; also one submit call per request"},
        "served_pinned": {"ms_per_batch": 1e3 * served_pinned_s,
                          "best_ms_per_batch": 1e3 * sv["served_pinned_best_s"],
                          "examined_GBps": examined / served_pinned_s / 1e9,
                          "credited_GBps": credited / served_pinned_s / 1e9,
                          "note": "the same bulk submit from a caller buffer in pinned host "
                                  "memory: the scheduler copies it to the device directly"},
        "batched_over_plain": plain_s / batched_s,
        "host_pipelined_over_plain": plain_s / pipelined_s,
        "served_over_plain": plain_s / served_s,
        "served_pinned_over_plain": plain_s / served_pinned_s,
        "cpu_python_bitmap": py_cpu,
        "resident_bytes": resident_bytes,
        "count": count,
        "states": pu.nfa.num_states,
        "batched_s": batched_s,
        "cpu_s": cpu_s,
        "threads": threads,
        "cpu_oracle": {"examined_GBps": examined / cpu_s / 1e9, "threads": threads,
                       "seconds": cpu_s},
    }
    _log(f"ragged nfa: batched {out['batched']['examined_GBps']:.2f} GB/s "
         f"({choice.label()}), plain {out['plain_per_call']['examined_GBps']:.3f} GB/s, "
         f"ratio {out['batched_over_plain']:.1f}x; host-staged pipelined "
         f"{1e3 * pipelined_s:.2f} ms/batch, {out['host_pipelined_over_plain']:.1f}x plain; "
         f"served {examined / served_s / 1e9:.2f} GB/s, {out['served_over_plain']:.1f}x plain "
         f"(from pinned memory {examined / served_pinned_s / 1e9:.2f} GB/s, "
         f"{out['served_pinned_over_plain']:.1f}x)")
    return out


def bench_plain_nfa(args, torch) -> dict:
    """Plain NFA drop-in (1024 states, 64 x 4 KiB): microseconds per call."""
    from gpufsm import run_batch, workloads as W
    from gpufsm.batch import BatchEngine, DeviceBatch

    from oracle import gate, native

    nfa, (data, off) = W.workload_plain()
    wf, wn, wex = native.nfa_batch(nfa, data, off)
    inputs = [data[off[i]:off[i + 1]].tobytes() for i in range(off.size - 1)]
    res = run_batch(nfa, inputs, backend="hip")
    gate.require_arrays(np.array([r.accepted for r in res], np.int32),
                        np.array([r.match_len for r in res], np.int32), wf, wn,
                        "hip plain nfa run_batch")
    calls = []
    for _ in range(20):
        t0 = time.perf_counter()
        run_batch(nfa, inputs, backend="hip")
        calls.append(time.perf_counter() - t0)
    s = torch.cuda.current_stream()
    engine = BatchEngine(tuner=None)
    batch = DeviceBatch(data, off, stream=s)
    engine.enqueue(nfa, batch, "auto", stream=s)
    f, n = batch.fetch(s)
    gate.require_arrays(f, n, wf, wn, "hip plain nfa engine")
    wall, ks = _time_steps(torch, lambda: engine.enqueue(nfa, batch, "auto", stream=s,
                                                         graph=True), 50, 5)
    engine.close()
    per_call = float(np.median(calls))
    return {
        "workload": "plain NFA drop-in: random NFA 1024 states x 4 edges, 5% accepting, "
                    "64 inputs x 4096 uniform bytes",
        "examined_bytes": int(wex.sum()),
        "credited_bytes": int(off[-1]),
        "plain_per_call_us": 1e6 * per_call,
        "plain_per_call_p90_us": 1e6 * float(np.percentile(calls, 90)),
        # enqueue(graph=True): one kernel here, so launched directly (a graph pays from
        # three kernels on); back to back, so this is the host time per call
        "resident_us": 1e6 * wall / 50,
        "resident_kernel_us": 1e3 * float(np.mean(ks)),
        "plain_credited_GBps": int(off[-1]) / per_call / 1e9,
    }


def bench_serving(args, torch) -> dict:
    """Serving loop: 8 resident automata, Poisson arrivals at a QPS sweep."""
    from gpufsm import workloads as W
    from gpufsm.serve import NativeServingLoop, ServingLoop, poisson_drive

    from oracle import gate, native

    automata, unions = W.serving_automata()
    if args.serve_impl == "python":
        loop = ServingLoop(automata, budget_bytes=32 << 30, slots=args.serve_slots,
                           max_batch_bytes=16 << 20, max_batch_count=1024)
    else:  # the library's native scheduler
        loop = NativeServingLoop(automata, budget_bytes=32 << 30, slots=args.serve_slots,
                                 max_batch_bytes=16 << 20, max_batch_count=1024,
                                 max_inflight_per_automaton=args.serve_inflight)
    sweep = []
    with loop:
        warm = W.serving_requests(automata, unions, 64, seed=99)
        poisson_drive(loop, warm, qps=2000, seed=99)
        for k, qps in enumerate(args.serve_qps):
            n = max(64, int(qps * args.serve_seconds))
            reqs = W.serving_requests(automata, unions, n, seed=1000 + k)
            t_start = time.perf_counter()
            done = poisson_drive(loop, reqs, qps=qps, seed=2000 + k)
            st = loop.stats(since=t_start)
            st["qps_offered"] = qps
            st["seconds"] = time.perf_counter() - t_start
            st["mean_request_bytes"] = float(np.mean([len(d) for _, d in reqs]))
            # per automaton: the large random DFAs are walked serially (a random n-state
            # DFA needs segments of several n bytes to split), so they set the tail
            per = {}
            for name in sorted(automata):
                lat = np.array([r.latency for r in done if r.name == name and r.status == "ok"])
                if lat.size:
                    per[name] = {"requests": int(lat.size),
                                 "p50_ms": float(np.percentile(lat, 50) * 1e3),
                                 "p99_ms": float(np.percentile(lat, 99) * 1e3)}
            st["by_automaton"] = per
            sweep.append(st)
            _log(f"serving qps={qps}: p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms, "
                 f"{st['credited_GBps']:.2f} GB/s credited, {st['rejected']} rejected")
            # gate a sample of the served verdicts against the oracle
            for r, (name, data) in list(zip(done, reqs))[:: max(1, n // 48)]:
                a = automata[name]
                d = np.frombuffer(data, dtype=np.uint8)
                o = np.array([0, len(data)], dtype=np.int64)
                fn = native.dfa_batch if name.startswith("dfa") else native.nfa_batch
                wf, wn, _ = fn(a, d, o)
                gate.require_arrays(np.array([int(r.accepted)], np.int32),
                                    np.array([r.match_len], np.int32), wf, wn,
                                    f"served {name} request")
    if hasattr(loop, "close"):
        loop.close()
    return {"workload": "8 resident automata (DFAs 256/1K/4K/16K states, pattern unions "
                        "1K/4K/16K/64K states), Poisson arrivals, log-normal request "
                        f"lengths (median 16 KiB, clipped 4 MiB), HBM budget 32 GiB, "
                        f"{args.serve_slots} slots",
            "loop": "native scheduler (gfsm_server_*)" if args.serve_impl != "python"
                    else "Python scheduler (ServingLoop)",
            "slots": args.serve_slots, "sweep": sweep}


def bench_divergent(args, torch) -> dict:
    """High-divergence NFA: 65,536-state union (8000 patterns, class widths 16-64), 95%
    pattern bytes; bitpacked vs chain vs worklist, and where the autotuner's choice flips.

    The workload as configured latches within a few bytes per input (wide classes complete
    short patterns quickly): it is run on the full 1024 x 1 MiB batch. The non-latching
    control (no accepting states) keeps thousands of states active on every byte; the oracle
    cannot walk a full GiB of it, so the control is timed on a bounded sample and each
    technique is gated on a smaller one. A density sweep (fraction of pattern bytes) on the
    control shows the crossover between the techniques.
    """
    from gpufsm import workloads as W
    from gpufsm.backends.hip import _lib
    from gpufsm.backends.hip.handles import DeviceNFA

    from oracle import gate, native

    stream = torch.cuda.current_stream().cuda_stream
    out: dict = {"workload": "65536-state pattern union, 8000 patterns (lengths 5-12 adjusted to "
                             "the state count), class widths 16-64, 95% pattern bytes"}

    def run(dev, batch, tech):
        batch.run(dev, _lib.NFA_TECHNIQUES[tech], stream)
        torch.cuda.synchronize()
        return batch.results()

    def timed(dev, batch, tech, reps):
        wall, ks = _time_steps(torch, lambda: batch.run(dev, _lib.NFA_TECHNIQUES[tech], stream),
                               reps, 1)
        return wall / reps

    # as configured (latching), full batch
    t0 = time.perf_counter()
    pu, (data, off) = W.workload_divergent(count=args.div_count, length=args.div_length)
    _log(f"divergent: {pu.nfa.num_states} states, {off[-1] / 2**20:.0f} MiB "
         f"({time.perf_counter() - t0:.1f} s to generate)")
    threads = _threads()
    t0 = time.perf_counter()
    wf, wn, wex = native.nfa_batch(pu.nfa, data, off, threads=threads)
    oracle_s = time.perf_counter() - t0
    _log(f"divergent: oracle {oracle_s:.1f} s, {int(wf.sum())} accepted, "
         f"{int(wex.sum())} bytes examined")
    batch = DeviceBatch(torch, data, off)
    latching = {"examined_bytes": int(wex.sum()), "credited_bytes": int(off[-1]),
                "accepted": int(wf.sum()), "oracle_s": oracle_s, "oracle_threads": threads,
                "count": int(args.div_count), "techniques": {}}
    t0 = time.perf_counter()
    with DeviceNFA(pu.nfa) as dev:
        _log(f"divergent: automaton resident ({time.perf_counter() - t0:.1f} s, "
             f"{dev.props.device_bytes / 1e6:.0f} MB)")
        latching["num_tiles"] = int(dev.props.num_tiles)
        latching["resident_bytes"] = int(dev.props.device_bytes)
        for tech in ("bitpacked", "chain", "worklist"):
            t0 = time.perf_counter()
            f, n = run(dev, batch, tech)
            _log(f"divergent latching {tech}: first run {time.perf_counter() - t0:.2f} s")
            gate.require_arrays(f, n, wf, wn, f"hip divergent latching {tech}")
            sec = timed(dev, batch, tech, 3)
            _log(f"divergent latching {tech}: {1e3 * sec:.2f} ms per batch")
            latching["techniques"][tech] = {"ms": 1e3 * sec,
                                            "examined_GBps": int(wex.sum()) / sec / 1e9,
                                            "credited_GBps": int(off[-1]) / sec / 1e9}
    out["latching"] = latching
    del batch

    # non-latching control and the density sweep: a batch wide enough to fill the GPU,
    # gated on its first inputs (the oracle walks thousands of states per byte here)
    ctl = {}
    pu0 = W.pattern_union(8000, (5, 12), (16, 64), seed=0, total_states=65536, accepting=False)
    with DeviceNFA(pu0.nfa) as dev:
        for frac in args.div_sweep:
            lens_t = np.full(args.div_sample_inputs, args.div_sample_len, dtype=np.int64)
            d_t, o_t = W.pattern_inputs(pu0, lens_t, frac, seed=7)
            k_gate = min(4, args.div_sample_inputs)
            t0 = time.perf_counter()
            gf, gn, gex = native.nfa_batch(pu0.nfa, d_t[: int(o_t[k_gate])], o_t[: k_gate + 1])
            _log(f"divergent control p={frac}: oracle on {k_gate} inputs "
                 f"{time.perf_counter() - t0:.1f} s")
            bt = DeviceBatch(torch, d_t, o_t)
            k_wl = min(args.div_sample_inputs, 16)
            bw = DeviceBatch(torch, d_t[: int(o_t[k_wl])], o_t[: k_wl + 1])
            row = {}
            for tech in ("bitpacked", "chain", "worklist"):
                bb = bw if tech == "worklist" else bt  # the list walk: a small prefix only
                f, n = run(dev, bb, tech)
                gate.require_arrays(f[:k_gate], n[:k_gate], gf, gn,
                                    f"hip divergent control {tech} p={frac}")
                sec = timed(dev, bb, tech, 1 if tech == "worklist" else 2)
                nbytes = int(bb.total)
                row[tech] = {"examined_GBps": nbytes / sec / 1e9, "sample_bytes": nbytes,
                             "inputs": int(bb.count)}
                _log(f"divergent control p={frac} {tech}: {row[tech]['examined_GBps']:.5f} "
                     f"GB/s over {nbytes} bytes")
            best = max(row, key=lambda t: row[t]["examined_GBps"])
            ctl[f"p_pattern={frac}"] = {"rates": row, "fastest": best,
                                        "gated_inputs": k_gate}
    out["non_latching_control"] = ctl
    return out


def bench_general(args, torch) -> dict:
    """General (not chain-structured) NFA with dense active sets: the bit-parallel step
    against the worklist.

    The high-divergence union (65,536 states, 8000 patterns, class widths 16-64) with
    optional positions (epsilon edges s -> s + 1 on a quarter of the positions, so neither
    chain nor the tiled shift step applies) and no accepting state: with 95% pattern bytes
    thousands of states stay active on every byte and every byte is walked. Each technique
    is gated against the C oracle on a prefix the oracle walks in seconds, then timed on a
    batch of the same number of inputs (the worklist's inputs are shorter: its time per byte
    grows with the active set, and rates are per byte).
    """
    from gpufsm import workloads as W
    from gpufsm.backends.hip import _lib
    from gpufsm.backends.hip.handles import DeviceNFA

    from oracle import gate, native

    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    pu = W.pattern_union(8000, (5, 12), (16, 64), seed=0, total_states=65536, accepting=False)
    nfa = W.optional_positions(pu, 0.25, seed=1)
    shapes = {"bitpacked": (args.gen_count, args.gen_length),
              "worklist": (args.gen_wl_count, args.gen_wl_length)}
    out: dict = {"workload": "65536-state union of 8000 patterns (class widths 16-64) with "
                             f"{int(nfa.eps_row_ptr[-1])} optional positions (epsilon edges), "
                             "no accepting state, 95% pattern bytes",
                 "techniques": {}}
    threads = _threads()
    with DeviceNFA(nfa) as dev:
        assert not dev.props.chain_ok and dev.props.general_ok
        out["resident_bytes"] = int(dev.props.device_bytes)
        _log(f"general nfa: {nfa.num_states} states, not chain-structured, folded layout "
             f"{dev.props.device_bytes / 1e6:.1f} MB ({time.perf_counter() - t0:.1f} s)")
        for tech, (count, length) in shapes.items():
            d, o = W.pattern_inputs(pu, np.full(count, length, dtype=np.int64), 0.95, seed=2)
            batch = DeviceBatch(torch, d, o)
            # gate on a prefix (8 inputs x 2 KiB)
            gd, go = W.pattern_inputs(pu, np.full(8, 2048, dtype=np.int64), 0.95, seed=3)
            gb = DeviceBatch(torch, gd, go)
            wf, wn, _ = native.nfa_batch(nfa, gd, go, threads=threads)
            gb.run(dev, _lib.NFA_TECHNIQUES[tech], stream)
            torch.cuda.synchronize()
            f, n = gb.results()
            gate.require_arrays(f, n, wf, wn, f"hip general nfa {tech} (prefix)")
            wall, ks = _time_steps(torch, lambda: batch.run(dev, _lib.NFA_TECHNIQUES[tech],
                                                            stream), 2, 1)
            sec = wall / 2
            f, _ = batch.results()
            assert not f.any()
            out["techniques"][tech] = {"inputs": count, "input_bytes": length,
                                       "ms_per_batch": 1e3 * sec,
                                       "kernel_ms": float(np.mean(ks)),
                                       "examined_GBps": int(o[-1]) / sec / 1e9}
            _log(f"general nfa {tech}: {count} x {length} B in {1e3 * sec:.2f} ms, "
                 f"{int(o[-1]) / sec / 1e9:.5f} GB/s examined")
    r = out["techniques"]
    out["bitpacked_over_worklist"] = r["bitpacked"]["examined_GBps"] / r["worklist"]["examined_GBps"]
    return out


def general_line(args, torch) -> dict:
    """The general-NFA control as the line (for profiling its kernels alone)."""
    r = bench_general(args, torch)
    t = r["techniques"]["bitpacked"]
    nbytes = t["inputs"] * t["input_bytes"]
    algorithmic = nbytes + r["resident_bytes"] + 16 * t["inputs"] + 8
    return {
        "metric": "input GB/s batched vs plain run_batch on MI355X; p50/p99 latency at fixed QPS",
        "value": t["examined_GBps"],
        "unit": "GB/s",
        "ms_per_step": t["ms_per_batch"],
        "dtype": "u8",
        "config": {"workload": f"general NFA control (resident; value = examined input GB/s of "
                               f"bitpacked): {r['workload']}, {t['inputs']} inputs x "
                               f"{t['input_bytes']} bytes; the worklist under extras",
                   "resolved_technique": "bitpacked"},
        "roofline": {"bound": "hbm", "achieved": algorithmic / (t["kernel_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": algorithmic / (t["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                     "traffic": None, "kernel": "nfa_gather_kernel", "kernel_ms": t["kernel_ms"],
                     "algorithmic_bytes_per_launch": algorithmic},
        "cpu_baseline": None,
        "extras": {"general_nfa": r},
    }


def ragged_line(args, torch) -> dict:
    """The ragged NFA mix as the headline: resident, autotuned, graph-captured batches."""
    r = bench_ragged(args, torch)
    batched_s = r["batched_s"] if args.profile else r["batched_s"]
    count = r["count"]
    kernel_ms = r["kernel_ms"] if args.profile else r["batched"]["kernel_ms"]
    compulsory = r["resident_bytes"] + 8 * (count + 1) + 8 * count
    achieved = (r["examined_bytes"] + compulsory) / (kernel_ms * 1e-3) / 1e9
    technique = r["technique"] if args.profile else r["batched"]["technique"]
    extras = {} if args.profile else {"ragged_nfa": r}
    return {
        "metric": "input GB/s batched vs plain run_batch on MI355X; p50/p99 latency at fixed QPS",
        "value": r["examined_bytes"] / batched_s / 1e9,
        "unit": "GB/s",
        "ms_per_step": 1e3 * batched_s,
        "dtype": "u8",
        "config": {
            "workload": f"ragged NFA mix (batched: automaton and inputs resident in HBM, "
                        f"autotuned, graph-captured; value = examined input GB/s): "
                        f"{r['states']}-state union of 2000 patterns (lengths 5-40, class "
                        f"widths 1-32, wildcard start), {count} inputs with log-normal "
                        f"lengths (median 4 KiB, sigma 1, 64 B-1 MiB), 70% pattern bytes; the "
                        f"plain per-call rate is under extras",
            "num_states": r["states"], "num_inputs": count,
            "credited_bytes": r["credited_bytes"], "examined_bytes": r["examined_bytes"],
            "technique": args.ragged_technique, "resolved_technique": technique,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": _traffic("chain", f"ragged_nfa_{count}"),
            "kernel": "nfa_chain (zero, prep, rounds, scatter, reg tier, lds tier, tile tier, "
                      "finish)",
            "kernel_ms": kernel_ms,
            "algorithmic_bytes_per_launch": r["examined_bytes"] + compulsory,
        },
        "cpu_baseline": {
            "value": r["examined_bytes"] / r["cpu_s"] / 1e9,
            "unit": "GB/s",
            "cores": r["threads"],
            "kind": "port",
            "sample": f"full batch ({count} inputs, {r['credited_bytes']} bytes), C oracle over "
                      f"{r['threads']} threads, one run",
        },
        "extras": extras,
    }


def served_line(args, torch) -> dict:
    """The ragged mix served through the native serving loop only (for profiling that path):
    the mix, the oracle, the autotuner's technique for the resident batch, then the bulk
    submits of :func:`_serve_bulk`; value = examined bytes per second of the pageable-buffer
    submits (median), first submit to last verdict."""
    from gpufsm import workloads as W
    from gpufsm.batch import BatchEngine, DeviceBatch
    from gpufsm.tune import Autotuner

    from oracle import native

    count = args.ragged_count
    pu, (data, off) = W.workload_ragged(count=count)
    threads = _threads()
    t0 = time.perf_counter()
    wf, wn, wex = native.nfa_batch(pu.nfa, data, off, threads=threads)
    cpu_s = time.perf_counter() - t0
    examined, credited = int(wex.sum()), int(off[-1])
    s = torch.cuda.current_stream()
    engine = BatchEngine(tuner=Autotuner(persist=False))
    batch = DeviceBatch(data, off, stream=s)
    choice = engine.enqueue(pu.nfa, batch, args.ragged_technique, stream=s)
    batch.fetch(s)
    engine.close()
    del batch
    inputs = [data[off[i]:off[i + 1]].tobytes() for i in range(count)]
    sv = _serve_bulk(torch, pu, data, off, inputs, wf, wn, choice.label())
    _log(f"served: {examined / sv['served_s'] / 1e9:.2f} GB/s from a pageable buffer, "
         f"{examined / sv['served_pinned_s'] / 1e9:.2f} GB/s pinned ({choice.label()})")
    return {
        "metric": "input GB/s batched vs plain run_batch on MI355X; p50/p99 latency at fixed QPS",
        "value": examined / sv["served_s"] / 1e9,
        "unit": "GB/s",
        "steps": 5,   # bulk submits from the pageable buffer (the median is the value)
        "warmup": 1,  # one untimed bulk submit
        "ms_per_step": 1e3 * sv["served_s"],
        "dtype": "u8",
        "config": {
            "workload": f"ragged NFA mix served: {count} requests submitted in one bulk call "
                        f"to the native serving loop, read in place from a pageable buffer "
                        f"(value), launches of up to 64 MiB; {pu.nfa.num_states}-state union "
                        f"of 2000 patterns; examined input GB/s, host copies included",
            "num_states": pu.nfa.num_states, "num_inputs": count,
            "credited_bytes": credited, "examined_bytes": examined,
            "technique": choice.label(),
        },
        "roofline": None,
        "cpu_baseline": {
            "value": examined / cpu_s / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"full batch ({count} inputs), C oracle over {threads} threads, one run",
        },
        "extras": {"served_pinned_GBps": examined / sv["served_pinned_s"] / 1e9,
                   "one_request_per_call_ms": 1e3 * sv["one_by_one_s"],
                   "launches": sv["launches"]},
    }


def divergent_line(args, torch) -> dict:
    """The high-divergence NFA as the line: the as-configured latching batch, each technique
    gated and timed; ``value`` is the fastest technique's examined input GB/s."""
    r = bench_divergent(args, torch)
    lat = r["latching"]
    best = max(lat["techniques"], key=lambda t: lat["techniques"][t]["examined_GBps"])
    ms = lat["techniques"][best]["ms"]
    count = lat["count"]
    algorithmic = lat["examined_bytes"] + lat["resident_bytes"] + 8 * (count + 1) + 8 * count
    achieved = algorithmic / (ms * 1e-3) / 1e9
    return {
        "metric": "input GB/s batched vs plain run_batch on MI355X; p50/p99 latency at fixed QPS",
        "value": lat["techniques"][best]["examined_GBps"],
        "unit": "GB/s",
        "ms_per_step": ms,
        "dtype": "u8",
        "config": {
            "workload": f"high-divergence NFA (resident in HBM; value = examined input GB/s of "
                        f"the fastest technique on the latching batch as configured): "
                        f"{r['workload']}, {count} inputs x {args.div_length} bytes; the "
                        f"non-latching control and its density sweep are under extras",
            "examined_bytes": lat["examined_bytes"], "credited_bytes": lat["credited_bytes"],
            "resolved_technique": best,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": f"nfa_{best}",
            "kernel_ms": ms,
            "algorithmic_bytes_per_launch": algorithmic,
        },
        "cpu_baseline": {
            "value": lat["examined_bytes"] / lat["oracle_s"] / 1e9,
            "unit": "GB/s",
            "cores": lat["oracle_threads"],
            "kind": "port",
            "sample": f"full latching batch ({count} inputs), C oracle over "
                      f"{lat['oracle_threads']} threads, one run",
        },
        "extras": {"divergent_nfa": r},
    }


GATE_COLUMNS = ("kind", "backend", "technique", "automaton", "inputs", "bytes", "mismatches",
                "status", "seconds")


def _environment() -> dict:
    """What the numbers depend on: host, runtime and device."""
    import platform

    env = {"python": platform.python_version(), "numpy": np.__version__,
           "host_cores": _threads(), "machine": platform.machine()}
    try:
        import torch

        env["torch"] = torch.__version__
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            env.update(device=p.name, compute_units=p.multi_processor_count,
                       hbm_bytes=int(p.total_memory),
                       arch=getattr(p, "gcnArchName", ""))
    except Exception as exc:  # pragma: no cover - depends on the machine
        env["torch"] = f"unavailable: {exc}"
    try:
        with open("/opt/rocm/.info/version") as f:
            env["rocm"] = f.read().strip()
    except OSError:
        env["rocm"] = "unknown"
    try:
        from gpufsm.backends.hip import _lib

        env["gpufsm_hip_abi"] = _lib.load().gfsm_abi_version()
        env["gpufsm_hip_lib"] = _lib.LIB_PATH
    except Exception as exc:
        env["gpufsm_hip_lib"] = f"unavailable: {exc}"
    return env


def gate_main(args) -> int:
    """Strict oracle gate over every registered HIP technique; schema-checked CSV.

    Walks the registry (so it checks whatever is registered, not a list in a comment) over
    the vendored golden corpus, seeded random NFAs with epsilon edges and wildcards, pattern
    unions (the chain-structured automata of the chain and bitpacked techniques) and random
    DFAs from 16 to 40000 states, comparing (accepted, match_len) with the C oracle. A
    required backend that is missing fails the run instead of skipping it
    (`scripts/oracle_gate.py:1-20` upstream). Exit 0 only if every check agreed.
    """
    import csv
    import random

    from gpufsm import workloads as W
    from gpufsm.core.dfa import DFA, random_dfa
    from gpufsm.core.nfa import NFA, NFABuilder
    from gpufsm.core.packing import pack_bytes
    from gpufsm.core.registry import Backend, Kind, is_available, list_techniques, \
        unavailable_reason
    from gpufsm import run_batch

    from oracle import native

    required = {b.strip().lower() for b in args.require.split(",") if b.strip()}
    env = _environment()
    print("== environment ==")
    for k, v in env.items():
        print(f"  {k:16s} {v}")
    print("== backends ==")
    missing = []
    for b in Backend:
        avail = is_available(b)
        print(f"  {b.value:8s} {'available' if avail else 'MISSING':10s} "
              f"{unavailable_reason(b) or ''}")
        if b.value in required and not avail:
            missing.append(b.value)
    if missing:
        print(f"FAIL: required backend(s) not available: {', '.join(missing)}")
        return 2

    rng = random.Random(args.gate_seed)
    nrng = np.random.default_rng(args.gate_seed)
    cases: list[tuple[str, object, list[bytes]]] = []
    golden = os.path.join(ROOT, "tests", "golden", "golden.json")
    if os.path.exists(golden):
        with open(golden) as f:
            g = json.load(f)
        for kind_key, make in (("nfa_cases", "nfa"), ("dfa_cases", "dfa")):
            for c in g.get(kind_key, [])[: args.gate_golden]:
                a = c[make]
                if make == "nfa":
                    auto = NFA(num_states=int(a["num_states"]), start_state=int(a["start_state"]),
                               accept=np.array(a["accept"], bool),
                               sym_row_ptr=np.array(a["sym_row_ptr"], np.int32),
                               sym_targets=np.array(a["sym_targets"], np.int32),
                               sym_symbols=np.array(a["sym_symbols"], np.int32),
                               eps_row_ptr=np.array(a["eps_row_ptr"], np.int32),
                               eps_targets=np.array(a["eps_targets"], np.int32))
                else:
                    auto = DFA(num_states=int(a["num_states"]), start_state=int(a["start_state"]),
                               accept=np.array(a["accept"], bool),
                               trans=np.array(a["trans"], np.int32))
                cases.append((f"golden:{c.get('id', '?')}", auto,
                              [bytes.fromhex(c["input_hex"])]))
    for n in (8, 64, 700):
        nb = NFABuilder()
        for _ in range(n):
            nb.add_state(accept=rng.random() < 0.02)
        nb.set_start(rng.randrange(n))
        for s_ in range(n):
            for _ in range(rng.randint(0, 3)):
                nb.add_transition(s_, 256 if rng.random() < 0.03 else rng.randrange(97, 105),
                                  rng.randrange(n))
            if rng.random() < 0.3:
                nb.add_epsilon(s_, rng.randrange(n))
        nfa = nb.build()
        inputs = [bytes(rng.randrange(97, 106) for _ in range(rng.randint(0, args.gate_length)))
                  for _ in range(args.gate_inputs)]
        cases.append((f"random_nfa:{n}", nfa, inputs))
    for patterns in (40, 600):
        pu = W.pattern_union(patterns, (3, 20), (1, 32), seed=patterns)
        lens = nrng.integers(0, args.gate_length * 8, size=args.gate_inputs)
        d, o = W.pattern_inputs(pu, lens, 0.7, seed=patterns)
        cases.append((f"pattern_union:{pu.nfa.num_states}", pu.nfa,
                      [d[o[i]:o[i + 1]].tobytes() for i in range(lens.size)]))
    for n in (16, 256, 4096, 40000):
        dfa = random_dfa(n, accept_prob=0.3 / n, seed=n)
        d, o = W.uniform_batch(args.gate_inputs, args.gate_length * 16, seed=n)
        cases.append((f"random_dfa:{n}", dfa, [d[o[i]:o[i + 1]].tobytes()
                                               for i in range(args.gate_inputs)]))

    rows, failures, checked = [], 0, 0
    print(f"== oracle gate: {len(cases)} automata ==")
    for name, auto, inputs in cases:
        kind = Kind.of(auto)
        data, off = pack_bytes(inputs)
        fn = native.nfa_batch if kind is Kind.NFA else native.dfa_batch
        wf, wn, _ = fn(auto, data, off)
        for backend in Backend:
            if backend is Backend.CPU or not is_available(backend):
                continue
            for tech in list_techniques(backend, kind):
                t0 = time.perf_counter()
                try:
                    res = run_batch(auto, inputs, backend=backend, technique=tech)
                except ValueError:
                    status, bad = "not-applicable", 0
                except Exception as exc:
                    status, bad = f"error:{type(exc).__name__}", len(inputs)
                else:
                    got_f = np.array([r.accepted for r in res], np.int32)
                    got_n = np.array([r.match_len for r in res], np.int32)
                    bad = int(((got_f != wf) | (got_n != wn)).sum())
                    status = "ok" if bad == 0 else "MISMATCH"
                secs = time.perf_counter() - t0
                checked += status != "not-applicable"
                failures += status not in ("ok", "not-applicable")
                rows.append({"kind": kind.value, "backend": backend.value, "technique": tech,
                             "automaton": name, "inputs": len(inputs),
                             "bytes": int(off[-1]), "mismatches": bad, "status": status,
                             "seconds": round(secs, 6)})
                if status not in ("ok", "not-applicable"):
                    print(f"  {status:14s} {kind.value}/{backend.value}/{tech} on {name}")
    summary = collections_counter(rows)
    for (kind, backend, tech), (n_ok, n_bad, n_na) in sorted(summary.items()):
        print(f"  {kind}/{backend}/{tech:10s} ok {n_ok:4d}  failed {n_bad:3d}  "
              f"not applicable {n_na:3d}")
    if args.gate_csv:
        with open(args.gate_csv, "w", newline="") as f:
            f.write("# " + json.dumps(env, sort_keys=True) + "\n")
            w = csv.DictWriter(f, fieldnames=GATE_COLUMNS)
            w.writeheader()
            w.writerows(rows)
        _check_gate_csv(args.gate_csv)
        print(f"  wrote {args.gate_csv} ({len(rows)} rows, schema checked)")
    if checked == 0:
        print("FAIL: nothing was checked")
        return 2
    print(f"{'PASS' if failures == 0 else 'FAIL'}: {checked} checks, {failures} failed")
    return 0 if failures == 0 else 1


def collections_counter(rows):
    out: dict = {}
    for r in rows:
        key = (r["kind"], r["backend"], r["technique"])
        ok, bad, na = out.get(key, (0, 0, 0))
        if r["status"] == "ok":
            ok += 1
        elif r["status"] == "not-applicable":
            na += 1
        else:
            bad += 1
        out[key] = (ok, bad, na)
    return out


def _check_gate_csv(path: str) -> None:
    """The written CSV has the environment line, the exact header and typed fields."""
    import csv

    with open(path) as f:
        first = f.readline()
        if not first.startswith("# "):
            raise ValueError("gate CSV: missing environment line")
        json.loads(first[2:])
        r = csv.DictReader(f)
        if tuple(r.fieldnames or ()) != GATE_COLUMNS:
            raise ValueError(f"gate CSV: header {r.fieldnames} != {GATE_COLUMNS}")
        for row in r:
            int(row["inputs"]), int(row["bytes"]), int(row["mismatches"]), float(row["seconds"])
            if row["kind"] not in ("nfa", "dfa") or not row["technique"] or not row["status"]:
                raise ValueError(f"gate CSV: bad row {row}")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dfa",
                    choices=["dfa", "ragged", "served", "divergent", "general"],
                    help="dfa: the DFA memory-bound face (default, with every extra); "
                         "ragged: the ragged NFA mix as the line; served: the ragged mix "
                         "served through the native loop only; divergent: the "
                         "high-divergence NFA as the line")
    ap.add_argument("--states", type=int, default=256)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--length", type=int, default=65536)
    ap.add_argument("--technique", default="auto")
    ap.add_argument("--pipeline", type=int, default=2,
                    help="DFA launches in flight (streams) for the headline; 1 = one batch at "
                         "a time")
    ap.add_argument("--coalesce", type=int, default=64,
                    help="DFA headline stream: batches per launch (copies back to back)")
    ap.add_argument("--ragged-count", type=int, default=16384)
    ap.add_argument("--ragged-technique", default="auto",
                    help='"auto" (autotuned) or a fixed label such as "chain:chunk=2048"')
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[1000.0, 4000.0, 16000.0])
    ap.add_argument("--serve-slots", type=int, default=4,
                    help="serving loop launch slots (streams with their own batch buffers)")
    ap.add_argument("--div-count", type=int, default=1024)
    ap.add_argument("--div-length", type=int, default=1 << 20)
    ap.add_argument("--div-sweep", type=float, nargs="+", default=[0.3, 0.6, 0.95])
    ap.add_argument("--div-sample-inputs", type=int, default=1024)
    ap.add_argument("--div-sample-len", type=int, default=1 << 14)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-inflight", type=int, default=2,
                    help="native serving loop: launches in flight per automaton")
    ap.add_argument("--serve-impl", choices=("native", "python"), default="native",
                    help="serving loop: the library's native scheduler or the Python one")
    ap.add_argument("--gen-count", type=int, default=1024)
    ap.add_argument("--gen-length", type=int, default=1 << 14)
    ap.add_argument("--gen-wl-count", type=int, default=1024)
    ap.add_argument("--gen-wl-length", type=int, default=1 << 10)
    ap.add_argument("--only", default="",
                    help="comma list of extra sections to run (ragged,plain,serving,"
                         "divergent,general); default all")
    ap.add_argument("--profile", action="store_true",
                    help="headline variant only, plus one device copy of the input bytes "
                         "(the calibration point for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--gate", action="store_true",
                    help="run only the strict oracle gate over every registered technique")
    ap.add_argument("--require", default="",
                    help="with --gate: comma list of backends that must be available")
    ap.add_argument("--gate-csv", default="", help="with --gate: write the results here")
    ap.add_argument("--gate-inputs", type=int, default=64)
    ap.add_argument("--gate-length", type=int, default=256)
    ap.add_argument("--gate-golden", type=int, default=10**6)
    ap.add_argument("--gate-seed", type=int, default=1048)
    args = ap.parse_args()
    if args.gate:
        return gate_main(args)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); inputs are "
              f"independent, so N GPUs would each run this batch: see the sharding notes.",
              file=sys.stderr)
        return 2
    if args.steps < 1:
        print("--steps must be >= 1", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU (torch.cuda.is_available() is False)", file=sys.stderr)
        return 2
    if args.workload == "ragged":
        out = ragged_line(args, torch)
    elif args.workload == "served":
        out = served_line(args, torch)
    elif args.workload == "divergent":
        out = divergent_line(args, torch)
    elif args.workload == "general":
        out = general_line(args, torch)
    else:
        out = bench_dfa(args, torch)
    if not args.profile and args.workload == "dfa":
        want = (set(filter(None, args.only.split(",")))
                or {"ragged", "plain", "serving", "divergent", "general"})
        if "ragged" in want:
            out["extras"]["ragged_nfa"] = bench_ragged(args, torch)
        if "plain" in want:
            out["extras"]["plain_nfa"] = bench_plain_nfa(args, torch)
        if "serving" in want:
            out["extras"]["serving"] = bench_serving(args, torch)
        if "divergent" in want:
            out["extras"]["divergent_nfa"] = bench_divergent(args, torch)
        if "general" in want:
            out["extras"]["general_nfa"] = bench_general(args, torch)
    line = {
        "metric": out["metric"],
        "value": out["value"],
        "unit": out["unit"],
        "n_gpus": 1,
        "steps": out.get("steps", args.steps),
        "warmup": out.get("warmup", args.warmup),
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": out["dtype"],
        "data": "synthetic",
        "config": out["config"],
        "roofline": out["roofline"],
        "cpu_baseline": out["cpu_baseline"],
    }
    if "single_batch" in out:
        line["single_batch"] = out["single_batch"]
    ex = out["extras"]
    rag = ex.get("ragged_nfa") or {}
    if rag:
        # the metric's comparison: batched (resident, and host-staged) over plain per-call
        line["batched_vs_plain"] = {
            "workload": f"ragged NFA mix ({rag['states']}-state union, {rag['count']} inputs), "
                        f"examined bytes",
            "plain_GBps": rag["plain_per_call"]["examined_GBps"],
            "batched_resident_GBps": rag["batched"]["examined_GBps"],
            "batched_host_pipelined_GBps": rag["batched_host_pipelined"]["examined_GBps"],
            "resident_over_plain": rag["batched_over_plain"],
            "host_pipelined_over_plain": rag["host_pipelined_over_plain"],
            "served_GBps": rag["served"]["examined_GBps"],
            "served_over_plain": rag["served_over_plain"],
            "served_pinned_GBps": rag["served_pinned"]["examined_GBps"],
            "served_pinned_over_plain": rag["served_pinned_over_plain"],
        }
    srv = ex.get("serving") or {}
    if srv:
        line["serving_latency"] = [{"qps": st["qps_offered"], "p50_ms": st["p50_ms"],
                                    "p99_ms": st["p99_ms"],
                                    "credited_GBps": st["credited_GBps"]}
                                   for st in srv["sweep"]]
    line["extras"] = ex
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
