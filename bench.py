#!/usr/bin/env python3
"""Benchmark of the gpufsm HIP hot path on one MI355X; prints ONE JSON line.

Workload (default ``dfa``): the DFA memory-bound face — a seeded random complete DFA with
4096 states (uint16 table with folded accept bits, 2 MiB), a batch of 4096 inputs of 64 KiB
uniform random bytes already resident in HBM. The headline ``value`` is measured on the
non-latching control (``accept_prob = 0``, as the upstream latch control does), where every
input byte is examined, so input GB/s counts bytes actually read. The latching variant as
configured (10% accepting states) is reported beside it with both examined and credited
bytes, and so is the plain per-call drop-in (``gpufsm.run_batch``: upload, copy, launch,
copy back).

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
    "split": "dfa_split (init, walk, fix x2, event, finish)",
}


def _measure_dfa(torch, batch, dfa, technique: str, data, offsets, steps: int, warmup: int,
                 what: str) -> dict:
    """Gate one DFA/technique on the full batch against the C oracle, then time it."""
    from gpufsm.backends.hip import _lib
    from gpufsm.backends.hip.handles import DeviceDFA

    from oracle import gate, native

    tech = _lib.DFA_TECHNIQUES[technique]
    stream = torch.cuda.current_stream().cuda_stream
    with DeviceDFA(dfa) as dev:
        batch.run(dev, tech, stream)
        torch.cuda.synchronize()
        f, m = batch.results()
        wf, wm, wex = native.dfa_batch(dfa, data, offsets)
        gate.require_arrays(f, m, wf, wm, f"hip dfa {technique} ({what})")
        wall, ks = _time_steps(torch, lambda: batch.run(dev, tech, stream), steps, warmup)
        entry_bytes = dev.props.entry_bytes
    examined = int(wex.sum())
    credited = int(offsets[-1])
    step_s = wall / steps
    out = {
        "workload": what,
        "technique": technique,
        "accepted": int(f.sum()),
        "examined_bytes": examined,
        "credited_bytes": credited,
        "ms_per_step": 1e3 * step_s,
        "kernel_ms": float(np.mean(ks)),
        "examined_GBps": examined / step_s / 1e9,
        "credited_GBps": credited / step_s / 1e9,
        "table_bytes": dfa.num_states * 256 * entry_bytes,
    }
    _log(f"{what} [{technique}]: {out['examined_GBps']:.1f} GB/s examined, "
         f"{out['ms_per_step']:.3f} ms/step")
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
                        args.warmup, f"dfa{n} non-latching")
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
                    f"dfa{other_n} {'latching 10%' if latching else 'non-latching'}"))
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
    kernel_ms = head["kernel_ms"]
    achieved = (examined + compulsory) / (kernel_ms * 1e-3) / 1e9
    tech = args.technique
    if tech == "auto":
        tech = "split" if length >= 4 * (1024 if n <= 320 else max(1024, 4 * n)) else (
            "lds" if n <= 320 else "gather")
    kernel_name = KERNELS.get(tech, tech)
    workload_name = f"dfa{n}_{count}x{length}_nonlatching"
    return {
        "metric": "input GB/s (examined bytes), DFA walk, inputs resident in HBM",
        "value": head["examined_GBps"],
        "unit": "GB/s",
        "ms_per_step": head["ms_per_step"],
        "dtype": "u8",
        "config": {
            "workload": f"DFA memory-bound face, non-latching control: random complete DFA "
                        f"{n} states x 256 (uint16 table, accept_prob 0), {count} inputs x "
                        f"{length} uniform bytes (seed 0)",
            "num_states": n, "num_inputs": count, "input_bytes": length,
            "technique": args.technique, "resolved_technique": tech,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": _traffic(tech, workload_name),
            "kernel": kernel_name,
            "kernel_ms": kernel_ms,
            "algorithmic_bytes_per_launch": examined + compulsory,
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dfa", choices=["dfa"])
    ap.add_argument("--states", type=int, default=256)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--length", type=int, default=65536)
    ap.add_argument("--technique", default="auto")
    ap.add_argument("--profile", action="store_true",
                    help="headline variant only, plus one device copy of the input bytes "
                         "(the calibration point for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
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
    out = bench_dfa(args, torch)
    line = {
        "metric": out["metric"],
        "value": out["value"],
        "unit": out["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": out["dtype"],
        "data": "synthetic",
        "config": out["config"],
        "roofline": out["roofline"],
        "cpu_baseline": out["cpu_baseline"],
        "extras": out["extras"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
