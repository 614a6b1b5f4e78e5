"""Command line: ``python -m gpufsm {env,list,verify,bench}``.

Interface parity: `src/gpufsm/cli.py:27-160` (the same four commands). Differences follow
from this package: ``env`` also reports the HIP library (path, ABI version) and the
device's architecture, and names why HIP is unavailable when it is; ``verify`` checks every registered HIP technique against the labelled example inputs
(the CPU oracle is test infrastructure here, not part of the package) and checks that all
techniques agree on seeded random batches of both automaton kinds.
"""

from __future__ import annotations

import argparse
import platform
import sys

from .api import benchmark, run, run_batch
from .backends import IMPORT_ERRORS
from .core.registry import (
    Backend,
    Kind,
    available_backends,
    list_kinds,
    list_techniques,
    unavailable_reason,
)
from .examples import EXAMPLES


def _techniques_line(backend: Backend, kind: Kind) -> str:
    return ", ".join(list_techniques(backend, kind)) or "-"


def _cmd_env(_: argparse.Namespace) -> int:
    print(f"python   : {platform.python_version()} ({sys.platform})")
    try:
        import numpy

        print(f"numpy    : {numpy.__version__}")
    except Exception:  # pragma: no cover
        print("numpy    : (missing)")
    from .backends.hip import _lib

    try:
        lib = _lib.load()
        print(f"hip lib  : {_lib.LIB_PATH} (ABI {lib.gfsm_abi_version()})")
        try:
            print(f"device 0 : {_lib.device_arch(0)}")
        except _lib.HIPError as exc:
            print(f"device 0 : none ({exc})")
    except _lib.HIPUnavailable as exc:
        print(f"hip lib  : unavailable ({exc})")
    avail = available_backends()
    print(f"backends : {', '.join(b.value for b in avail) or '(none)'}")
    for b in avail:
        for k in list_kinds(b):
            print(f"  {b.value:7s} {k.value}: {_techniques_line(b, k)}")
    missing = [b for b in Backend if b not in avail]
    if missing:
        print(f"missing  : {', '.join(b.value for b in missing)}")
        for b in missing:
            why = IMPORT_ERRORS.get(b.value) or unavailable_reason(b)
            print(f"  {b.value:7s} {why}" if why else f"  {b.value:7s} not part of this build")
    return 0


def _cmd_list(_: argparse.Namespace) -> int:
    avail = available_backends()
    for b in Backend:
        status = "available" if b in avail else "unavailable"
        kinds = list_kinds(b)
        if not kinds:
            print(f"{b.value:7s} [{status}]: -")
            continue
        for k in kinds:
            print(f"{b.value:7s} [{status}] {k.value}: {_techniques_line(b, k)}")
    return 0


def _verify_random(backend: Backend) -> int:
    """Every technique of each kind gives the same verdicts on seeded random batches."""
    import numpy as np

    from . import workloads as W

    failures = 0
    pu = W.pattern_union(200, seed=1)
    lens = np.random.default_rng(1).integers(0, 4000, size=64)
    data, off = W.pattern_inputs(pu, lens, 0.7, seed=2)
    cases = [(Kind.NFA, pu.nfa, data, off)]
    dfa = W.dense_dfa(256, 0.001, seed=3)
    ddata, doff = W.uniform_batch(32, 20000, seed=4)
    cases.append((Kind.DFA, dfa, ddata, doff))
    for kind, automaton, d, o in cases:
        inputs = [d[o[i]:o[i + 1]].tobytes() for i in range(o.size - 1)]
        first = None
        for tech in list_techniques(backend, kind):
            try:
                got = [(r.accepted, r.match_len)
                       for r in run_batch(automaton, inputs, backend=backend, technique=tech)]
            except ValueError:
                print(f"     {backend.value}/{kind.value}/{tech:10s} not applicable")
                continue
            except Exception as exc:
                failures += 1
                print(f"     {backend.value}/{kind.value}/{tech:10s} ERROR "
                      f"{type(exc).__name__}: {exc}")
                continue
            if first is None:
                first = (tech, got)
                print(f"     {backend.value}/{kind.value}/{tech:10s} reference for agreement "
                      f"({sum(a for a, _ in got)}/{len(got)} accepted)")
                continue
            agree = got == first[1]
            failures += not agree
            print(f"     {backend.value}/{kind.value}/{tech:10s} "
                  f"{'agrees' if agree else 'DIFFERS from ' + first[0]}")
    return failures


def _cmd_verify(args: argparse.Namespace) -> int:
    """Each available technique against the labelled examples, then cross-agreement."""
    failures = 0
    backends = [b for b in available_backends() if b != Backend.CPU]
    if not backends:
        print("no GPU backend is available: nothing to verify")
        for b in Backend:
            why = IMPORT_ERRORS.get(b.value) or unavailable_reason(b)
            if why:
                print(f"  {b.value}: {why}")
        return 1
    for name, factory in EXAMPLES.items():
        nfa, labelled = factory()
        for data, expected in labelled:
            for b in backends:
                for tech in list_techniques(b, Kind.NFA):
                    label = f"{b.value}/{tech}"
                    try:
                        res = run(nfa, data, backend=b, technique=tech)
                    except ValueError:
                        continue  # technique does not apply to this automaton
                    except Exception as exc:
                        failures += 1
                        print(f"[ERROR] {label:16s} {name:16s} {data!r}: "
                              f"{type(exc).__name__}: {exc}")
                        continue
                    ok = res.accepted == expected
                    failures += not ok
                    print(f"[{'ok' if ok else 'FAIL'}] {label:16s} {name:16s} {data!r:16s} -> "
                          f"{res.accepted} (want {expected})")
    if not args.quick:
        for b in backends:
            failures += _verify_random(b)
    print(f"\n{failures} failure(s).")
    return 1 if failures else 0


def _cmd_bench(args: argparse.Namespace) -> int:
    factory = EXAMPLES.get(args.example)
    if factory is None:
        print(f"unknown example {args.example!r}; choose from {list(EXAMPLES)}", file=sys.stderr)
        return 2
    nfa, _ = factory()
    data = (b"abcd" * (args.size // 4 + 1))[: args.size]
    stats = benchmark(nfa, data, backend=Backend(args.backend), technique=args.technique,
                      repeats=args.repeats, warmup=args.warmup)
    print(f"{stats.backend}/{stats.technique}: mean {stats.mean_ms:.4f} ms "
          f"(CI95 +/- {stats.ci95_ms:.4f}, n={stats.n}); accepted={stats.accepted}, "
          f"match_len={stats.match_len}")
    return 0


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="gpufsm", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("env", help="interpreter, library, device and backends").set_defaults(
        fn=_cmd_env)
    sub.add_parser("list", help="registered techniques per backend").set_defaults(fn=_cmd_list)
    v = sub.add_parser("verify", help="check every technique against labelled examples")
    v.add_argument("--quick", action="store_true", help="labelled examples only")
    v.set_defaults(fn=_cmd_verify)
    bp = sub.add_parser("bench", help="time one example automaton")
    bp.add_argument("--example", default="ab_star_c_plus_d")
    bp.add_argument("--size", type=int, default=4096)
    bp.add_argument("--backend", default="hip", choices=[b.value for b in Backend])
    bp.add_argument("--technique", default=None)
    bp.add_argument("--repeats", type=int, default=10)
    bp.add_argument("--warmup", type=int, default=3)
    bp.set_defaults(fn=_cmd_bench)
    args = ap.parse_args(argv)
    return args.fn(args)


if __name__ == "__main__":  # pragma: no cover
    sys.exit(main())
