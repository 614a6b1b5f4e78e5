"""Autotuner: picks the technique and launch tuning per automaton and batch shape.

Every technique is bit-exact (the parity suite checks each against the oracle), so the
choice only moves time. The tuner times each candidate on a bounded sample of the real batch
(a prefix of its inputs, run in place: same device buffers, smaller count), keeps the
fastest, and records every candidate's measured examined-byte rate next to the choice, in a
JSON table keyed by a shape signature (the sample is described in
:meth:`Autotuner.sample_count`):

    kind | device arch | states (power-of-two bucket) | structure flags | mean input length
    (bucket) | batch size (bucket)

so a later batch of the same class reuses the decision without timing anything. The table
persists across processes (``GFSM_TUNE_TABLE`` or ``~/.cache/gpufsm/hip_autotune.json``).

Candidates
  NFA, chain-structured: ``chain`` with chunks of 2, 4 and 8 KiB, ``chain`` with one item
  per wave in the register tier (for active sets beyond 32 listed states), ``chain``
  without the byte-pair lists, ``bitpacked`` (dense active sets) and, up to 4096 states,
  ``worklist``.
  NFA, general: ``worklist`` (sparse active sets), ``bitpacked`` (the bit-parallel step,
  dense active sets; up to 65,536 states).
  Any NFA of at most 512 states with the folded layout: also ``register`` (one lane per
  input).
  DFA: ``gather``; ``lds`` when the table fits LDS; ``split`` (segment 1, 2 and 4 KiB when
  the table fits LDS, the library's choice otherwise) when inputs average at least 1 KiB.

No upstream counterpart: upstream picks techniques by hand (`bench/` drivers).
"""

from __future__ import annotations

import json
import math
import os
import threading
from dataclasses import dataclass, field

from .backends.hip import _lib
from .backends.hip.handles import make_tuning


@dataclass(frozen=True)
class Choice:
    """A technique id plus tuning fields (zeros = library default)."""

    technique: str
    chunk_bytes: int = 0
    segment_bytes: int = 0
    flags: int = 0

    def label(self) -> str:
        parts = [self.technique]
        if self.chunk_bytes:
            parts.append(f"chunk={self.chunk_bytes}")
        if self.segment_bytes:
            parts.append(f"segment={self.segment_bytes}")
        if self.flags & _lib.TUNE_CHAIN_FIRST:
            parts.append("first")
        if self.flags & _lib.TUNE_CHAIN_WAVE64:
            parts.append("wave64")
        if self.flags & _lib.TUNE_CHAIN_TRIPLE:
            parts.append("triple")
        return ":".join(parts)

    def tuning(self, max_input_bytes: int = 0):
        """The ``gfsm_tuning`` of this choice for a batch whose longest input is
        ``max_input_bytes`` (0 = not known)."""
        return make_tuning(self.chunk_bytes, self.segment_bytes, self.flags, max_input_bytes)

    @classmethod
    def parse(cls, label: str) -> Choice:
        """Inverse of :meth:`label`: ``"chain:chunk=2048"``, ``"split:segment=1024"``, ..."""
        name, *parts = label.split(":")
        kw = {"chunk_bytes": 0, "segment_bytes": 0, "flags": 0}
        for part in parts:
            if part == "first":
                kw["flags"] |= _lib.TUNE_CHAIN_FIRST
            elif part == "wave64":
                kw["flags"] |= _lib.TUNE_CHAIN_WAVE64
            elif part == "triple":
                kw["flags"] |= _lib.TUNE_CHAIN_TRIPLE
            elif part.startswith("chunk="):
                kw["chunk_bytes"] = int(part[6:])
            elif part.startswith("segment="):
                kw["segment_bytes"] = int(part[8:])
            else:
                raise ValueError(f"bad technique label {label!r}")
        return cls(name, **kw)


@dataclass
class Decision:
    choice: Choice
    rates: dict[str, float] = field(default_factory=dict)  # label -> examined GB/s


def _bucket(x: int) -> int:
    return 0 if x <= 0 else 1 << int(math.ceil(math.log2(x)))


def default_table_path() -> str:
    env = os.environ.get("GFSM_TUNE_TABLE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "gpufsm", "hip_autotune.json")


def candidates(kind: str, props, mean_len: float) -> list[Choice]:
    if kind == "nfa":
        lanes = ([Choice("register")] if props.general_ok and props.num_states <= 512
                 else [])
        if not props.chain_ok:
            out = [Choice("worklist")]
            if props.general_ok and props.num_states <= 65536:
                out.append(Choice("bitpacked"))
            return out + lanes
        out = [Choice("chain", chunk_bytes=c) for c in (2048, 4096, 8192)]
        out += [Choice("chain", chunk_bytes=4096, flags=_lib.TUNE_CHAIN_WAVE64),
                Choice("chain", chunk_bytes=4096, flags=_lib.TUNE_CHAIN_FIRST),
                Choice("bitpacked")]
        if props.start_always and props.num_tiles <= 16:
            # the lane tier (byte-triple lists; one item per lane, so shorter chunks)
            out += [Choice("chain", chunk_bytes=c, flags=_lib.TUNE_CHAIN_TRIPLE)
                    for c in (512, 1024, 2048)]
        if props.num_states <= 4096:  # the general list walk only competes on small unions
            out.append(Choice("worklist"))
        return out + lanes
    out = [Choice("gather")]
    if props.lds_ok:
        out.append(Choice("lds"))
    if mean_len >= 1024:
        if props.lds_ok:
            out += [Choice("split", segment_bytes=s) for s in (1024, 2048, 4096)]
        else:
            out.append(Choice("split"))
    return out


class Autotuner:
    """Chooses a :class:`Choice` per shape class; see the module docstring."""

    def __init__(self, path: str | None = None, sample_bytes: int = 128 << 20,
                 repeats: int = 3, persist: bool = True,
                 dfa_sample_bytes: int = 1 << 30) -> None:
        self.path = path or default_table_path()
        self.sample_bytes = int(sample_bytes)
        self.dfa_sample_bytes = int(dfa_sample_bytes)
        self.repeats = int(repeats)
        self.persist = persist
        self._lock = threading.Lock()
        self.table: dict[str, dict] = {}
        if persist and os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    def sample_count(self, kind: str, offsets, count: int) -> int:
        """How many leading inputs of a batch a candidate is timed on.

        NFA: the prefix of at most ``sample_bytes`` (at least one input). The chain lane
        tier runs one (input, chunk) item per lane and needs about as many items as the
        device has lanes (some 260,000, so 128 MiB of 512-byte chunks) before its rate
        settles; a smaller prefix would misjudge it against the register tier. DFA: the
        whole batch up to ``dfa_sample_bytes``. The per-input DFA walk runs one lane per
        input and its rate grows with inputs per CU (256 to 1024), while the segmented walk
        fills the device from any count; a byte-bounded prefix of a batch of short inputs
        would leave the per-input walk a few lanes per CU and misjudge it.
        """
        import numpy as np

        limit = self.dfa_sample_bytes if kind == "dfa" else self.sample_bytes
        k = int(np.searchsorted(offsets, limit, side="right")) - 1
        return max(1, min(count, k))

    @staticmethod
    def signature(kind: str, arch: str, props, count: int, total: int) -> str:
        mean = total // count if count else 0
        if kind == "nfa":
            struct = (f"chain{props.chain_ok}always{props.start_always}"
                      f"general{props.general_ok}")
        else:
            struct = f"w{props.entry_bytes}lds{props.lds_ok}"
        return (f"{kind}|{arch}|n{_bucket(props.num_states)}|{struct}|len{_bucket(mean)}"
                f"|cnt{_bucket(count)}")

    def lookup(self, key: str) -> Decision | None:
        entry = self.table.get(key)
        if not entry:
            return None
        c = entry["choice"]
        return Decision(Choice(c["technique"], c.get("chunk_bytes", 0),
                               c.get("segment_bytes", 0), c.get("flags", 0)),
                        dict(entry.get("rates", {})))

    def decide(self, key: str, kind: str, props, count: int, total: int, timer) -> Decision:
        """The recorded decision for ``key``, or time the candidates with ``timer``.

        ``timer(choice, give_up) -> (seconds, examined_bytes)`` runs one candidate over the
        sample (repeating it only while it stays under ``give_up`` seconds, if given); it may
        raise for a candidate that does not apply, which then drops out.
        """
        with self._lock:
            hit = self.lookup(key)
            if hit is not None:
                return hit
        mean = total / count if count else 0.0
        best: tuple[float, Choice] | None = None
        rates: dict[str, float] = {}
        for cand in candidates(kind, props, mean):
            try:
                secs, examined = timer(cand, None if best is None else 8.0 * best[0])
            except _lib.HIPError:
                continue
            rate = examined / secs / 1e9 if secs > 0 else float("inf")
            rates[cand.label()] = round(rate, 4)
            if best is None or secs < best[0]:
                best = (secs, cand)
        if best is None:
            raise RuntimeError(f"no technique could run shape {key}")
        dec = Decision(best[1], rates)
        with self._lock:
            self.table[key] = {"choice": {"technique": dec.choice.technique,
                                          "chunk_bytes": dec.choice.chunk_bytes,
                                          "segment_bytes": dec.choice.segment_bytes,
                                          "flags": dec.choice.flags},
                               "rates": rates}
            self.save()
        return dec

    def save(self) -> None:
        if not self.persist:
            return
        try:
            os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # an unwritable cache only costs re-tuning in the next process
