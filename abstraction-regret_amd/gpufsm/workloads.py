"""Seeded synthetic automata and input batches for the benchmark workloads.

Every generator draws from ``numpy.random.default_rng(seed)`` streams only, so a workload
is fully determined by its arguments. Parameters the workload descriptions leave open are
fixed here and named in the docstrings:

* ``plain_nfa``       random NFA, 1024 states, 4 labelled edges per state (uniform byte,
                      uniform target), 5% accepting states drawn without replacement from
                      the non-start states, start 0, no epsilon, no wildcard.
* ``dense_dfa``       ``random_dfa`` (`src/gpufsm/core/dfa.py:44-57`) with 10% accepting
                      states; ``accept_prob=0`` is the non-latching control (the upstream
                      control of `scripts/dfa_latch_control.py:102`).
* ``pattern_union``   unanchored homogeneous NFA: start state 0 with a wildcard self-loop,
                      then one chain of states per pattern; position ``i`` of a pattern is a
                      byte class whose width is uniform in ``[wmin, wmax]`` (width 1 is a
                      literal), class bytes drawn without replacement; the last position of
                      every pattern accepts (none with ``accepting=False``).
* ``pattern_inputs``  each byte is, with probability ``p_pattern``, a uniform byte of the
                      class of a uniformly chosen pattern position, otherwise a uniform byte
                      (i.i.d. bytes, drawn from that mixture by an inverse-CDF table).
* ``lognormal_lengths`` lengths ``exp(N(ln median, sigma))`` rounded and clipped.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .core.dfa import DFA, random_dfa
from .core.nfa import ANY_SYMBOL, NFA


def plain_nfa(num_states: int = 1024, edges_per_state: int = 4, accept_frac: float = 0.05,
              seed: int = 0) -> NFA:
    """The plain drop-in NFA (1024 states, 4 random edges each, 5% accepting)."""
    rng = np.random.default_rng(seed)
    n = num_states
    targets = rng.integers(0, n, size=n * edges_per_state, dtype=np.int32)
    labels = rng.integers(0, 256, size=n * edges_per_state, dtype=np.int32)
    accept = np.zeros(n, dtype=bool)
    n_acc = int(round(accept_frac * n))
    accept[1 + rng.choice(n - 1, size=n_acc, replace=False)] = True
    row = (np.arange(n + 1, dtype=np.int64) * edges_per_state).astype(np.int32)
    return NFA(num_states=n, start_state=0, accept=accept, sym_row_ptr=row,
               sym_targets=targets, sym_symbols=labels,
               eps_row_ptr=np.zeros(n + 1, dtype=np.int32),
               eps_targets=np.zeros(0, dtype=np.int32))


def dense_dfa(num_states: int, accept_prob: float = 0.10, seed: int = 0) -> DFA:
    return random_dfa(num_states, accept_prob=accept_prob, seed=seed)


def uniform_batch(count: int, length: int, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """``count`` inputs of ``length`` uniform bytes: ``(data uint8, offsets int64)``."""
    rng = np.random.default_rng(seed)
    data = rng.integers(0, 256, size=count * length, dtype=np.uint8)
    offsets = np.arange(count + 1, dtype=np.int64) * length
    return data, offsets


@dataclass
class PatternUnion:
    nfa: NFA
    lengths: np.ndarray        # positions per pattern
    class_start: np.ndarray    # [positions + 1] offsets into class_bytes
    class_bytes: np.ndarray    # concatenated class members (uint8)


def pattern_union(num_patterns: int, len_range: tuple[int, int] = (5, 40),
                  width_range: tuple[int, int] = (1, 32), *, seed: int = 0,
                  total_states: int | None = None, accepting: bool = True) -> PatternUnion:
    """Union of ``num_patterns`` random class patterns behind an always-active start.

    ``total_states`` (including the start) adjusts pattern lengths, within ``len_range``,
    until the state count is exactly that number.
    """
    rng = np.random.default_rng(seed)
    lo, hi = len_range
    lengths = rng.integers(lo, hi + 1, size=num_patterns).astype(np.int64)
    if total_states is not None:
        want = total_states - 1
        if not lo * num_patterns <= want <= hi * num_patterns:
            raise ValueError("total_states is not reachable with these lengths")
        diff = int(want - lengths.sum())
        while diff != 0:
            i = int(rng.integers(0, num_patterns))
            if diff > 0 and lengths[i] < hi:
                lengths[i] += 1
                diff -= 1
            elif diff < 0 and lengths[i] > lo:
                lengths[i] -= 1
                diff += 1
    positions = int(lengths.sum())
    n = positions + 1
    widths = rng.integers(width_range[0], width_range[1] + 1, size=positions)
    class_start = np.zeros(positions + 1, dtype=np.int64)
    np.cumsum(widths, out=class_start[1:])
    class_bytes = np.empty(int(class_start[-1]), dtype=np.uint8)
    for p in range(positions):
        class_bytes[class_start[p]:class_start[p + 1]] = rng.choice(256, size=int(widths[p]),
                                                                    replace=False)
    # states: 0 = start; pattern positions numbered consecutively from 1
    pat_first = np.zeros(num_patterns, dtype=np.int64)
    np.cumsum(lengths[:-1], out=pat_first[1:])
    pat_first += 1
    is_first = np.zeros(n, dtype=bool)
    is_first[pat_first] = True
    is_last = np.zeros(n, dtype=bool)
    is_last[pat_first + lengths - 1] = True
    # edges: start -> ANY self-loop and class(first position) of every pattern;
    #        state s (not last) -> s + 1 on class(s + 1)
    start_deg = 1 + int(widths[pat_first - 1].sum())
    deg = np.zeros(n, dtype=np.int64)
    deg[0] = start_deg
    inner = np.flatnonzero(~is_last[1:]) + 1
    deg[inner] = widths[inner]  # state s -> s+1 uses the class of position (s+1)-1 = s
    row = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=row[1:])
    nnz = int(row[-1])
    targets = np.empty(nnz, dtype=np.int32)
    labels = np.empty(nnz, dtype=np.int32)
    targets[0] = 0
    labels[0] = ANY_SYMBOL
    k = 1
    for f in pat_first:
        cls = class_bytes[class_start[f - 1]:class_start[f]]
        targets[k:k + cls.size] = f
        labels[k:k + cls.size] = cls
        k += cls.size
    for s in inner:
        cls = class_bytes[class_start[s]:class_start[s + 1]]
        a = row[s]
        targets[a:a + cls.size] = s + 1
        labels[a:a + cls.size] = cls
    if row[-1] > np.iinfo(np.int32).max:
        raise ValueError("too many edges")
    accept = is_last.copy() if accepting else np.zeros(n, dtype=bool)
    nfa = NFA(num_states=n, start_state=0, accept=accept, sym_row_ptr=row.astype(np.int32),
              sym_targets=targets, sym_symbols=labels,
              eps_row_ptr=np.zeros(n + 1, dtype=np.int32),
              eps_targets=np.zeros(0, dtype=np.int32))
    return PatternUnion(nfa=nfa, lengths=lengths, class_start=class_start,
                        class_bytes=class_bytes)


def optional_positions(pu: PatternUnion, p_optional: float = 0.25, seed: int = 0) -> NFA:
    """The union with optional positions: with probability ``p_optional`` a pattern position
    s that is not its pattern's last gets an epsilon edge s -> s + 1, so reaching s also
    counts as having matched s + 1 (the regex ``c?`` for the class of s + 1).

    The epsilon edges make the automaton general (not chain-structured), which is the
    non-chain control for the general NFA techniques; the active sets stay as large as the
    union's, or larger.
    """
    rng = np.random.default_rng(seed)
    nfa = pu.nfa
    n = int(nfa.num_states)
    first = np.zeros(pu.lengths.size, dtype=np.int64)
    np.cumsum(pu.lengths[:-1], out=first[1:])
    first += 1
    last = first + pu.lengths - 1
    is_last = np.zeros(n, dtype=bool)
    is_last[last] = True
    cand = np.flatnonzero(~is_last[1:]) + 1
    src = cand[rng.random(cand.size) < p_optional]
    counts = np.zeros(n, dtype=np.int64)
    counts[src] = 1
    erow = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=erow[1:])
    return NFA(num_states=n, start_state=0, accept=np.asarray(nfa.accept, dtype=bool).copy(),
               sym_row_ptr=nfa.sym_row_ptr, sym_targets=nfa.sym_targets,
               sym_symbols=nfa.sym_symbols, eps_row_ptr=erow,
               eps_targets=(src + 1).astype(np.int32))


def lognormal_lengths(count: int, median: int, sigma: float, lo: int, hi: int,
                      seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.normal(np.log(median), sigma, size=count))
    return np.clip(np.rint(raw), lo, hi).astype(np.int64)


def pattern_inputs(pu: PatternUnion, lengths: np.ndarray, p_pattern: float,
                   seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Inputs of the given lengths, biased toward the union's class bytes.

    Each byte is, independently, with probability ``p_pattern`` a uniform member of the class
    of a uniformly chosen pattern position, else a uniform byte. Bytes are i.i.d., so they
    are drawn directly from that 256-way mixture through a 65536-entry inverse-CDF table
    (probabilities rounded to 1/65536 by largest remainders): one table lookup per byte.
    """
    rng = np.random.default_rng(seed)
    offsets = np.zeros(lengths.size + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    total = int(offsets[-1])
    positions = pu.class_start.size - 1
    width = np.diff(pu.class_start).astype(np.float64)
    owner = np.repeat(np.arange(positions), np.diff(pu.class_start))
    prob = np.full(256, (1.0 - p_pattern) / 256.0)
    np.add.at(prob, pu.class_bytes.astype(np.int64), p_pattern / positions / width[owner])
    slots = prob * 65536.0
    counts = np.floor(slots).astype(np.int64)
    rest = 65536 - int(counts.sum())
    if rest > 0:
        counts[np.argsort(-(slots - counts), kind="stable")[:rest]] += 1
    table = np.repeat(np.arange(256, dtype=np.uint8), counts)
    data = np.empty(total, dtype=np.uint8)
    chunk = 1 << 26
    for a in range(0, total, chunk):
        b = min(total, a + chunk)
        np.take(table, rng.integers(0, 65536, size=b - a, dtype=np.uint16), out=data[a:b])
    return data, offsets


# ----------------------------------------------------------------- named workloads ------

def workload_plain(seed: int = 0):
    """Plain NFA drop-in: 1024-state NFA, 64 inputs x 4096 uniform bytes."""
    return plain_nfa(seed=seed), uniform_batch(64, 4096, seed=seed)


def workload_dfa(num_states: int = 4096, count: int = 4096, length: int = 65536,
                 latching: bool = True, seed: int = 0):
    """DFA memory-bound face: random DFA, ``count`` x ``length`` uniform bytes."""
    dfa = dense_dfa(num_states, 0.10 if latching else 0.0, seed=seed)
    return dfa, uniform_batch(count, length, seed=seed)


def workload_ragged(count: int = 16384, num_patterns: int = 2000, seed: int = 0,
                    accepting: bool = True):
    """Ragged batched NFA: 2000-pattern union (lengths 5-40, widths 1-32), 16384 inputs
    with log-normal lengths (median 4 KiB, sigma 1.0, clipped to 64 B - 1 MiB), 70% of
    bytes from pattern classes."""
    pu = pattern_union(num_patterns, (5, 40), (1, 32), seed=seed, accepting=accepting)
    lengths = lognormal_lengths(count, 4096, 1.0, 64, 1 << 20, seed=seed + 1)
    return pu, pattern_inputs(pu, lengths, 0.70, seed=seed + 2)


def workload_divergent(count: int = 1024, length: int = 1 << 20, seed: int = 0,
                       accepting: bool = True):
    """High-divergence NFA: 65536-state union of 8000 patterns (lengths 5-12 adjusted to
    the state count, widths 16-64), ``count`` x ``length`` inputs, 95% pattern bytes."""
    pu = pattern_union(8000, (5, 12), (16, 64), seed=seed, total_states=65536,
                       accepting=accepting)
    lengths = np.full(count, length, dtype=np.int64)
    return pu, pattern_inputs(pu, lengths, 0.95, seed=seed + 2)


def serving_automata(seed: int = 0) -> tuple[dict, dict]:
    """The serving mix: 4 random DFAs (256, 1024, 4096, 16384 states; accept probability
    1e-4, so most requests are scanned to the end) and 4 pattern-union NFAs built as in
    the ragged workload (about 1K, 4K, 16K and 64K states). Returns (automata by name,
    pattern unions by name)."""
    automata: dict = {}
    unions: dict = {}
    for k, n in enumerate((256, 1024, 4096, 16384)):
        automata[f"dfa{n}"] = dense_dfa(n, 1e-4, seed=seed + k)
    for k, n in enumerate((1024, 4096, 16384, 65536)):
        patterns = max(1, round(n / 22.5))
        pu = pattern_union(patterns, (5, 40), (1, 32), seed=seed + 10 + k, total_states=n)
        automata[f"nfa{n}"] = pu.nfa
        unions[f"nfa{n}"] = pu
    return automata, unions


def serving_requests(automata: dict, unions: dict, count: int, seed: int = 0,
                     median: int = 16384, sigma: float = 1.0,
                     max_len: int = 4 << 20) -> list[tuple[str, bytes]]:
    """``count`` requests: uniform automaton choice, log-normal lengths (median 16 KiB,
    clipped to 4 MiB); NFA requests draw 70% of bytes from the union's classes."""
    rng = np.random.default_rng(seed)
    names = sorted(automata)
    lengths = lognormal_lengths(count, median, sigma, 1, max_len, seed=seed + 1)
    out = []
    for k in range(count):
        name = names[int(rng.integers(0, len(names)))]
        n = int(lengths[k])
        if name in unions:
            d, _ = pattern_inputs(unions[name], np.array([n], dtype=np.int64), 0.70,
                                  seed=seed * 100003 + k)
        else:
            d = rng.integers(0, 256, size=n, dtype=np.uint8)
        out.append((name, d.tobytes()))
    return out
