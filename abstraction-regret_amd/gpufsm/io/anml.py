"""ANML loader: homogeneous state-transition elements -> edge-labelled CSR ``NFA``.

Interface and semantics follow `src/gpufsm/io/anml.py:52-212` (``parse_symbol_set``,
``load_anml``):

* an STE matches a symbol set; ``activate-on-match`` edges connect STEs; an STE becomes
  active when a predecessor was active and the byte is in its set; ``report-on-match``
  STEs accept;
* ``start="all-input"`` STEs may match at every position, ``start="start-of-data"`` only at
  position 0;
* the conversion puts each STE's set on its incoming edges, behind three synthetic states:
  the NFA start ``root`` with epsilon edges to ``every`` (a wildcard self-loop, seeding the
  all-input STEs at each byte) and to ``first`` (no self-loop, seeding the start-of-data
  STEs once). State 0, 1, 2 are root, every, first; STE i is state 3 + i in file order.

Symbol sets: ``*`` (any byte: ``ANY_SYMBOL``), a single atom, or a bracketed class of atoms
and ``lo-hi`` ranges, optionally negated with ``^`` over 0..255 (``[^]`` is every byte).
Atoms are ``0xHH`` / ``\\xHH`` (exactly two hex digits, either case) or one literal
character. Anything else is refused, and so is every file construct that would change the
language if skipped: elements outside the STE subset (counters, boolean gates), STEs
without or with duplicate ids, activations of unknown elements. An activation target may
carry a port (``id:port``); the id names the element.

``fold_starts=True`` drops the synthetic states when the file has no start-of-data STE: the
start state is then ``every`` itself, with no epsilon edges, which lets chain-shaped
automata (pattern sets) use the chain and bitpacked techniques. Verdicts are the same.
"""

from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from pathlib import Path

from ..core.nfa import ANY_SYMBOL, NFA, NFABuilder

_HEX = re.compile(r"(?:0x|\\x)([0-9a-fA-F]{2})", re.IGNORECASE)
_KNOWN = frozenset({"anml", "automata-network", "description", "state-transition-element",
                    "activate-on-match", "report-on-match"})


def _atom(tok: str) -> int:
    m = _HEX.fullmatch(tok)
    if m:
        return int(m.group(1), 16)
    if len(tok) == 1:
        return ord(tok)
    raise ValueError(f"unsupported ANML symbol atom {tok!r}")


def _tokens(body: str, whole: str) -> list[str]:
    out, i = [], 0
    while i < len(body):
        if body[i:i + 2].lower() in ("0x", "\\x"):
            tok = body[i:i + 4]
            if not _HEX.fullmatch(tok):
                raise ValueError(f"malformed hex atom {tok!r} in ANML symbol-set {whole!r}")
            out.append(tok)
            i += 4
        else:
            out.append(body[i])
            i += 1
    return out


def parse_symbol_set(text: str) -> set[int]:
    """Bytes (0..255) of an ANML symbol-set attribute; ``{ANY_SYMBOL}`` for ``*``."""
    s = text.strip()
    if s == "*":
        return {ANY_SYMBOL}
    if not (s.startswith("[") and s.endswith("]")):
        return {_atom(s)}
    body = s[1:-1]
    if body == "":
        raise ValueError(f"empty ANML symbol-set {s!r}: the STE could never match")
    negated = body.startswith("^")
    toks = _tokens(body[1:] if negated else body, s)
    members: set[int] = set()
    k = 0
    while k < len(toks):
        if k + 2 < len(toks) and toks[k + 1] == "-":
            lo, hi = _atom(toks[k]), _atom(toks[k + 2])
            if hi < lo:
                raise ValueError(f"inverted ANML range in {s!r}")
            members.update(range(lo, hi + 1))
            k += 3
        else:
            members.add(_atom(toks[k]))
            k += 1
    return set(range(256)) - members if negated else members


def _tag(element) -> str:
    return element.tag.rsplit("}", 1)[-1]


def load_anml(path: str | Path, fold_starts: bool = False) -> NFA:
    """Load an ANML file (homogeneous-STE subset) into an edge-labelled ``NFA``."""
    root = ET.parse(path).getroot()  # noqa: S314 - local automaton files
    extra = sorted({_tag(e) for e in root.iter()} - _KNOWN)
    if extra:
        raise ValueError(f"{path}: unsupported ANML elements {extra} (only the homogeneous "
                         f"STE subset is supported; counters and boolean gates change the "
                         f"language)")
    stes = [e for e in root.iter() if _tag(e) == "state-transition-element"]
    if not stes:
        raise ValueError(f"no state-transition-element found in {path}")
    ids = [e.get("id") for e in stes]
    missing = sum(i is None for i in ids)
    if missing:
        raise ValueError(f"{path}: {missing} state-transition-element(s) without an 'id' "
                         f"attribute")
    seen: set[str] = set()
    dup = sorted({i for i in ids if i in seen or seen.add(i)})
    if dup:
        raise ValueError(f"{path}: duplicate state-transition-element ids {dup[:5]}")
    sets: dict[str, set[int]] = {}
    for e in stes:
        try:
            sets[e.get("id")] = parse_symbol_set(e.get("symbol-set", "*"))
        except ValueError as exc:
            raise ValueError(f"{path}: STE {e.get('id')!r}: {exc}") from None
    modes = {e.get("id"): (e.get("start") or "none").lower() for e in stes}
    fold = fold_starts and "start-of-data" not in modes.values()

    b = NFABuilder()
    if fold:
        every = b.add_state()
        first = None
        b.set_start(every)
    else:
        root_state, every, first = b.add_state(), b.add_state(), b.add_state()
        b.set_start(root_state)
        b.add_epsilon(root_state, every)
        b.add_epsilon(root_state, first)
    b.add_transition(every, ANY_SYMBOL, every)
    state = {i: b.add_state() for i in ids}

    def enter(target: str, src: int) -> None:
        for c in sets[target]:
            b.add_transition(src, c, state[target])

    for e in stes:
        sid = e.get("id")
        if modes[sid] == "all-input":
            enter(sid, every)
        elif modes[sid] == "start-of-data":
            enter(sid, first)
        for child in e:
            kind = _tag(child)
            if kind == "report-on-match":
                b.set_accept(state[sid], True)
            elif kind == "activate-on-match":
                target = child.get("element")
                if target is None:
                    raise ValueError(f"{path}: STE {sid!r}: activate-on-match without "
                                     f"'element'")
                if target not in state and ":" in target:
                    target = target.split(":", 1)[0]
                if target not in state:
                    raise ValueError(f"{path}: STE {sid!r} activates unknown element "
                                     f"{child.get('element')!r}")
                enter(target, state[sid])
    return b.build()
