"""Host-side packing of inputs and accept sets.

Interface parity: `src/gpufsm/core/packing.py:21-98` (``WORD_BITS``, ``words_for``,
``symbols``, ``pack_inputs``, ``pack_accept``, ``accept_bitmask``). Those keep the
reference's int32 layout for callers that use them directly.

The HIP backend does not use the int32 form. It hands the device the bytes themselves
(``uint8``, four times less traffic) and ``int64`` offsets, so a batch may exceed 2 GiB:
:func:`pack_bytes`.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .nfa import NFA

WORD_BITS = 64
"""Bits per word of a packed state set."""


def words_for(num_states: int) -> int:
    """``ceil(num_states / 64)``."""
    return (num_states + WORD_BITS - 1) // WORD_BITS


def symbols(input_bytes: bytes) -> np.ndarray:
    """One input as an int32 array (reference layout)."""
    return np.frombuffer(input_bytes, dtype=np.uint8).astype(np.int32)


def pack_inputs(inputs: list[bytes]) -> tuple[np.ndarray, np.ndarray]:
    """Batch as int32 symbols plus int32 offsets (reference layout, capped at 2**31-1 bytes)."""
    total = sum(len(b) for b in inputs)
    limit = int(np.iinfo(np.int32).max)
    if total > limit:
        raise ValueError(
            f"batch of {total} bytes exceeds the int32 offset range ({limit}); "
            f"split it into smaller batches"
        )
    data, offsets = pack_bytes(inputs)
    return data.astype(np.int32), offsets.astype(np.int32)


def pack_bytes(inputs: list[bytes]) -> tuple[np.ndarray, np.ndarray]:
    """Batch as one ``uint8`` buffer plus ``int64`` offsets (``len(inputs) + 1`` entries).

    Input ``i`` is ``data[offsets[i]:offsets[i + 1]]``. An all-empty batch gives an empty
    buffer.
    """
    lengths = np.fromiter((len(b) for b in inputs), dtype=np.int64, count=len(inputs))
    offsets = np.zeros(len(inputs) + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    if offsets[-1] > 0:
        data = np.frombuffer(b"".join(inputs), dtype=np.uint8)
    else:
        data = np.zeros(0, dtype=np.uint8)
    return data, offsets


def pack_accept(
    nfa: NFA, dtype: type = np.uint64, nwords: int | None = None
) -> np.ndarray[Any, Any]:
    """Accept set as ``nwords`` packed words, state ``s`` at bit ``s % 64`` of word ``s // 64``."""
    if nwords is None:
        nwords = words_for(nfa.num_states)
    bits = np.zeros(nwords * WORD_BITS, dtype=np.uint8)
    acc = np.asarray(nfa.accept, dtype=bool)
    bits[: acc.size] = acc
    packed = np.packbits(bits, bitorder="little").view(np.uint64)
    return packed.astype(dtype) if dtype is not np.uint64 else packed.copy()


def accept_bitmask(nfa: NFA) -> int:
    """Accept set as one Python int (bit ``s`` = state ``s``)."""
    mask = 0
    for s in np.flatnonzero(np.asarray(nfa.accept, dtype=bool)):
        mask |= 1 << int(s)
    return mask
