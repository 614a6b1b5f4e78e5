"""On-disk automaton formats (`src/gpufsm/io/`): ANML, Micron's homogeneous-automaton XML."""

from .anml import load_anml, parse_symbol_set

__all__ = ["load_anml", "parse_symbol_set"]
