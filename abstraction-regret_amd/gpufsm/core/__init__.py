"""Core data model: automata (CSR NFA, dense DFA), results, packing and the registry.

numpy only; nothing here touches the GPU. Mirrors `src/gpufsm/core/__init__.py`.
"""
