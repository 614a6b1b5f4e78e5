"""CPU oracle for the gpufsm HIP backend. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this package, and there only as the checker of the HIP results. Nothing in the
``gpufsm`` package imports it; the HIP path fails loudly when its library is missing and
never falls back to these functions.

Contents and the upstream code each restates:

* ``reference.py``  ``epsilon_closure`` / ``simulate`` / ``simulate_dfa``
                    (`src/gpufsm/reference.py:20-81`); pinned by the vendored golden
                    corpus, subset construction, DFA-as-NFA embedding and pattern-union
                    brute force (``tests/test_oracle.py``).
* ``bitmap.py``     packed-bitset spec ``BitmapProgram`` (`src/gpufsm/core/bitmap.py:20-120`);
                    pinned by the golden corpus.
* ``fsm_oracle.c``  the same semantics in C over a thread pool, with examined-byte counts
  (``native.py``)   (scan and label-indexed modes); pinned against ``reference.py`` on the
                    golden corpus and on random automata.
* ``cpu_backend.py`` registers the above as ``Backend.CPU`` techniques ``reference`` and
                    ``bitmap`` (`src/gpufsm/backends/cpu.py:28-67`) for tests.
* ``gate.py``       the oracle gate ``compare`` / ``matches`` / ``require``
                    (`src/gpufsm/bench/oracle.py:14-80`).

Parity status of every function is recorded in ``tests/test_oracle.py``; none is unpinned.
"""
