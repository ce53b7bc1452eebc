"""CPU oracle for the light-field projection path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  It restates the reference's MATLAB algorithms in FP64
(``sid_oracle.py``, numpy) and in plain C (``sid_oracle.c``, built into
``liboracle.so`` by ``build_oracle.py``).
"""
