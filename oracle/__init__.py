"""Oracle: CPU FP64 restatement of the reference split-step solver.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (``rapport_amd``)
imports, links or executes anything under ``oracle/``. The only callers are
``tests/``, ``__graft_entry__.smoke()`` (as the checker) and the
``cpu_baseline`` leg of ``bench.py``.

Modules
-------
``reference_numpy``  NumPy/SciPy FP64 restatement of ``D3/sim/{config,keldysh,
                     grids,operators,kernels,integrator,filament_sim}.py``
                     (``D3`` = ``.github/deliverable (3)``), each function citing
                     the reference file:line it follows.
``rate_kernel.c``    the same carrier rate-equation recurrence
                     (``D3/sim/kernels.py:13-106``) in plain C, built into
                     ``oracle/_build/librate_oracle.so`` by ``oracle/Makefile``;
                     used to time the CPU leg of the plasma pass and cross-check
                     the NumPy loop.
"""
