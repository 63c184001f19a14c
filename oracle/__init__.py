"""Oracle: CPU FP64 restatement of the reference split-step solver.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (``rapport_amd``)
imports, links or executes anything under ``oracle/``. The only callers are
``tests/``, ``__graft_entry__.smoke()`` (as the checker) and the
``cpu_baseline`` leg of ``bench.py``.

Modules
-------
``reference_numpy``    NumPy/SciPy FP64 restatement of ``D3/sim/{config,keldysh,
                       grids,operators,kernels,integrator,filament_sim}.py``
                       (``D3`` = ``.github/deliverable (3)``), each function citing
                       the reference file:line it follows.
``adjoint_numpy``      FP64 discrete adjoint of the oracle's forward step (the
                       reference has no gradients), pinned by finite differences.
``postprocess_numpy``  the 0-D carrier re-integration (``D3/sim/figures_article.py``)
                       and the 1-D SPM + shock model (``D3/sim/self_steepening.py``).

The reference path is Python over cupy (no C/C++ sources to compile), so the
oracle is NumPy/SciPy only.
"""
