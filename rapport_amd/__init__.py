"""rapport_amd: MI355X (gfx950) split-step filamentation propagator for fused silica.

Drop-in for the reference solver's Python entry points (``run``, ``Config``,
``build_grids``, ``Integrator``, operators) with the per-step arithmetic in
hand-written HIP kernels behind the C ABI of ``include/rapport_amd.h``.
"""
from .filament_sim import *  # noqa: F401,F403
from .filament_sim import __all__  # noqa: F401
