"""rapport_amd: MI355X (gfx950) split-step filamentation propagator for fused silica.

Drop-in for the reference solver's Python entry points (``run``, ``Config``,
``build_grids``, ``Integrator``, operators) with the per-step arithmetic in
hand-written HIP kernels behind the C ABI of ``include/rapport_amd.h``.

Submodules imported by name: ``batch`` (many instances per launch, HIP graph) with
``autotune``, ``adjoint`` (hand-written adjoint behind torch.autograd), ``reintegrate``
(0-D carrier re-integration of saved intensity histories) and ``shock`` (1-D SPM + shock
model).
"""
from .filament_sim import *  # noqa: F401,F403
from .filament_sim import __all__  # noqa: F401
