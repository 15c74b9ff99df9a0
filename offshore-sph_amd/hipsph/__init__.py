"""hipsph — WCSPH step engine for AMD MI355X (gfx950).

Host-side mirror of the KoningJasper/Offshore-SPH plugin surface
(Solver / Method / Kernels / Integrators / Helpers / Common) over the
hand-written HIP kernels of libhipsph.so (C ABI: include/hipsph.h).
"""
from .Common import ParticleType, particle_dtype, computed_dtype  # noqa: F401
