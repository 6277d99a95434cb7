"""relax_amd -- MI355X (gfx950) implementation of RELAX's per-time-step
elastic + relaxation solve.

Modules mirror the reference's Fortran modules on this path:
  fourier         fft3, fft2                          (fourier.f90)
  green           elasticresponse, surfacetractioncowling, cerrutimodified,
                  greenfunctioncowling                (green.f90)
  elastic3d       stressupdate, equivalentbodyforce, fieldadd, tensorfieldadd,
                  tensoramplitude, momentdensityshear (elastic3d.f90)
  viscoelastic3d  viscouseigenstress                  (viscoelastic3d.f90)
  friction3d      frictioneigenstress, monitorfriction, frictionadd
  model           Model: the device-resident RK2 loop (relax.f90:555-890)

Arrays are numpy float32 in the reference's memory order (C-order shapes
(sx3, sx2, sx1+2) for vector fields, (sx3, sx2, sx1, 6) for TENSOR fields).
Everything runs through librelax_hip.so; there is no CPU fallback.
"""
from ._lib import LAYER, PATCH, PLANE, SOURCE, TENSOR_LAYER, WEAK, RelaxError, load  # noqa: F401

__all__ = ["LAYER", "PATCH", "PLANE", "SOURCE", "TENSOR_LAYER", "WEAK", "RelaxError", "load"]
