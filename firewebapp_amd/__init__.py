"""firewebapp_amd — MI355X (gfx950) implementation of the FireWebApp Tier-3
``spread_3d`` time step.

* ``firewebapp_amd.physics_3d``: drop-in modules with the names and call
  signatures of the reference's ``model_outdoor.physics_3d`` package; every
  kernel runs in libfirewave.so (HIP) through the C ABI of
  ``include/firewave.h``.
* ``firewebapp_amd.spread_3d``: the GPU time-step driver (``run_3d_spread``).
* ``firewebapp_amd.build``: builds libfirewave.so in-tree.
"""
__all__ = ["physics_3d"]
