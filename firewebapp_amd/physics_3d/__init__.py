"""Drop-in replacement of ``model_outdoor.physics_3d`` (hot-path subset).

Each module keeps the reference's function names, argument order, in-place
semantics and exception types; the work runs in HIP kernels.  NumPy float64
arguments run the FP64 kernels (parity mode); CUDA float32 tensors run the
FP32 kernels in place.
"""
