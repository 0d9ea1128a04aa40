"""CPU oracle: an FP64 NumPy restatement of the FireWebApp Tier-3 step.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP path: it
may be imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` and by nothing else.  The product
(``firewebapp_amd``) never imports it and has no CPU fallback.

Each function restates one reference kernel (file:line cited in its
docstring, paths relative to the FireWebApp tree) with the reference's
per-element operation order, so that FP64 results agree with the golden
vectors (``tests/golden/kernel_vectors.json``) bit-for-bit for the pure
stencils and to ~1e-12 for kernels with transcendental calls.

Pinning status of every function is listed in ``oracle/PINNING.md``.
"""
