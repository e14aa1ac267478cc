"""MI355X (gfx950) backend for trench01's per-batch charging hot path.

HIP kernels behind a C-ABI (libtrench_hip.so, include/trench_hip.h), with a
host-side mirror of the reference's tracer API (``HipTracer``,
``make_particle_tracer``) and batch loop.
"""
from .tracer import HipTracer, make_particle_tracer, trace, BACKENDS  # noqa: F401
