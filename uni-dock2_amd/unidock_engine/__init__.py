"""MI355X (gfx950) engine for the Uni-Dock2 docking hot path.

Drop-in for ``unidock_engine.api.python.pipeline`` of Uni-Dock2; the search,
scoring, optimisation and clustering run as HIP kernels in libud2.so behind a
C ABI (include/ud2.h).
"""
__all__ = ["topology", "synthetic"]
