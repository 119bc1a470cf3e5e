"""MI355X-native device path for Octopus's real-space H-apply: finite-difference stencil,
Hamiltonian apply (kinetic + local + Kleinman-Bylander nonlocal) and Taylor propagator.

Host-side setup (mesh, stencil index, projector matrices, seeded workloads) is plain
numpy; every apply runs in liboctopus_hip.so through its C ABI (include/octhip.h).
"""
from . import rng, mesh, stencil, projectors, workloads  # noqa: F401

__all__ = ["rng", "mesh", "stencil", "projectors", "workloads"]
