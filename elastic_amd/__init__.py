"""MI355X-native (gfx950) 2D elastic reverse-time migration.

The propagation, adjoint and imaging kernels are hand-written HIP in
``csrc/`` behind the C ABI of ``include/rtmlib_modC.h`` (librtm_hip.so);
``lib`` is its ctypes binding and ``synth`` builds the seeded workloads.
"""
from . import lib, synth  # noqa: F401

__all__ = ["lib", "synth"]
