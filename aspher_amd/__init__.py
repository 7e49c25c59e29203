"""MI355X-native (gfx950) spectral contact path of ASPHER.

Public API (same names, signatures and defaults as the reference's pybind11
module ``aspher``): ContactSolver, solve_nested, ContactResult,
configure_allocator. The compute runs in libaspher_hip.so (hand-written HIP
kernels + C++17 host engine) reached through the C ABI of include/aspher_hip.h.
"""
from . import surfaces
from .api import (ContactResult, ContactSolver, clear_operator_cache, configure_allocator,
                  device_count, solve_ensemble, solve_ensemble_device, solve_nested,
                  solve_nested_device)
from .surfaces import rough_surface, westergaard_gap

__all__ = ["ContactResult", "ContactSolver", "solve_nested", "solve_nested_device",
           "solve_ensemble", "solve_ensemble_device", "configure_allocator",
           "clear_operator_cache", "device_count", "rough_surface", "westergaard_gap",
           "surfaces"]
