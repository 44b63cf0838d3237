"""psb_project_amd -- MI355X-native SPH / PCISPH timestep of PSB_Project.

The compute path is the HIP library libpsb_sph.so (csrc/, C-ABI in
include/psb_sph.h).  This package is its Python host side: ctypes bindings,
device buffers, seeded scenes and the simulator classes that mirror the
reference's System / SPHSystem / PCI_SPHSystem API.
"""
from ._lib import (EOS_LINEAR, EOS_TAIT, FLAG_NEIGHBOR_OVERFLOW, FLAG_NONFINITE,  # noqa: F401
                   FLAG_NOT_CONVERGED, FLAG_OUT_OF_GRID, FLAG_ZERO_DENSITY, MODE_CONFIG,
                   MODE_REFERENCE, SOLVER_PCISPH, SOLVER_WCSPH, WARM_MLP, WARM_NONE,
                   WARM_PREVIOUS, GridDesc, Neighbors, Params, PsbError, StepStats, check, lib)
from .device import Context, DeviceArray, device_count  # noqa: F401

__all__ = ["Context", "DeviceArray", "Params", "PsbError", "lib", "check", "device_count"]
