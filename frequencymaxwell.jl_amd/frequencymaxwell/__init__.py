"""FrequencyMaxwell on MI355X: the reference's solve API over a HIP CBS engine.

The Python layer mirrors FrequencyMaxwell.jl's exports
(src/FrequencyMaxwell.jl:59-111); every numerical step of the solve runs in
``lib/libcbs_mi355x.so`` (hand-written HIP for gfx950).
"""

from ._lib import CBSError, Context, default_context, load as load_library
from .boundary_conditions import (AbsorbingBoundaryCondition, AbstractBoundaryCondition,
                                  AttenuationProfile, ExponentialProfile,
                                  PeriodicBoundaryCondition, TanhProfile, TukeyProfile,
                                  padding_pixels, requires_averaging, requires_padding,
                                  subpixel_shift)
from .engine import (CBSProblem, CBSServer, DeviceField, GreenSpec, make_opts, request_footprint,
                     solve_batch as solve_problems)
from .fields import (ElectromagneticField, crop_to_ROI, extract_plane, field_energy,
                     field_intensity, poynting_vector)
from .medium import Medium, grid_size, permittivity, permittivity_bg, phantom_bead
from .solver import (CBSLinearOperator, CBSPreconditioner, ConvergentBornSolver, DyadicGreen,
                     compute_magnetic_field, compute_ROI, conv, domain_size, grid_spacing,
                     h_scale, reset, solve, solve_batch)
from .sources import PlaneWaveSource, generate_incident_field, source_wavelength

GPU_BACKENDS = {"mi355x": CBSProblem}


def is_backend_available(device) -> bool:
    """is_backend_available (FrequencyMaxwell.jl:150-152)."""
    return str(device).lstrip(":").lower() in GPU_BACKENDS


__all__ = [
    "AbsorbingBoundaryCondition", "AbstractBoundaryCondition", "AttenuationProfile",
    "CBSError", "CBSLinearOperator", "CBSPreconditioner", "CBSProblem", "CBSServer", "Context",
    "DeviceField",
    "ConvergentBornSolver", "DyadicGreen", "ElectromagneticField", "ExponentialProfile",
    "GreenSpec", "Medium", "PeriodicBoundaryCondition", "PlaneWaveSource", "TanhProfile",
    "TukeyProfile", "compute_ROI", "compute_magnetic_field", "conv", "crop_to_ROI",
    "default_context", "domain_size", "extract_plane", "field_energy", "field_intensity", "generate_incident_field", "grid_size", "grid_spacing",
    "h_scale", "is_backend_available", "load_library", "make_opts", "padding_pixels",
    "permittivity", "permittivity_bg", "phantom_bead", "poynting_vector", "requires_averaging", "requires_padding",
    "request_footprint", "reset", "solve", "solve_batch", "solve_problems", "source_wavelength", "subpixel_shift",
]
