"""caustics_amd — MI355X-native microlensing magnification hot path.

Public API mirrors fbartolic/caustics (src/caustics/__init__.py:2-23) for the
magnification path: the root-solver operator and the lensing functions built on it.
All compute runs in hand-written HIP kernels for gfx950 (libcaustics_hip.so).
"""
from .ehrlich_aberth_primitive import ehrlich_aberth, poly_roots
from .extended_source import mag_extended_source
from .lightcurve import (hexadecapole_gate, hexadecapole_vjp, mag, mag_fwd_vjp,
                         mag_point_source_lightcurve)
from .linalg import marginalized_log_likelihood
from .point_source import (critical_and_caustic_curves, images_point_source,
                           images_point_source_sequential, lens_eq, lens_eq_det_jac,
                           mag_point_source)
from .trajectory import AnnualParallaxTrajectory

__all__ = [
    "ehrlich_aberth",
    "poly_roots",
    "lens_eq",
    "lens_eq_det_jac",
    "mag_point_source",
    "images_point_source",
    "images_point_source_sequential",
    "critical_and_caustic_curves",
    "mag_point_source_lightcurve",
    "mag_extended_source",
    "hexadecapole_gate",
    "hexadecapole_vjp",
    "mag",
    "mag_fwd_vjp",
    "marginalized_log_likelihood",
    "AnnualParallaxTrajectory",
]
