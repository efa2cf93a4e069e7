"""Fixes on the DEM step (reference: EXTRA_FIX/)."""

from .base import Fix
from .damping_cundall import FixDampingCundall
from .deform_pressure import FixDeformPressure
from .nve_sphere import FixNVESphere
from .print import FixPrint
from .wall_gran import FixWallGran, FixWallGranRegion

__all__ = ["Fix", "FixNVESphere", "FixDampingCundall", "FixDeformPressure", "FixWallGran",
           "FixWallGranRegion", "FixPrint"]
