"""taichimps_amd: the taichimps per-step DEM path on MI355X (gfx950).

The reference's Python API (Simulation, AtomSystem, Domain, NeighborList,
ContactHistory, the GRANULAR pair styles, the EXTRA_FIX fixes and Computes)
over hand-written HIP kernels in libtmps_hip.so, reached through its C ABI.
"""

from .atom import AtomSystem
from .computes import ComputeContactAtom, ComputeFabric, ComputeStressAtom, Computes
from .constants import f32split, f64
from .contact_history import ContactHistory
from .domain import Domain
from .EXTRA_FIX import (FixDampingCundall, FixDeformPressure, FixNVESphere, FixPrint,
                        FixWallGran, FixWallGranRegion)
from .GRANULAR import GranHertzHistory, GranHookeHistory, GranularPair, PairGranular
from .neighbor import NeighborList
from .simulation import Simulation

__all__ = [
    "AtomSystem", "ComputeContactAtom", "ComputeFabric", "ComputeStressAtom", "Computes",
    "ContactHistory", "Domain", "FixDampingCundall",
    "FixDeformPressure", "FixNVESphere", "FixPrint", "FixWallGran", "FixWallGranRegion", "GranHertzHistory", "GranHookeHistory",
    "GranularPair", "NeighborList", "PairGranular", "Simulation", "f32split", "f64",
]
