"""Computes (reference: computes/): global thermodynamic sums and the
per-contact analysis computes."""

from .contact_atom import ComputeContactAtom
from .fabric import ComputeFabric
from .stress_atom import ComputeStressAtom
from .thermo import BOLTZMANN, Computes

__all__ = ["BOLTZMANN", "ComputeContactAtom", "ComputeFabric", "ComputeStressAtom", "Computes"]
