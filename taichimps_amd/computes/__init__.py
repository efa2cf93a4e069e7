"""Global thermodynamic computes (reference: computes/)."""

from .thermo import BOLTZMANN, Computes

__all__ = ["Computes", "BOLTZMANN"]
