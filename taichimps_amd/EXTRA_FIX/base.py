"""Fix hook protocol (reference: EXTRA_FIX/base.py:15-57)."""

from __future__ import annotations

from abc import ABC
from typing import Any

from ..constants import f64


class Fix(ABC):
    def __init__(self, domain=None, float_type: Any = f64) -> None:
        self.domain = domain
        self.float_type = float_type
        self.history_update = True
        self.force_reneighbor = False
        self.timestep: int = 0

    def setup(self, nsteps_total: int = 0, dt: float = 0.0) -> None:
        """Called once at the start of a run (LAMMPS Fix::init)."""

    def initial_integrate(self, atom, dt: float) -> None:
        """First half of velocity Verlet."""

    def post_force(self, atom, dt: float) -> None:
        """After the pair forces."""

    def final_integrate(self, atom, dt: float) -> None:
        """Second half of velocity Verlet."""

    def end_of_step(self, atom, dt: float) -> None:
        """End of the timestep."""

    def _atoms(self):
        """The AtomSystem this fix's domain is bound to (most recent first)."""
        if self.domain is None:
            return None
        bound = self.domain._bound_atoms()
        return bound[-1] if bound else None
