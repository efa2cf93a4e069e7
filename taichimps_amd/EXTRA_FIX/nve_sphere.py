"""fix nve/sphere: velocity Verlet for translation and rotation, I = 0.4 m r^2
(reference: EXTRA_FIX/nve_sphere.py). No PBC wrap per step (:68-72)."""

from __future__ import annotations

from typing import Any

from ..constants import f64
from .base import Fix


class FixNVESphere(Fix):
    def __init__(self, domain, float_type: Any = f64) -> None:
        super().__init__(domain, float_type=float_type)

    def initial_integrate(self, atom, dt: float) -> None:
        """v += dtf f/m; x += dt v; omega += dtf torque/I (nve_sphere.py:43-76)."""
        if atom.nlocal == 0:
            return
        atom._native.nve_initial(dt)

    def final_integrate(self, atom, dt: float) -> None:
        """v += dtf f/m; omega += dtf torque/I (nve_sphere.py:78-96)."""
        if atom.nlocal == 0:
            return
        atom._native.nve_final(dt)
