"""Kinetic energy, virial and pressure tensor (reference: computes/thermo.py).

The virial and kinetic sums run on the device as a pairwise tree over atoms
(per-atom leaves, adjacent pairs combined level by level), which makes them
deterministic; the reference's own order is a strided chunk sum.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..constants import f64

BOLTZMANN = 1.3806504e-23  # LAMMPS units si


class Computes:
    def __init__(self, float_type: Any = f64) -> None:
        self.float_type = float_type
        self.virial_tensor = np.zeros(6)

    def _ke(self, atom):
        if atom.nlocal == 0:
            return 0.0, 0.0
        return atom._native.ke()

    def ke_trans(self, atom) -> float:
        """sum 0.5 m v.v (thermo.py:46-60)."""
        return self._ke(atom)[0]

    def ke_rot(self, atom) -> float:
        """sum 0.5 (0.4 m r^2) omega.omega."""
        return self._ke(atom)[1]

    def ke_total(self, atom) -> float:
        a, b = self._ke(atom)
        return float(a + b)

    def compute_pressure_tensor(self, atom, domain, kinetic: bool = True) -> np.ndarray:
        """[Pxx, Pyy, Pzz, Pxy, Pxz, Pyz] = (virial + m v v) / V (thermo.py:141-170)."""
        if atom.nlocal == 0:
            return np.zeros(6, dtype=np.float64)
        atom._bind_domain(domain)
        domain.pull()
        vol = domain.volume
        if vol <= 0.0:
            return np.zeros(6, dtype=np.float64)
        self.virial_tensor = atom._native.virial(kinetic)
        return self.virial_tensor / vol

    def kinetic_tensor(self, atom) -> np.ndarray:
        """sum m v_a v_b, no factor 1/2 (thermo.py:180-201)."""
        if atom.nlocal == 0:
            return np.zeros(6, dtype=np.float64)
        n = atom.nlocal
        v = atom.v.to_numpy()[:n]
        m = atom.rmass.to_numpy()[:n]
        mv = m[:, None] * v
        return np.array([
            float(np.sum(mv[:, 0] * v[:, 0])), float(np.sum(mv[:, 1] * v[:, 1])),
            float(np.sum(mv[:, 2] * v[:, 2])), float(np.sum(mv[:, 0] * v[:, 1])),
            float(np.sum(mv[:, 0] * v[:, 2])), float(np.sum(mv[:, 1] * v[:, 2])),
        ])

    def temperature(self, atom) -> float:
        """sum m v^2 / (dof kB), dof = 3N - 3 (thermo.py:203-209)."""
        if atom.nlocal == 0:
            return 0.0
        dof = 3.0 * atom.nlocal - 3.0
        if dof <= 0.0:
            return 0.0
        return float(np.sum(self.kinetic_tensor(atom)[:3]) / (dof * BOLTZMANN))

    def compute_coordination_number(self, atom, neighbor) -> np.ndarray:
        """Neighbor-list counts per atom, as the reference returns them (thermo.py:211-215)."""
        if atom.nlocal == 0:
            return np.zeros(0, dtype=np.int32)
        return neighbor.num_neighbors.to_numpy()[: atom.nlocal]
