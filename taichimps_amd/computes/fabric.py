"""ComputeFabric: contact-normal fabric tensor (reference: computes/fabric.py,
LAMMPS compute fabric).

phi_ab = (1 / Nc) * sum over contacts of n_a n_b, n the unit vector from the
higher-index atom to the lower one. The sums run on the device as per-atom
leaves (each atom's partners j > i, ascending) in a pairwise tree over atoms,
so the result does not depend on thread scheduling; the reference adds with
float atomics in an unspecified order.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..constants import f64


class ComputeFabric:
    def __init__(self, float_type: Any = f64) -> None:
        self.float_type = float_type
        self.phi = np.zeros(6, dtype=np.float64)  # xx, yy, zz, xy, xz, yz
        self.num_contacts = 0

    def compute(self, atom, domain, neighbor) -> tuple[np.ndarray, int]:
        """(phi[6], Nc); zeros and 0 without contacts (fabric.py:26-97)."""
        if atom.nlocal == 0:
            return np.zeros(6, dtype=np.float64), 0
        atom._bind_domain(domain)
        phi, nc = atom._native.fabric()
        self.phi = phi
        self.num_contacts = nc
        return phi.copy(), nc
