"""ComputeContactAtom: per-atom contact count, the coordination number
(reference: computes/contact_atom.py, LAMMPS compute contact/atom).

One device thread per atom counts the partners in its full neighbor-list row
that overlap it at the current positions; the full list holds every pair from
both sides, so no atomics are needed for the two-sided tally.
"""

from __future__ import annotations

import numpy as np

from ..fields import HostField


class ComputeContactAtom:
    def __init__(self, max_atoms: int = 2000000) -> None:
        self.max_atoms = int(max_atoms)
        self.contact_count = HostField(self.max_atoms, np.int32)

    def compute(self, atom, domain, neighbor) -> np.ndarray:
        """int32[nlocal]: partners j with |d|^2 < (ri + rj)^2 and |d|^2 > 1e-18
        (contact_atom.py:19-64)."""
        n = atom.nlocal
        if n == 0:
            return np.zeros(0, dtype=np.int32)
        if n > self.max_atoms:
            raise ValueError(f"nlocal {n} exceeds ComputeContactAtom max_atoms {self.max_atoms}")
        atom._bind_domain(domain)
        counts, _ = atom._native.contact_atom(n)
        self.contact_count.from_numpy(counts)
        return counts.copy()
