"""ComputeStressAtom: per-atom stress tensor (reference: computes/stress_atom.py,
LAMMPS compute stress/atom).

The per-atom stress is the negated per-atom pair virial, minus m v_a v_b when
the kinetic term is asked for, in pressure*volume units (not divided by a
per-atom volume). The macro stress is its sum over atoms over the box volume,
the negative of the pressure tensor; the sum is a pairwise tree on the device.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..constants import f64


class ComputeStressAtom:
    """Per-atom stress tensor (6 components, [xx, yy, zz, xy, xz, yz])."""

    def __init__(self, max_atoms: int, float_type: Any = f64) -> None:
        self.float_type = float_type
        self.max_atoms = int(max_atoms)
        self.stress = np.zeros((self.max_atoms, 6), dtype=np.float64)
        self.macro_stress = np.zeros(6, dtype=np.float64)

    def _run(self, atom, domain, kinetic: bool, per_atom: bool):
        n = atom.nlocal
        if n > self.max_atoms:
            raise ValueError(f"nlocal {n} exceeds ComputeStressAtom max_atoms {self.max_atoms}")
        atom._bind_domain(domain)
        st, tot = atom._native.stress_atom(n, keflag=kinetic, per_atom=per_atom)
        if per_atom:
            self.stress[:n] = st
        self.macro_stress = tot
        return st, tot

    def compute(self, atom, domain, kinetic: bool = False) -> np.ndarray:
        """(nlocal, 6) per-atom stress (stress_atom.py:54-79). kinetic=False
        is `compute stress/atom NULL pair`: the pair virial only."""
        if atom.nlocal == 0:
            return np.zeros((0, 6), dtype=np.float64)
        st, _ = self._run(atom, domain, kinetic, True)
        return st.copy()

    def compute_macro(self, atom, domain, kinetic: bool = False) -> np.ndarray:
        """Sum of the per-atom stress over the box volume (stress_atom.py:81-103):
        the negative of the pressure tensor."""
        if atom.nlocal == 0:
            return np.zeros(6, dtype=np.float64)
        atom._bind_domain(domain)
        domain.pull()
        vol = domain.volume
        if vol <= 0.0:
            return np.zeros(6, dtype=np.float64)
        _, tot = self._run(atom, domain, kinetic, False)
        return tot / vol
