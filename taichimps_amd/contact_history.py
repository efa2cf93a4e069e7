"""Contact (shear) history and its carry-over across neighbor rebuilds
(reference: contact_history.py).

On the device each full-list slot holds a touching flag (the reference's
``partner == tag[j]``) and the shear of its pair; both slots of a pair hold the
same values. ``save_state`` snapshots the list, ``restore_state`` merges the
rebuilt list with the snapshot by partner (contact_history.py:91-139). The
``partner`` / ``shear`` views present the reference's flat half-list layout.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .constants import f64


class _PairField:
    def __init__(self, hist, kind: str):
        self._hist = hist
        self._kind = kind

    def to_numpy(self) -> np.ndarray:
        h = self._hist
        cap = h.max_pairs
        if self._kind == "partner":
            out = np.full(cap, -1, dtype=np.int32)
        else:
            out = np.zeros((cap, 3))
        if h._atom is None or self._kind == "rolling":
            return out
        _, _, partner, shear = h._atom._native.neigh_pairs()
        m = len(partner)
        if self._kind == "partner":
            out[:m] = partner
        else:
            out[:m] = shear
        return out

    def from_numpy(self, arr) -> None:
        h = self._hist
        if h._atom is None or self._kind == "rolling":
            return
        N = h._atom._native
        _, _, partner, shear = N.neigh_pairs()
        m = len(partner)
        a = np.asarray(arr)
        if self._kind == "partner":
            partner = a[:m].astype(np.int32)
        else:
            shear = a[:m].astype(np.float64)
        N.history_set_pairs(partner, shear)

    def fill(self, value) -> None:
        shape = (self._hist.max_pairs,) if self._kind == "partner" else (self._hist.max_pairs, 3)
        self.from_numpy(np.full(shape, value))

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class ContactHistory:
    """Tangential shear displacement per contact, keyed to neighbor-list slots."""

    def __init__(self, max_atoms: int, max_neighbors: int = 64, float_type: Any = f64) -> None:
        self.max_atoms = max_atoms
        self.max_neighbors = max_neighbors
        self.float_type = float_type
        self.max_pairs = max_atoms * max_neighbors
        self.partner = _PairField(self, "partner")
        self.shear = _PairField(self, "shear")
        self.rolling = _PairField(self, "rolling")
        self.setup_pending = False
        self._atom = None

    def _bind(self, atom) -> None:
        self._atom = atom

    def reset(self) -> None:
        """All slots not touching (contact_history.py:82-89)."""
        if self._atom is not None:
            self._atom._native.history_reset()

    def save_state(self, atom, nlist: Any) -> None:
        """Snapshot the touching contacts before a rebuild (:141-145)."""
        if atom.nlocal == 0 or self.setup_pending:
            return
        self._bind(atom)
        atom._native.history_save()

    def restore_state(self, atom, nlist: Any) -> None:
        """Carry the snapshot onto the rebuilt list (:147-154)."""
        if atom.nlocal == 0:
            return
        self._bind(atom)
        atom._native.history_restore()
        self.setup_pending = False

    def touching_by_atom(self, nlist: Any, nlocal: int) -> list[list[tuple[int, np.ndarray]]]:
        """The touching contacts of each atom as (partner tag, shear), :156-174."""
        out: list[list[tuple[int, np.ndarray]]] = [[] for _ in range(nlocal)]
        if self._atom is None:
            return out
        pi, _, partner, shear = self._atom._native.neigh_pairs()
        for nc in np.nonzero(partner >= 0)[0]:
            i = int(pi[nc])
            if i < nlocal:
                out[i].append((int(partner[nc]), shear[nc]))
        return out

    def load_restart(self, partner_tags, values) -> None:
        """Restart seeding (contact_history.py:176-214) is outside the device path."""
        raise NotImplementedError("restart files are not part of the device DEM step")
