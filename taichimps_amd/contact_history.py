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
        self._seed = None  # restart seed waiting for the atoms

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
        self._push_seed(atom)
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
        """Seed the saved history from a restart file, keyed by partner tag
        (contact_history.py:176-214, FixNeighHistory::unpack_restart).

        Entry k of the lists belongs to local atom k (the order the atoms were
        given to add_particles). Each entry under atom i for partner j is in
        the i-to-j convention. As in the reference, a pair (i, j > i) takes
        the value stored under atom i, its lower-index owner. The first list
        build after this (setup_pending skips the save before it) carries the
        seeded contacts onto the new list.
        """
        n = min(len(partner_tags), self.max_atoms)
        rows = []
        for i in range(n):
            t = np.asarray(partner_tags[i], dtype=np.int64).reshape(-1)
            m = len(t)
            if m > self.max_neighbors:
                raise ValueError(
                    f"atom {i} has {m} stored contacts but the history only has "
                    f"{self.max_neighbors} slots per atom")
            v = (np.asarray(values[i], dtype=np.float64).reshape(m, -1)[:, :3] if m
                 else np.zeros((0, 3)))
            rows.append((t, v))
        self._seed = rows
        self.setup_pending = True
        if self._atom is not None and self._atom.nlocal > 0:
            self._push_seed(self._atom)

    def _push_seed(self, atom) -> None:
        """Hand a pending restart seed to the device as per-atom rows of
        partner indices (both atoms of each pair, the owner's value)."""
        rows = getattr(self, "_seed", None)
        if rows is None:
            return
        self._seed = None
        self._bind(atom)
        n = atom.nlocal
        tags = atom.tag.to_numpy()[:n].astype(np.int64)
        order = np.argsort(tags, kind="stable")
        owner, other, val = [], [], []
        for i, (t, v) in enumerate(rows[:n]):
            if len(t) == 0:
                continue
            pos = np.minimum(np.searchsorted(tags[order], t), n - 1)
            found = tags[order][pos] == t
            j = order[pos]
            own = found & (j > i)  # pairs whose lower-index atom is i
            owner.append(np.full(int(own.sum()), i, dtype=np.int64))
            other.append(j[own])
            val.append(v[own])
        if owner:
            a = np.concatenate(owner)
            b = np.concatenate(other)
            sv = np.concatenate(val)
        else:
            a = b = np.zeros(0, dtype=np.int64)
            sv = np.zeros((0, 3))
        # both rows of each pair hold the lower-index-first value
        ra, rb = np.concatenate([a, b]), np.concatenate([b, a])
        rv = np.concatenate([sv, sv])
        srt = np.lexsort((rb, ra))
        ra, rb, rv = ra[srt], rb[srt], rv[srt]
        if len(ra) > 1:
            dup = (ra[1:] == ra[:-1]) & (rb[1:] == rb[:-1])
            if np.any(dup):
                keep = np.concatenate([[True], ~dup])
                ra, rb, rv = ra[keep], rb[keep], rv[keep]
        start = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(ra, minlength=n), out=start[1:])
        if np.any(np.diff(start) > self.max_neighbors):
            raise ValueError("an atom has more restored contacts than the history has slots "
                             "per atom")
        atom._native.history_seed(start, rb, rv)
