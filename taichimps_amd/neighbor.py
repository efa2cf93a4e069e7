"""Cell-list neighbor list with the skin trigger (reference: neighbor.py).

The device keeps a full per-atom list (both i->j and j->i, rows ascending in
j) built by a counting sort over the reference's lamda-space grid (200 000
cell cap included). The reference's half list (j > i) is exposed through
``num_neighbors``, ``neighbors``, ``pair_i``/``pair_j`` and ``npairs``, in the
reference's flat layout, so other code reads the same arrays.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .atom import AtomSystem
from .constants import f64
from .domain import Domain
from .fields import ArrayView, ScalarView


class NeighborList:
    def __init__(
        self,
        domain: Domain,
        skin: float = 0.001,
        max_atoms: int = 100000,
        max_neighbors_per_atom: int = 64,
        max_neighbors: int | None = None,
        float_type: Any = f64,
    ) -> None:
        self.domain = domain
        self.skin = skin
        self.max_atoms = max_atoms
        self.max_neighbors_per_atom = (max_neighbors if max_neighbors is not None
                                       else max_neighbors_per_atom)
        self.float_type = float_type
        # neigh_modify defaults (neighbor.py:88-95)
        self.delay: int = 0
        self.every: int = 1
        self.check: bool = True
        self.last_build_step: int = -1
        self.current_step: int = 0
        self.has_built: bool = False
        self.build_count: int = 0
        self.max_grid_cells: int = 200000
        self._atom: AtomSystem | None = None

        self.npairs = ScalarView(lambda: self._native().neigh_info()[2] if self._atom else 0)
        self.num_neighbors = ArrayView(self._num_neighbors)
        self.neighbors = ArrayView(self._neighbors)
        self.pair_i = ArrayView(lambda: self._pairs()[0])
        self.pair_j = ArrayView(lambda: self._pairs()[1])
        self.pair_start = ArrayView(self._pair_start)
        self.grid_dim = ScalarView(lambda: self._native().neigh_info()[0])
        self.cell_size = ScalarView(lambda: self._native().neigh_info()[1])
        self.cell_of = ArrayView(lambda: self._native().neigh_bins(self._atom.nlocal)[0])
        self.cell_start = ArrayView(lambda: self._native().neigh_bins(self._atom.nlocal)[1])
        self.cell_particles = ArrayView(lambda: self._native().neigh_bins(self._atom.nlocal)[2])
        self.x0 = ArrayView(lambda: self._native().get_field(9, self._atom.max_atoms))

    # ------------------------------------------------------------ plumbing
    def _bind(self, atom: AtomSystem) -> None:
        atom._bind_domain(self.domain)
        self._atom = atom

    def _native(self):
        if self._atom is None:
            raise RuntimeError("the neighbor list has not been used with an AtomSystem yet")
        return self._atom._native

    def _push(self) -> None:
        N = self._native()
        N.neigh_config(self.skin, self.delay, self.every, self.check, self.max_grid_cells,
                       self.max_neighbors_per_atom)
        N.set_counters(has_built=self.has_built, build_count=self.build_count,
                       last_build_step=self.last_build_step, current_step=self.current_step)

    def _pull(self) -> None:
        c = self._native().counters()
        self.has_built = bool(c["has_built"])
        self.build_count = c["build_count"]
        self.last_build_step = c["last_build_step"]
        self.current_step = c["current_step"]

    def _pairs(self):
        if self._atom is None:
            return (np.zeros(0, np.int32),) * 2 + (np.zeros(0, np.int32), np.zeros((0, 3)))
        return self._native().neigh_pairs()

    def _num_neighbors(self):
        out = np.zeros(self.max_atoms, dtype=np.int32)
        if self._atom is not None and self._atom.nlocal:
            n = self._atom.nlocal
            out[:n] = self._native().neigh_counts(n)
        return out

    def _pair_start(self):
        counts = self._num_neighbors()
        out = np.zeros(self.max_atoms + 1, dtype=np.int32)
        out[1:] = np.cumsum(counts)
        return out

    def _neighbors(self):
        out = np.zeros((self.max_atoms, self.max_neighbors_per_atom), dtype=np.int32)
        pi, pj, _, _ = self._pairs()
        if len(pi):
            start = np.searchsorted(pi, pi, side="left")
            slot = np.arange(len(pi)) - start
            out[pi, slot] = pj
        return out

    # ------------------------------------------------------------------ API
    def decide(self, atom: AtomSystem, timestep: int | None = None) -> bool:
        """LAMMPS delay / every / check criteria (neighbor.py:371-403)."""
        self._bind(atom)
        self._push()
        rebuild = self._native().neigh_decide(timestep)
        self._pull()
        return rebuild

    def check_and_build(self, atom: AtomSystem, timestep: int | None = None) -> bool:
        if self.decide(atom, timestep):
            self.build(atom)
            return True
        return False

    def build(self, atom: AtomSystem) -> None:
        """Grid, bins, list, x0 and counters (neighbor.py:414-435)."""
        if atom.nlocal == 0:
            return
        self._bind(atom)
        self._push()
        self._native().neigh_build()
        self._pull()
