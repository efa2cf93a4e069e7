"""Inputs of the restart-file fixtures (G6): a small system as plain arrays,
wrapped in minimal objects with the attributes a restart writer reads
(AtomSystem fields, Domain box, NeighborList pair arrays, ContactHistory
partner/shear, PairGranular sub-models). Shared by make_restart_golden.py and
tests/test_restart.py so both writers see identical inputs."""

from __future__ import annotations

import numpy as np

N = 10
TIMESTEP = 1234
DT = 1.0e-6
BOXLO = (0.0, -1.0, 0.5)
BOXHI = (4.0, 3.5, 5.0)
# (i, j > i, touching) in the device's pair order: i ascending, j ascending
PAIRS = [(0, 1, True), (0, 3, True), (0, 6, False), (1, 2, False), (2, 5, True),
         (3, 4, False), (4, 9, True), (5, 6, False), (7, 8, True), (8, 9, False)]


class _F:
    def __init__(self, a):
        self._a = np.asarray(a)

    def to_numpy(self):
        return self._a.copy()

    def __getitem__(self, idx):
        # a 0-d field reads as field[None], a scalar
        return self._a[()] if idx is None and self._a.ndim == 0 else self._a[idx]


class Atom:
    def __init__(self, d):
        self.nlocal = N
        self.x, self.v, self.omega = _F(d["x"]), _F(d["v"]), _F(d["omega"])
        self.radius, self.rmass = _F(d["radius"]), _F(d["rmass"])
        self.tag, self.atom_type = _F(d["tag"]), _F(d["atom_type"])


class Domain:
    periodicity = np.array([1, 1, 0], dtype=np.int32)
    triclinic = False
    boxlo = np.array(BOXLO)
    boxhi = np.array(BOXHI)
    tilt = np.zeros(3)


class NeighborList:
    def __init__(self, d):
        self.npairs = _F(len(d["pair_i"]))
        self.pair_i, self.pair_j = _F(d["pair_i"]), _F(d["pair_j"])


class History:
    """partner (tag or -1) and shear per half-list pair."""

    def __init__(self, d):
        self.partner, self.shear = _F(d["partner"]), _F(d["shear"])

    def touching_by_atom(self, nlist, nlocal):
        out = [[] for _ in range(nlocal)]
        pi = nlist.pair_i.to_numpy()
        partner, shear = self.partner.to_numpy(), self.shear.to_numpy()
        for nc in np.nonzero(partner >= 0)[0]:
            if pi[nc] < nlocal:
                out[int(pi[nc])].append((int(partner[nc]), shear[nc]))
        return out

    def _bind(self, atom):
        pass


class Pair:
    limit_damping = 0
    submodels = {
        "normal": ("hertz/material", [1.0e7, 0.5, 0.3]),
        "damping": ("coeff_restitution", []),
        "tangential": ("mindlin", [-1.0, 1.0, 0.5]),
        "rolling": ("none", []),
        "twisting": ("none", []),
        "heat": ("none", []),
    }


def arrays() -> dict:
    rng = np.random.default_rng(606)
    d = {
        "x": rng.uniform(0.0, 4.0, size=(N, 3)),
        "v": rng.normal(0.0, 0.1, size=(N, 3)),
        "omega": rng.normal(0.0, 2.0, size=(N, 3)),
        "radius": rng.uniform(0.45, 0.55, size=N),
        "tag": (np.arange(N, dtype=np.int32) * 7 + 101)[::-1].copy(),
        "atom_type": np.array([1, 2] * (N // 2), dtype=np.int32),
    }
    d["rmass"] = 4.0 / 3.0 * np.pi * d["radius"] ** 3 * 2.5
    d["pair_i"] = np.array([p[0] for p in PAIRS], dtype=np.int32)
    d["pair_j"] = np.array([p[1] for p in PAIRS], dtype=np.int32)
    touching = np.array([p[2] for p in PAIRS])
    d["partner"] = np.where(touching, d["tag"][d["pair_j"]], -1).astype(np.int32)
    d["shear"] = np.where(touching[:, None], rng.normal(0.0, 1e-3, size=(len(PAIRS), 3)), 0.0)
    return d
