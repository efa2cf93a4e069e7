"""Array views with the interface of the reference's Taichi fields
(to_numpy / from_numpy / fill / indexing), backed by device memory owned by
libtmps_hip.so."""

from __future__ import annotations

import numpy as np


class DeviceField:
    """A per-atom field of an AtomSystem (atom.py:36-53)."""

    def __init__(self, atom, field_id: int, width: int, dtype=np.float64, writable=True):
        self._atom = atom
        self._fid = field_id
        self._width = width
        self._dtype = dtype
        self._writable = writable

    @property
    def shape(self):
        n = self._atom.max_atoms
        return (n,) if self._width == 1 else (n, self._width)

    def to_numpy(self) -> np.ndarray:
        return self._atom._native.get_field(self._fid, self._atom.max_atoms)

    def from_numpy(self, arr) -> None:
        if not self._writable:
            raise ValueError("this field is read-only")
        a = np.asarray(arr, dtype=self._dtype)
        count = min(len(a), self._atom.max_atoms)
        self._atom._native.set_field(self._fid, a[:count], count)

    def fill(self, value) -> None:
        self.from_numpy(np.full(self.shape, value, dtype=self._dtype))

    def __getitem__(self, idx):
        return self.to_numpy()[idx]

    def __setitem__(self, idx, value):
        a = self.to_numpy()
        a[idx] = value
        self.from_numpy(a)


class HostField:
    """A field the path never reads on the device (atom_type, mask)."""

    def __init__(self, n: int, dtype):
        self._a = np.zeros(n, dtype=dtype)

    @property
    def shape(self):
        return self._a.shape

    def to_numpy(self) -> np.ndarray:
        return self._a.copy()

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr, dtype=self._a.dtype)
        self._a[: len(a)] = a[: len(self._a)]

    def fill(self, value) -> None:
        self._a[:] = value

    def __getitem__(self, idx):
        return self._a[idx]

    def __setitem__(self, idx, value):
        self._a[idx] = value


class ScalarView:
    """A 0-d value read as ``field[None]`` (e.g. NeighborList.npairs)."""

    def __init__(self, getter):
        self._get = getter

    def __getitem__(self, idx):
        return self._get()

    def to_numpy(self):
        return np.asarray(self._get())


class ArrayView:
    """A read-only array computed on demand (neighbor-list and bin views)."""

    def __init__(self, getter):
        self._get = getter

    def to_numpy(self) -> np.ndarray:
        return np.asarray(self._get())

    def __getitem__(self, idx):
        return self.to_numpy()[idx]
