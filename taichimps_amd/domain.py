"""Simulation box and periodic boundaries (reference: domain.py).

The box lives twice, as in the reference: a host mirror the host code reads,
and the device copy the kernels read (inside every AtomSystem bound to this
Domain). A device-side box update (the deform/pressure servo) leaves the host
mirror stale until ``pull()``; reading it before then raises, exactly as the
reference's ownership guard does (domain.py:93-167).

Only orthogonal boxes are on this build's device path; a non-zero tilt raises
NotImplementedError.
"""

from __future__ import annotations

import weakref
from typing import Any

import numpy as np

from .constants import f64


class Domain:
    """Box [boxlo, boxhi] with per-axis boundary 'p' (periodic) or 'f' (fixed)."""

    def __init__(
        self,
        boxlo: list[float] | tuple[float, float, float],
        boxhi: list[float] | tuple[float, float, float],
        boundary: list[str] | tuple[str, str, str] = ("p", "p", "p"),
        tilt: list[float] | tuple[float, float, float] | None = None,
        float_type: Any = f64,
    ) -> None:
        self.float_type = float_type
        self._boxlo = np.array([float(v) for v in boxlo], dtype=np.float64)
        self._boxhi = np.array([float(v) for v in boxhi], dtype=np.float64)
        self._prd = self._boxhi - self._boxlo
        self._periodicity = np.array([1 if b == "p" else 0 for b in boundary], dtype=np.int32)
        self._tilt = np.zeros(3, dtype=np.float64)
        if tilt is not None:
            self._set_tilt(tilt)
        self._h_rate = np.zeros(6, dtype=np.float64)
        self.vremap = False
        self._host_stale = False
        self._auth = None
        self._atoms: list[weakref.ref] = []

    # ------------------------------------------------------------ binding
    def _attach(self, atom) -> None:
        for r in self._atoms:
            if r() is atom:
                return
        self._atoms.append(weakref.ref(atom))
        if self._host_stale:
            self.pull()
        self._push(atom)

    def _bound_atoms(self):
        out = []
        for r in self._atoms:
            a = r()
            if a is not None:
                out.append(a)
        return out

    def _push(self, atom) -> None:
        atom._native.set_box(self._boxlo, self._boxhi, self._periodicity, self._h_rate, self.vremap)

    def _sync(self) -> None:
        """Publish the host box to every bound system (domain.py:235-254)."""
        self._host_stale = False
        for a in self._bound_atoms():
            self._push(a)

    def _set_tilt(self, tilt) -> None:
        t = np.array([float(v) for v in tilt], dtype=np.float64)
        if np.any(t != 0.0):
            raise NotImplementedError("triclinic boxes are not on the device DEM path")
        self._tilt = t

    # ------------------------------------------------- host/device ownership
    def mark_device_authoritative(self, atom=None) -> None:
        """A kernel changed the box; the host mirror is out of date (domain.py:93-102)."""
        self._host_stale = True
        if atom is not None:
            self._auth = weakref.ref(atom)

    def pull(self) -> None:
        """Bring the host mirror in step with the device (domain.py:104-128)."""
        if not self._host_stale:
            return
        atom = self._auth() if self._auth is not None else None
        if atom is None:
            bound = self._bound_atoms()
            atom = bound[0] if bound else None
        if atom is not None:
            lo, hi, hr, vr = atom._native.get_box()
            self._boxlo, self._boxhi, self._h_rate = lo, hi, hr
            self._prd = self._boxhi - self._boxlo
            self.vremap = vr
            for a in self._bound_atoms():
                if a is not atom:
                    self._push(a)
        self._host_stale = False

    def _fresh(self, name: str) -> None:
        if self._host_stale:
            raise RuntimeError(
                f"Domain.{name} was read while the device holds the current box. "
                "Call domain.pull() first, at a point where synchronising with "
                "the device is acceptable -- not on the per-step path."
            )

    @property
    def boxlo(self) -> np.ndarray:
        self._fresh("boxlo")
        return self._boxlo

    @property
    def boxhi(self) -> np.ndarray:
        self._fresh("boxhi")
        return self._boxhi

    @property
    def prd(self) -> np.ndarray:
        self._fresh("prd")
        return self._prd

    @property
    def tilt(self) -> np.ndarray:
        self._fresh("tilt")
        return self._tilt

    @property
    def h_rate(self) -> np.ndarray:
        self._fresh("h_rate")
        return self._h_rate

    @property
    def periodicity(self) -> np.ndarray:
        return self._periodicity

    @property
    def triclinic(self) -> bool:
        return bool(np.any(self.tilt != 0.0))

    @property
    def xy(self) -> float:
        return float(self.tilt[0])

    @property
    def xz(self) -> float:
        return float(self.tilt[1])

    @property
    def yz(self) -> float:
        return float(self.tilt[2])

    @property
    def volume(self) -> float:
        prd = self.prd
        return float(prd[0] * prd[1] * prd[2])

    # ------------------------------------------------------------ setters
    def set_box(self, boxlo, boxhi, tilt=None) -> None:
        self._boxlo = np.array([float(v) for v in boxlo], dtype=np.float64)
        self._boxhi = np.array([float(v) for v in boxhi], dtype=np.float64)
        self._prd = self._boxhi - self._boxlo
        if tilt is not None:
            self._set_tilt(tilt)
        self._sync()

    def set_tilt(self, tilt) -> None:
        self._set_tilt(tilt)
        self._sync()

    def set_box_and_h_rate(self, boxlo, boxhi, tilt, h_rate, vremap: bool) -> None:
        self._boxlo = np.array([float(v) for v in boxlo], dtype=np.float64)
        self._boxhi = np.array([float(v) for v in boxhi], dtype=np.float64)
        self._prd = self._boxhi - self._boxlo
        if tilt is not None:
            self._set_tilt(tilt)
        self._h_rate = np.asarray(h_rate, dtype=np.float64)
        self.vremap = bool(vremap)
        self._sync()

    def set_h_rate(self, h_rate, vremap: bool) -> None:
        self._h_rate = np.asarray(h_rate, dtype=np.float64)
        self.vremap = bool(vremap)
        self._sync()

    def set_boundary(self, boundary) -> None:
        self._periodicity = np.array([1 if b == "p" else 0 for b in boundary], dtype=np.int32)
        self._sync()

    def bounding_box(self) -> tuple[np.ndarray, np.ndarray]:
        return self._boxlo.copy(), self._boxhi.copy()

    @property
    def pbc_x(self) -> bool:
        return bool(self._periodicity[0])

    @property
    def pbc_y(self) -> bool:
        return bool(self._periodicity[1])

    @property
    def pbc_z(self) -> bool:
        return bool(self._periodicity[2])

    # -------------------------------------------------------------- device
    def pbc(self, atom) -> None:
        """Wrap positions into the box (+ `remap v` velocity shift), domain.py:573-584."""
        atom._bind_domain(self)
        if atom.nlocal == 0:
            return
        atom._native.pbc()
