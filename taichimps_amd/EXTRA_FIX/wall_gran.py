"""fix wall/gran: frictional contact with a flat, axis-aligned wall
(reference: EXTRA_FIX/wall_gran.py, LAMMPS fix_wall_gran.cpp).

The wall is a contact partner of zero radius and infinite mass: radsum =
Reff = r_i, meff = m_i, partner velocity = the wall's, and the moment arm is
the full radius. `hooke/history` is a hooke normal force with mass-velocity
damping and the linear_history_classic tangential force with one shear
history per atom; `history=False` is the history-free `hooke` style. The
contact runs on the device as one thread per atom (tmps_wall_post_force); the
wall's position and velocity for the step (`wiggle`, `shear`) are computed
here, as the reference's post_force does.
"""

from __future__ import annotations

import math
from typing import Any

import numpy as np

from .. import _native
from ..constants import f64
from .base import Fix


class _HistoryView:
    def __init__(self, fix, which):
        self._fix = fix
        self._which = which

    def to_numpy(self) -> np.ndarray:
        fix = self._fix
        n = fix.max_atoms
        if fix._wall is None:
            return np.zeros((n, 3)) if self._which == 0 else np.zeros(n, dtype=np.int32)
        m = min(n, fix._wall_cap)
        sh, tc = fix._wall.history(m)
        if self._which == 0:
            out = np.zeros((n, 3))
            out[:m] = sh
        else:
            out = np.zeros(n, dtype=np.int32)
            out[:m] = tc
        return out

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class FixWallGran(Fix):
    """Granular contact with one axis-aligned plane wall (wall_gran.py:35-111).

    wall_axis 0/1/2; wall_side -1 for a wall below the particles (normal
    +axis), +1 above them; kn, gamman, kt, gammat, xmu as the pair styles
    (dampflag 0 zeroes gammat); shear_axis / shear_vel give an in-plane wall
    velocity, wiggle_amplitude / wiggle_period an oscillation along the
    normal."""

    def __init__(
        self,
        domain,
        wall_axis: int,
        wall_side: int,
        wall_coord: float,
        kn: float,
        gamman: float,
        kt: float,
        gammat: float,
        xmu: float,
        dampflag: int = 1,
        history: bool = True,
        limit_damping: bool = False,
        shear_axis: int = -1,
        shear_vel: float = 0.0,
        wiggle_amplitude: float = 0.0,
        wiggle_period: float = 0.0,
        max_atoms: int = 100000,
        float_type: Any = f64,
    ) -> None:
        super().__init__(domain, float_type=float_type)
        if wall_axis not in (0, 1, 2):
            raise ValueError("wall_axis must be 0, 1 or 2")
        if wall_side not in (-1, 1):
            raise ValueError("wall_side must be -1 (low) or +1 (high)")
        if shear_axis == wall_axis:
            raise ValueError("shear must act in the plane of the wall, not along its normal")
        if wiggle_amplitude != 0.0 and shear_vel != 0.0:
            raise ValueError("cannot wiggle and shear the same wall, as in LAMMPS")
        if wiggle_amplitude != 0.0 and wiggle_period <= 0.0:
            raise ValueError("wiggle needs a positive period")
        self.wall_axis = int(wall_axis)
        self.wall_side = int(wall_side)
        self.wall_coord = float(wall_coord)
        self.kn = float(kn)
        self.gamman = float(gamman)
        self.kt = float(kt)
        self.dampflag = int(dampflag)
        self.gammat = 0.0 if self.dampflag == 0 else float(gammat)
        self.xmu = float(xmu)
        self.limit_damping = 1 if limit_damping else 0
        self.use_history = 1 if history else 0
        self.shear_axis = int(shear_axis)
        self.shear_vel = float(shear_vel)
        self.wiggle_amplitude = float(wiggle_amplitude)
        self.wiggle_period = float(wiggle_period)
        self.max_atoms = int(max_atoms)
        self.elapsed = 0.0
        self._wall = None
        self._wall_sys = None
        self._wall_cap = 0
        self.shear = _HistoryView(self, 0)
        self.touch = _HistoryView(self, 1)

    def _params(self) -> _native.WallParams:
        p = _native.WallParams()
        p.axis, p.side = self.wall_axis, self.wall_side
        p.use_history, p.limit_damping = self.use_history, self.limit_damping
        p.shear_axis = self.shear_axis
        p.kn, p.kt, p.gamman, p.gammat, p.xmu = (self.kn, self.kt, self.gamman, self.gammat,
                                                 self.xmu)
        return p

    def _bind(self, atom):
        if self._wall is None or self._wall_sys is not atom._native:
            if atom.nlocal > self.max_atoms:
                raise ValueError(f"nlocal {atom.nlocal} exceeds FixWallGran max_atoms "
                                 f"{self.max_atoms}")
            self._wall = _native.NativeWall(atom._native, self._params())
            self._wall_sys = atom._native
            self._wall_cap = atom.max_atoms
        return self._wall

    def reset_history(self, nlocal: int) -> None:
        """All atoms out of contact (wall_gran.py:113-118)."""
        if self._wall is not None:
            self._wall.reset()

    def post_force(self, atom, dt: float) -> None:
        """Wall position and velocity for this step, then the contact
        (wall_gran.py:213-236)."""
        if atom.nlocal == 0:
            return
        coord = self.wall_coord
        vwall_n = 0.0
        if self.wiggle_amplitude != 0.0:
            omega_w = 2.0 * math.pi / self.wiggle_period
            arg = omega_w * self.elapsed
            coord += self.wiggle_amplitude * (1.0 - math.cos(arg))
            vwall_n = self.wiggle_amplitude * omega_w * math.sin(arg)
        self._bind(atom).post_force(dt, coord, vwall_n, self.shear_vel,
                                    1 if self.history_update else 0)
        self.elapsed += dt

    def setup(self, nsteps_total: int = 0, dt: float = 0.0) -> None:
        """Restart the clock the wall motion is referenced to."""
        self.elapsed = 0.0

    def _fused(self, atom):
        """(device wall, position, in-plane velocity) for a device-resident
        run's step, or None for a wiggling plane (its position follows the
        clock on the host)."""
        if self.wiggle_amplitude != 0.0:
            return None
        return self._bind(atom), self.wall_coord, self.shear_vel

    def _advance(self, calls: int, dt: float) -> None:
        """The clock after `calls` post_force calls made on the device."""
        self.elapsed += calls * dt


class FixWallGranRegion(FixWallGran):
    """fix wall/gran/region with a cylinder region (reference:
    EXTRA_FIX/wall_gran_region.py, LAMMPS fix_wall_gran_region.cpp).

    The curved side of a cylinder along `axis` (0/1/2 or "x"/"y"/"z") with
    cross-section centre (c1, c2) and radius `radius`; side "in" (or -1) holds
    the atoms inside, "out" (or +1) outside. `axis_lo` / `axis_hi` bound the
    atoms considered (they are not end caps). `fstyle` hooke / hertz with or
    without `/history`; kt NULL = 2/7 kn, gammat NULL = gamman / 2; the hertz
    styles scale stiffness and damping by the contact radius sqrt(r delta).
    One shear history per atom (the side is a single surface)."""

    def __init__(
        self,
        domain,
        axis: int | str = 2,
        c1: float = 0.0,
        c2: float = 0.0,
        radius: float = 1.0,
        side: str | int = "in",
        axis_lo: float | None = None,
        axis_hi: float | None = None,
        fstyle: str = "hooke",
        kn: float = 1e5,
        gamman: float = 0.0,
        kt: float | None = None,
        gammat: float | None = None,
        xmu: float = 0.0,
        dampflag: int = 1,
        history: bool | None = None,
        limit_damping: bool = False,
        max_atoms: int = 100000,
        float_type: Any = f64,
    ) -> None:
        Fix.__init__(self, domain, float_type=float_type)
        if isinstance(axis, str):
            axis = {"x": 0, "y": 1, "z": 2}[axis.lower()]
        self.axis = int(axis)
        self.c1, self.c2, self.radius = float(c1), float(c2), float(radius)
        if isinstance(side, str):
            self.side = -1 if side.lower() in ("in", "inside", "-1") else 1
        else:
            self.side = -1 if side < 0 else 1
        self.axis_lo = float(axis_lo) if axis_lo is not None else -1e30
        self.axis_hi = float(axis_hi) if axis_hi is not None else 1e30
        self.has_axis_bounds = 1 if (axis_lo is not None or axis_hi is not None) else 0
        self.fstyle = fstyle.lower()
        self.is_hertz = 1 if "hertz" in self.fstyle else 0
        self.kn = float(kn)
        self.gamman = float(gamman)
        self.kt = float(kt) if kt is not None else (2.0 / 7.0 * self.kn)
        self.dampflag = int(dampflag)
        gammat_val = float(gammat) if gammat is not None else (0.5 * self.gamman)
        self.gammat = 0.0 if self.dampflag == 0 else gammat_val
        self.xmu = float(xmu)
        self.limit_damping = 1 if limit_damping else 0
        self.use_history = (1 if (self.fstyle.endswith("history") if history is None
                                  else history) else 0)
        # FixWallGran's attributes a cylinder does not use
        self.wall_axis, self.wall_side, self.wall_coord = self.axis, self.side, 0.0
        self.shear_axis, self.shear_vel = -1, 0.0
        self.wiggle_amplitude = self.wiggle_period = 0.0
        self.max_atoms = int(max_atoms)
        self.elapsed = 0.0
        self._wall = None
        self._wall_sys = None
        self._wall_cap = 0
        self.shear = _HistoryView(self, 0)
        self.touch = _HistoryView(self, 1)

    def _params(self) -> _native.WallParams:
        p = _native.WallParams()
        p.geometry = 1
        p.axis, p.side = self.axis, self.side
        p.use_history, p.limit_damping = self.use_history, self.limit_damping
        p.shear_axis = -1
        p.kn, p.kt, p.gamman, p.gammat, p.xmu = (self.kn, self.kt, self.gamman, self.gammat,
                                                 self.xmu)
        p.hertz, p.has_bounds = self.is_hertz, self.has_axis_bounds
        p.c1, p.c2, p.radius = self.c1, self.c2, self.radius
        p.axis_lo, p.axis_hi = self.axis_lo, self.axis_hi
        return p

    def post_force(self, atom, dt: float) -> None:
        """The cylinder contact on every atom (wall_gran_region.py:273-290)."""
        if atom.nlocal == 0:
            return
        self._bind(atom).post_force(dt, 0.0, 0.0, 0.0, 1 if self.history_update else 0)
