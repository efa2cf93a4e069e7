"""fix deform/pressure: per-axis strain or pressure control of the box
(reference: EXTRA_FIX/deform_pressure.py).

The x/y/z styles none, erate, trate, vel, wiggle, final, delta, scale,
pressure and pressure/mean, with couple, max/rate, normalize/pressure, nevery
and remap x / v / none, run entirely on the device: the virial(+kinetic)
reduction, the servo update, the affine `remap x` and the box publish happen
without a host round trip. The reference's host-only forms (tilt factors,
variable, volume, box, vol/balance/p) raise NotImplementedError.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass
from typing import Any

from .. import _native
from ..constants import f64
from .base import Fix

IX, IY, IZ, IYZ, IXZ, IXY = 0, 1, 2, 3, 4, 5
DIM_INDEX = {"x": IX, "y": IY, "z": IZ}
TILT_INDEX = {"yz": IYZ, "xz": IXZ, "xy": IXY}
STRAIN_STYLES = ("erate", "trate", "vel", "wiggle", "variable", "final", "scale", "delta")
PRESSURE_STYLES = ("pressure", "pressure/mean")
DIM_STYLES = ("none", "volume", *STRAIN_STYLES, *PRESSURE_STYLES)
TILT_STYLES = ("none", "erate", "trate", "vel", "wiggle", "variable", "final", "delta", "pressure")
BOX_STYLES = ("none", "volume", "pressure")
COUPLE_CHOICES = ("none", "xyz", "xy", "yz", "xz")
DEVICE_STYLE_CODE = {"none": 0, "erate": 1, "trate": 2, "vel": 3, "wiggle": 4, "final": 5,
                     "delta": 6, "scale": 7, "pressure": 8, "pressure/mean": 9}
COUPLE_CODE = {"none": 0, "xyz": 1, "xy": 2, "yz": 3, "xz": 4}
REMAP_CODE = {"x": 0, "v": 1, "none": 2}
NEEDS_RUN_LENGTH = ("final", "scale", "delta")


@dataclass
class AxisSet:
    """Per-dimension control settings, LAMMPS `Set` (deform_pressure.py:90-132)."""

    style: str = "none"
    rate: float = 0.0
    vel: float = 0.0
    scale: float = 1.0
    flo: float = 0.0
    fhi: float = 0.0
    dlo: float = 0.0
    dhi: float = 0.0
    ftilt: float = 0.0
    dtilt: float = 0.0
    amplitude: float = 0.0
    tperiod: float = 0.0
    hstr: str = ""
    hratestr: str = ""
    ptarget: float = 0.0
    pgain: float = 0.0
    coupled: bool = False


class FixDeformPressure(Fix):
    def __init__(
        self,
        domain,
        axes: dict[str, dict[str, Any]] | None = None,
        box: dict[str, Any] | None = None,
        couple: str = "none",
        max_rate: float = 0.0,
        normalize_pressure: bool = False,
        vol_balance_p: bool = False,
        remap: str = "x",
        nevery: int = 1,
        flip: bool = True,
        var_eval: Callable[[str], float] | None = None,
        float_type: Any = f64,
        p_target: float | None = None,
        p_gain: float | None = None,
    ) -> None:
        super().__init__(domain, float_type)
        self.nevery = int(nevery)
        if self.nevery < 1:
            raise ValueError("fix deform/pressure nevery must be >= 1")
        if couple not in COUPLE_CHOICES:
            raise ValueError(f"couple must be one of {COUPLE_CHOICES}, got {couple!r}")
        if remap not in ("x", "v", "none"):
            raise ValueError("remap must be 'x', 'v' or 'none'")
        self.couple = couple
        self.remap = remap
        self.flip = bool(flip)
        self.max_rate = float(max_rate)
        self.normalize_pressure = bool(normalize_pressure)
        self.vol_balance_p = bool(vol_balance_p)
        self.var_eval = var_eval
        self.sets = [AxisSet() for _ in range(6)]
        self.set_box = AxisSet()
        if axes is None and box is None:
            # legacy isotropic form (deform_pressure.py:190-199)
            target = 50000.0 if p_target is None else float(p_target)
            gain = 1e-4 if p_gain is None else float(p_gain)
            for s in self.sets[:3]:
                s.style = "pressure"
                s.ptarget = target
                s.pgain = gain
            if couple == "none":
                self.couple = "xyz"
        else:
            self._configure(axes or {}, box)
        self._validate()
        self.on_device = True
        self.device_reason = ""
        self._step_count = 0
        self._nsteps = 0
        self.nsteps_total = 0
        self._started = False
        self._atom = None

    # ------------------------------------------------------------ settings
    def _configure(self, axes, box) -> None:
        for name, spec in axes.items():
            if name in DIM_INDEX:
                idx, allowed = DIM_INDEX[name], DIM_STYLES
            elif name in TILT_INDEX:
                idx, allowed = TILT_INDEX[name], TILT_STYLES
            else:
                raise ValueError(f"Unknown deform dimension {name!r}")
            style = spec["style"]
            if style == "erate/rescale":
                raise ValueError("fix deform erate/rescale exists for NEMD with fix nvt/sllod, "
                                 "which taichimps does not have")
            if style not in allowed:
                raise ValueError(f"Unsupported fix deform/pressure style {style!r} for {name}; "
                                 f"implemented: {allowed}")
            s = self.sets[idx]
            s.style = style
            for key, value in spec.items():
                if key == "style":
                    continue
                if not hasattr(s, key):
                    raise ValueError(f"Unknown parameter {key!r} for dimension {name}")
                setattr(s, key, value if isinstance(value, str) else float(value))
        if box is not None:
            if box["style"] not in BOX_STYLES:
                raise ValueError(f"Unsupported fix deform/pressure box style {box['style']!r}; "
                                 f"implemented: {BOX_STYLES}")
            self.set_box.style = box["style"]
            for key, value in box.items():
                if key != "style":
                    setattr(self.set_box, key, float(value))

    def _validate(self) -> None:
        """deform_pressure.py:284-356, then the device-path limits."""
        names = ["x", "y", "z", "yz", "xz", "xy"]
        for i, s in enumerate(self.sets):
            if s.style in PRESSURE_STYLES and s.pgain <= 0.0:
                raise ValueError(f"fix deform/pressure gain for {names[i]} must be positive")
            if s.style == "variable" and not (s.hstr and s.hratestr):
                raise ValueError(f"fix deform/pressure variable style for {names[i]} needs two "
                                 "variable names")
            if s.style == "wiggle" and s.tperiod <= 0.0:
                raise ValueError(f"fix deform/pressure wiggle for {names[i]} needs a period")
        has_pressure = any(s.style in PRESSURE_STYLES for s in self.sets)
        has_pressure = has_pressure or self.set_box.style == "pressure"
        if self.normalize_pressure and not has_pressure:
            raise ValueError("normalize/pressure only applies to pressure-controlled dimensions")
        if self.max_rate < 0.0:
            raise ValueError("max/rate must be positive")
        if self.normalize_pressure and self.max_rate == 0.0:
            for s in self.sets[:3]:
                if s.style in PRESSURE_STYLES and s.ptarget == 0.0:
                    raise ValueError("cannot normalize the error for a zero target pressure "
                                     "without a max/rate")
        if self.couple != "none":
            idx = [i for i, c in enumerate("xyz") if c in self.couple]
            ref = next((i for i in idx if self.sets[i].style in PRESSURE_STYLES), None)
            if ref is None:
                raise ValueError("couple needs at least one of the coupled dimensions to be "
                                 "pressure controlled")
            for i in idx:
                s = self.sets[i]
                s.coupled = True
                if s.style == "none":
                    s.style = self.sets[ref].style
                    s.ptarget = self.sets[ref].ptarget
                    s.pgain = self.sets[ref].pgain
                elif s.style not in PRESSURE_STYLES:
                    raise ValueError("cannot couple non-pressure-controlled dimensions")
                elif (s.pgain != self.sets[ref].pgain or s.ptarget != self.sets[ref].ptarget):
                    raise ValueError("coupled dimensions must have identical target and gain")
        for i, s in enumerate(self.sets):
            if i >= 3 and s.style != "none":
                raise NotImplementedError("tilt-factor control is not on the device DEM path")
            if s.style in ("variable", "volume"):
                raise NotImplementedError(f"the {s.style} style is not on the device DEM path")
        if self.set_box.style != "none":
            raise NotImplementedError("the box style is not on the device DEM path")
        if self.vol_balance_p:
            raise NotImplementedError("vol/balance/p is not on the device DEM path")

    def params(self) -> _native.ServoParams:
        p = _native.ServoParams()
        p.enabled = 1
        p.nevery = self.nevery
        p.couple = COUPLE_CODE[self.couple]
        p.remap = REMAP_CODE[self.remap]
        p.normalize = 1 if self.normalize_pressure else 0
        p.max_rate = self.max_rate
        for i in range(3):
            s = self.sets[i]
            p.style[i] = DEVICE_STYLE_CODE[s.style]
            p.rate[i], p.vel[i], p.scale[i] = s.rate, s.vel, s.scale
            p.flo[i], p.fhi[i], p.dlo[i], p.dhi[i] = s.flo, s.fhi, s.dlo, s.dhi
            p.amp[i], p.tper[i] = s.amplitude, s.tperiod
            p.ptarget[i], p.pgain[i] = s.ptarget, s.pgain
        return p

    # ------------------------------------------------------------- counters
    def _sync_counters(self) -> None:
        if self._atom is not None:
            c = self._atom._native.counters()
            self._step_count = c["servo_step_count"]
            self._nsteps = c["servo_nsteps"]

    @property
    def step_count(self) -> int:
        self._sync_counters()
        return self._step_count

    @property
    def nsteps(self) -> int:
        self._sync_counters()
        return self._nsteps

    # ------------------------------------------------------------------ run
    def _bind(self, atom) -> None:
        if atom is None:
            return
        atom._bind_domain(self.domain)
        if self._atom is not atom:
            self._atom = atom
            atom._native.servo_set(self.params())
            # a fresh fix counts its own steps (deform_pressure.py:235-237)
            atom._native.set_counters(servo_step_count=0, servo_nsteps=0)

    def note_atom_count(self, nlocal: int) -> None:
        """Kept for the reference's Simulation protocol (deform_pressure.py:388-390)."""

    def setup(self, nsteps_total: int = 0, dt: float = 0.0) -> None:
        """Latch the box at the start of a run (deform_pressure.py:392-445)."""
        self.nsteps_total = int(nsteps_total)
        for s in self.sets[:3]:
            if s.style in NEEDS_RUN_LENGTH and self.nsteps_total <= 0:
                raise ValueError(f"fix deform/pressure style {s.style!r} needs the run length; "
                                 "call setup(nsteps_total) before running")
        atom = self._atom or self._atoms()
        if atom is None:
            return
        self._bind(atom)
        self.domain.pull()
        N = atom._native
        N.servo_set(self.params())
        N.servo_setup(self.nsteps_total, dt)
        self._started = True
        self.domain.mark_device_authoritative(atom)

    def end_of_step(self, atom, dt: float) -> None:
        """One servo update on the device (deform_pressure.py:1113-1219)."""
        if self.domain is None:
            return
        self._bind(atom)
        if not self._started:
            self.setup(0, dt)
        atom._native.servo_end_of_step(dt)
        self.domain.mark_device_authoritative(atom)
