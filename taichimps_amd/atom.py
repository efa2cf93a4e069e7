"""AtomSystem: the per-atom state (reference: atom.py), held in device memory
by libtmps_hip.so."""

from __future__ import annotations

from typing import Any

import numpy as np

from . import _native
from .constants import f64, precision_code
from .fields import DeviceField, HostField


class AtomSystem:
    """
    Per-atom fields of atom_style sphere: x, v, f, omega, torque, radius, rmass,
    tag (device), atom_type, mask (host), and the per-atom pair virial
    [xx, yy, zz, xy, xz, yz].
    """

    def __init__(self, max_atoms: int, float_type: Any = f64) -> None:
        self.max_atoms: int = int(max_atoms)
        self.nlocal: int = 0
        self.float_type = float_type
        self.precision = precision_code(float_type)
        self._native = _native.NativeSystem(self.max_atoms, 64, self.precision)
        self._domain = None

        self.x = DeviceField(self, _native.F_X, 3)
        self.v = DeviceField(self, _native.F_V, 3)
        self.f = DeviceField(self, _native.F_F, 3)
        self.omega = DeviceField(self, _native.F_OMEGA, 3)
        self.torque = DeviceField(self, _native.F_TORQUE, 3)
        self.radius = DeviceField(self, _native.F_RADIUS, 1)
        self.rmass = DeviceField(self, _native.F_RMASS, 1)
        self.tag = DeviceField(self, _native.F_TAG, 1, dtype=np.int32)
        self.virial = DeviceField(self, _native.F_VIRIAL, 6)
        self.atom_type = HostField(self.max_atoms, np.int32)
        self.mask = HostField(self.max_atoms, np.int32)

    # -------------------------------------------------------------- binding
    def _bind_domain(self, domain) -> None:
        if domain is None or self._domain is domain:
            return
        self._domain = domain
        domain._attach(self)

    # ------------------------------------------------------------ particles
    def add_particles(
        self,
        x: Any,
        radius: Any,
        density: float | Any,
        v: Any | None = None,
        omega: Any | None = None,
        tag: Any | None = None,
        atom_type: int | Any = 1,
    ) -> None:
        """
        Add particles; rmass = (4/3) pi r^3 density (atom.py:55-145).
        """
        x_np = np.asarray(x, dtype=np.float64)
        if x_np.ndim == 1:
            x_np = x_np.reshape(1, 3)
        n_new = len(x_np)
        if self.nlocal + n_new > self.max_atoms:
            raise ValueError(
                f"Exceeded max_atoms capacity: {self.nlocal + n_new} > {self.max_atoms}"
            )
        start, end = self.nlocal, self.nlocal + n_new
        v_np = np.zeros_like(x_np) if v is None else np.asarray(v, dtype=np.float64).reshape(-1, 3)
        omega_np = (np.zeros_like(x_np) if omega is None
                    else np.asarray(omega, dtype=np.float64).reshape(-1, 3))
        rad_np = np.asarray(radius, dtype=np.float64)
        if rad_np.ndim == 0:
            rad_np = np.full(n_new, rad_np)
        if np.isscalar(density):
            dens_np = np.full(n_new, density, dtype=np.float64)
        else:
            dens_np = np.asarray(density, dtype=np.float64)
        if np.isscalar(atom_type):
            type_np = np.full(n_new, atom_type, dtype=np.int32)
        else:
            type_np = np.asarray(atom_type, dtype=np.int32)
        vol = (4.0 / 3.0) * np.pi * (rad_np**3)
        mass_np = vol * dens_np
        tag_np = (np.arange(start + 1, end + 1, dtype=np.int32) if tag is None
                  else np.asarray(tag, dtype=np.int32))

        N = self._native

        def merged(fid, new):
            if start == 0:
                return new
            return np.concatenate([N.get_field(fid, start), new])

        # radius and mass first: the packed device layouts carry them
        N.set_field(_native.F_RADIUS, merged(_native.F_RADIUS, rad_np), end)
        N.set_field(_native.F_RMASS, merged(_native.F_RMASS, mass_np), end)
        N.set_field(_native.F_X, merged(_native.F_X, x_np), end)
        N.set_field(_native.F_V, merged(_native.F_V, v_np), end)
        N.set_field(_native.F_OMEGA, merged(_native.F_OMEGA, omega_np), end)
        N.set_field(_native.F_TAG, merged(_native.F_TAG, tag_np), end)
        self.atom_type._a[start:end] = type_np
        self.mask._a[start:end] = 1
        self.nlocal = end
        N.set_nlocal(end)

    def filter_particles(self, keep_mask: np.ndarray) -> int:
        """Compacting the arrays (atom.py:147-199) is outside this build's path."""
        raise NotImplementedError(
            "AtomSystem.filter_particles is not part of the device DEM step")

    def clear_forces(self) -> None:
        """Zero f, torque and the per-atom virial (atom.py:201-220)."""
        self._native.clear_forces()
