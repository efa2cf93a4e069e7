"""fix damping/cundall: f_d *= 1 - gamma_lin sign(f_d v_d), same for torque and
omega (reference: EXTRA_FIX/damping_cundall.py)."""

from __future__ import annotations

from typing import Any

from ..constants import f64
from .base import Fix


class FixDampingCundall(Fix):
    def __init__(self, domain=None, gamma_lin: float = 0.3, gamma_ang: float = 0.3,
                 float_type: Any = f64) -> None:
        super().__init__(domain, float_type)
        self.gamma_lin = float(gamma_lin)
        self.gamma_ang = float(gamma_ang)

    def post_force(self, atom, dt: float) -> None:
        if atom.nlocal == 0:
            return
        atom._native.cundall(self.gamma_lin, self.gamma_ang)
