"""pair_style gran/hooke/history (reference: GRANULAR/hooke_history.py).

Hookean normal spring-dashpot, incremental tangential shear history with the
Coulomb cap; the per-contact law runs in the device gather kernel (csrc/contact.h).
"""

from __future__ import annotations

from typing import Any

from .. import _native
from ..constants import f64
from .base import GranularPair


class GranHookeHistory(GranularPair):
    """Constructor order kn, gamman, kt, gammat, xmu (hooke_history.py:48-68)."""

    STYLE = _native.PAIR_HOOKE_HISTORY

    def __init__(self, domain, kn: float, gamman: float, kt: float, gammat: float, xmu: float,
                 dampflag: int = 1, limit_damping: bool = False, float_type: Any = f64) -> None:
        super().__init__(domain, float_type=float_type)
        self.kn = float(kn)
        self.gamman = float(gamman)
        self.kt = float(kt)
        self.dampflag = int(dampflag)
        # LAMMPS: "if (dampflag == 0) gammat = 0.0"
        self.gammat = 0.0 if self.dampflag == 0 else float(gammat)
        self.xmu = float(xmu)
        self.limit_damping = 1 if limit_damping else 0

    def params(self) -> _native.PairParams:
        p = _native.PairParams()
        p.style = self.STYLE
        p.limit_damping = self.limit_damping
        p.kn, p.kt, p.gamman, p.gammat, p.xmu = self.kn, self.kt, self.gamman, self.gammat, self.xmu
        return p
