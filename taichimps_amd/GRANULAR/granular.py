"""pair_style granular with selectable sub-models (reference: GRANULAR/granular.py).

    normal      hooke | hertz | hertz/material
    damping     none | velocity | mass_velocity | viscoelastic | tsuji | coeff_restitution
    tangential  none | linear_nohistory | linear_history | mindlin

The constructor derives the kernel constants exactly as the reference does
(granular.py:122-269); the per-contact law runs in the device gather kernel.
Rolling and twisting resistance are not implemented (the reference raises on
them too, granular.py:18-20,202-204).
"""

from __future__ import annotations

import math
from typing import Any

from .. import _native
from ..constants import f64
from .base import GranularPair

NORMAL = {"hooke": 0, "hertz": 1, "hertz/material": 2}
DAMPING = {"none": 0, "velocity": 1, "mass_velocity": 2, "viscoelastic": 3, "tsuji": 4,
           "coeff_restitution": 5}
TANGENTIAL = {"none": 0, "linear_nohistory": 1, "linear_history": 2, "mindlin": 3}

_TWOROOTFIVEBYSIX = 1.82574185835055380345  # 2*sqrt(5/6), gran_sub_mod_damping.cpp
_ROOTTHREEBYTWO = 1.22474487139158894067  # sqrt(3/2)
_FOURTHIRDS = 4.0 / 3.0


def mix_stiffness_e(e1: float, e2: float, poiss1: float, poiss2: float) -> float:
    """GranSubMod::mix_stiffnessE (granular.py:94-96)."""
    return 1.0 / ((1 - poiss1 * poiss1) / e1 + (1 - poiss2 * poiss2) / e2)


def mix_stiffness_g(e1: float, e2: float, poiss1: float, poiss2: float) -> float:
    """GranSubMod::mix_stiffnessG (granular.py:99-103)."""
    return 1.0 / (2 * (2 - poiss1) * (1 + poiss1) / e1 + 2 * (2 - poiss2) * (1 + poiss2) / e2)


class PairGranular(GranularPair):
    def __init__(
        self,
        domain,
        normal: str,
        normal_coeffs: list[float],
        tangential: str = "none",
        tangential_coeffs: list[float | None] | None = None,
        damping: str = "viscoelastic",
        limit_damping: bool = False,
        float_type: Any = f64,
    ) -> None:
        super().__init__(domain, float_type=float_type)
        if normal not in NORMAL:
            raise ValueError(f"Unsupported pair granular normal model {normal!r}; "
                             f"implemented: {sorted(NORMAL)}")
        if damping not in DAMPING:
            raise ValueError(f"Unsupported pair granular damping model {damping!r}; "
                             f"implemented: {sorted(DAMPING)}")
        if tangential not in TANGENTIAL:
            raise ValueError(f"Unsupported pair granular tangential model {tangential!r}; "
                             f"implemented: {sorted(TANGENTIAL)}")
        self.normal_id = NORMAL[normal]
        self.damping_id = DAMPING[damping]
        self.tangential_id = TANGENTIAL[tangential]
        self.limit_damping = 1 if limit_damping else 0

        self.emod = 0.0
        self.poiss = 0.0
        if normal == "hertz/material":
            if len(normal_coeffs) != 3:
                raise ValueError("normal hertz/material takes E, damp, poisson")
            self.emod, cor, self.poiss = (float(c) for c in normal_coeffs)
            self.kn = _FOURTHIRDS * mix_stiffness_e(self.emod, self.emod, self.poiss, self.poiss)
        else:
            if len(normal_coeffs) != 2:
                raise ValueError(f"normal {normal} takes k, damp")
            self.kn, cor = (float(c) for c in normal_coeffs)
        self.normal_damp = float(cor)
        self.damp = self._damping_coefficient(normal)

        self.kt = 0.0
        self.xt = 0.0
        self.mu = 0.0
        self.uses_history = self.tangential_id in (2, 3)
        if self.tangential_id != 0:
            self._setup_tangential(tangential, tangential_coeffs or [])
        # what a restart file stores (GranularModel::write_restart): sub-model
        # names and raw coefficients, NULL as LAMMPS's -1 sentinel
        self.submodels: dict[str, tuple[str, list[float]]] = {
            "normal": (normal, [float(c) for c in normal_coeffs]),
            "damping": (damping, []),
            "tangential": (tangential, [-1.0 if c is None else float(c)
                                        for c in (tangential_coeffs or [])]),
            "rolling": ("none", []),
            "twisting": ("none", []),
            "heat": ("none", []),
        }

    def _damping_coefficient(self, normal: str) -> float:
        """The damping sub-model's constant (granular.py:207-238)."""
        if self.damping_id in (DAMPING["tsuji"], DAMPING["coeff_restitution"]):
            cor = self.normal_damp
            if not 0.0 < cor <= 1.0:
                raise ValueError(
                    "damping tsuji/coeff_restitution needs the normal model's damp coefficient "
                    f"to be a restitution coefficient in (0, 1]; got {cor}")
            logcor = math.log(cor)
            if self.damping_id == DAMPING["tsuji"]:
                d = (1.2728 - 4.2783 * cor + 11.087 * cor**2 - 22.348 * cor**3
                     + 27.467 * cor**4 - 18.022 * cor**5 + 4.8218 * cor**6)
                return d * math.sqrt(2.0)
            if normal == "hooke":
                return -2 * logcor / math.sqrt(math.pi**2 + logcor**2)
            return -_ROOTTHREEBYTWO * _TWOROOTFIVEBYSIX * logcor / math.sqrt(math.pi**2 + logcor**2)
        return self.normal_damp

    def _setup_tangential(self, tangential: str, coeffs) -> None:
        """Tangential constants (granular.py:240-269); NULL -> 8 G* for mindlin."""
        if self.tangential_id == TANGENTIAL["linear_nohistory"]:
            if len(coeffs) != 2 or any(c is None for c in coeffs):
                raise ValueError("tangential linear_nohistory takes xt, mu")
            self.xt, self.mu = (float(c) for c in coeffs)
            return
        if len(coeffs) != 3:
            raise ValueError(f"tangential {tangential} takes k (or NULL), xt, mu")
        k_raw, xt, mu = coeffs
        if xt is None or mu is None:
            raise ValueError(f"tangential {tangential}: xt and mu cannot be NULL")
        self.xt = float(xt)
        self.mu = float(mu)
        if k_raw is None:
            if self.normal_id != NORMAL["hertz/material"]:
                raise ValueError("a NULL tangential stiffness requires normal hertz/material, "
                                 "which is where E and Poisson's ratio come from")
            if self.tangential_id == TANGENTIAL["mindlin"]:
                self.kt = 8.0 * mix_stiffness_g(self.emod, self.emod, self.poiss, self.poiss)
            else:
                raise ValueError("tangential linear_history needs an explicit stiffness")
        else:
            self.kt = float(k_raw)

    def params(self) -> _native.PairParams:
        p = _native.PairParams()
        p.style = _native.PAIR_GRANULAR
        p.limit_damping = self.limit_damping
        p.normal_model = self.normal_id
        p.damping_model = self.damping_id
        p.tangential_model = self.tangential_id
        p.g_kn, p.g_damp, p.g_kt, p.g_xt, p.g_mu = self.kn, self.damp, self.kt, self.xt, self.mu
        return p
