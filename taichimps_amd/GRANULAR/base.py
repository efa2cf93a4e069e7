"""Pair-style plug-in boundary (reference: GRANULAR/base.py:17-38)."""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any

from .. import _native
from ..constants import f64


class GranularPair(ABC):
    """Base class of the contact laws; compute() runs the device gather kernel."""

    def __init__(self, domain, float_type: Any = f64) -> None:
        self.domain = domain
        self.float_type = float_type

    @abstractmethod
    def params(self) -> _native.PairParams:
        """The coefficients as the kernels read them."""

    def _install(self, atom) -> None:
        N = atom._native
        key = bytes(self.params())
        if getattr(N, "_pair_key", None) != key:
            N.pair_set(self.params())
            N._pair_key = key

    def compute(self, atom, nlist, history, dt: float, shearupdate: bool = True) -> None:
        """
        Forces, torques and per-atom virial of every contact, added to the
        atom's current values; contact history advances when shearupdate.
        """
        if atom.nlocal == 0:
            return
        atom._bind_domain(self.domain)
        if history is not None:
            history._bind(atom)
        self._install(atom)
        atom._native.pair_compute(dt, shearupdate, accumulate=True)
