"""pair_style gran/hertz/history (reference: GRANULAR/hertz_history.py).

Same as gran/hooke/history with every stiffness and damping term scaled by
sqrt(Reff * delta); the law runs in the device gather kernel (csrc/contact.h).
"""

from __future__ import annotations

from .. import _native
from .hooke_history import GranHookeHistory


class GranHertzHistory(GranHookeHistory):
    """Constructor order kn, gamman, kt, gammat, xmu (hertz_history.py:36-55)."""

    STYLE = _native.PAIR_HERTZ_HISTORY
