"""Contact-force plug-ins (reference: GRANULAR/)."""

from .base import GranularPair
from .granular import PairGranular
from .hertz_history import GranHertzHistory
from .hooke_history import GranHookeHistory

__all__ = ["GranularPair", "GranHookeHistory", "GranHertzHistory", "PairGranular"]
