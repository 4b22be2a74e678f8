from __future__ import annotations

from ._harmony_integrate import harmony_integrate

__all__ = ["harmony_integrate"]
