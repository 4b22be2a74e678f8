"""Preprocessing namespace (mirrors ``rapids_singlecell.pp``, Harmony only)."""

from __future__ import annotations

from .preprocessing import harmony_integrate

__all__ = ["harmony_integrate"]
