"""Preprocessing namespace (mirrors ``rapids_singlecell.pp`` for the Harmony
path: the PCA that produces its input, the integration itself and the
neighbour graph built on its output)."""

from __future__ import annotations

from .preprocessing import harmony_integrate, neighbors, pca

__all__ = ["harmony_integrate", "neighbors", "pca"]
