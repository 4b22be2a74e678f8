from __future__ import annotations

from ._harmony_integrate import harmony_integrate
from ._neighbors import neighbors
from ._pca import pca

__all__ = ["harmony_integrate", "neighbors", "pca"]
